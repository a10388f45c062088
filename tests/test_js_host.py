"""The JS host (fenix_amd/js: fenix.js over the Node-API addon fenix_napi.node) with the reference's
simulation API (src/engine/simulation/runtime/combustion-volume-simulation/types.ts:62-76).

CPU: the module loads, packs emitter sources bit for bit like the Python host (both follow
packEmitterSources.ts:25-121), and surfaces library errors as thrown Errors. GPU: the same scene
stepped through Node and through the Python host gives bit-identical fields, flags and image."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from fenix_amd import abi
from fenix_amd.emitters import BurstEmitter, IgniterEmitter, ScalarEmitter, VelocityEmitter, pack_emitter_sources

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
JS = os.path.join(ROOT, "fenix_amd", "js", "fenix.js")
ADDON = os.path.join(ROOT, "fenix_amd", "js", "fenix_napi.node")

pytestmark = pytest.mark.skipif(shutil.which("node") is None or not os.path.exists(ADDON),
                                reason="node or the built addon is missing")

SOURCES = [
    ScalarEmitter((0.5, 0.2, 0.45), 0.15, startTime=0.0, duration=9999.0, densityRate=1.1, heatRate=1.8, fuelRate=1.4,
                  noiseScale=6.0, noiseMix=0.3, seed=5.0),
    VelocityEmitter((0.5, 0.25, 0.5), 0.2, mode="turbulent", speed=4.0, direction=(0.1, 1.0, 0.0), falloff=0.6,
                    tightness=0.2, seed=9.0),
    VelocityEmitter((0.45, 0.3, 0.5), 0.12, mode="radial", speed=2.0, direction=(0.0, 0.0, 0.0), falloff=1.0,
                    tightness=None, seed=3.0),
    IgniterEmitter((0.5, 0.22, 0.5), 0.1, intensity=0.8, startTime=0.0, duration=5.0, seed=1.0),
    BurstEmitter((0.52, 0.35, 0.48), 0.09, startTime=0.05, duration=0.2, densityAmount=0.3, heatAmount=1.2,
                 fuelAmount=0.4, reactionAmount=0.7, expansionSpeed=30.0, liftSpeed=6.0, turbulenceSpeed=12.0,
                 falloff=0.3, noiseScale=20.0, noiseMix=0.8, seed=11.0),
]


def js_sources():
    out = []
    for s in SOURCES:
        d = {k: v for k, v in vars(s).items()}
        d["kind"] = s.kind
        d["position"] = list(s.position)
        if hasattr(s, "direction"):
            d["direction"] = list(s.direction)
        out.append(d)
    return out


def run_node(script, *args, timeout=120):
    r = subprocess.run(["node", "-e", script, "--", *args], capture_output=True, text=True, timeout=timeout, cwd=ROOT)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_module_exports_the_reference_api_and_packs_like_the_python_host(tmp_path):
    src = tmp_path / "sources.json"
    src.write_text(json.dumps(js_sources()))
    out = run_node(
        "const f = require(process.argv[1]); const fs = require('fs');"
        "const p = f.packEmitterSources(JSON.parse(fs.readFileSync(process.argv[2])));"
        "fs.writeFileSync(process.argv[3], Buffer.from(p.buffer));"
        "console.log(Object.keys(f).sort().join(','))", JS, str(src), str(tmp_path / "packed.bin"))
    assert out.strip().split(",") == sorted(["createCombustionVolumeSimulation", "packEmitterSources", "renderVolume",
                                             "FLOATS_PER_EMITTER", "PRECISION", "native"])
    packed = np.fromfile(tmp_path / "packed.bin", dtype=np.float32)
    np.testing.assert_array_equal(packed, pack_emitter_sources(SOURCES))


def test_library_errors_are_thrown_with_the_library_message():
    """An invalid resolution is rejected by fx_create before any device work; the JS host throws
    Error(message) as the reference's TypeScript throws on creation failures."""
    out = run_node(
        "const f = require(process.argv[1]);"
        "try { f.createCombustionVolumeSimulation(0, {resolution: {width: 2, height: 16, depth: 16}}); console.log('no error') }"
        "catch (e) { console.log(e instanceof Error, e.message) }", JS)
    assert out.startswith("true createCombustionVolumeSimulation: resolution must be at least 4"), out


_GPU_SCRIPT = r"""
const f = require(process.argv[1]);
const fs = require('fs');
const sources = JSON.parse(fs.readFileSync(process.argv[2]));
const out = process.argv[3];
const sim = f.createCombustionVolumeSimulation(0, {
  resolution: {width: 40, height: 48, depth: 32}, worldSize: 40, windStrength: 0.3, wind: [1, 0, 0.2],
  qualitySettings: {finePostIterations: 2, coarseIterations: 12},
});
sim.setRuntimeParams({buoyancy: 4.1, vorticityStrength: 1.6});
sim.setScalarAdvectionMode('maccormack');
sim.updateSources(sources);
for (let k = 0; k < 8; k++) sim.step(null, k / 30, 1 / 30);
const rb = sim.getRenderBuffers();
const dbg = sim.getDebugBuffers();
sim.step(null, 8 / 30, 1 / 30);
const parts = [rb.density.read(), rb.temperature.read(), rb.fuel.read(), rb.reaction.read(), dbg.velocityMagnitude.read(),
               dbg.pressure.read()];
const flags = rb.activeBrickFlags.read();
const cam = {position: [24, 26, -30], target: [0, 14, 0], forward: [-0.55, -0.27, 0.79], right: [-0.82, 0, -0.57],
             up: [-0.15, 0.96, 0.22], aspect: 1, tanHalfFovY: 0.445};
const img = f.renderVolume(sim, cam, {stepCount: 64, scatteringForward: 0.32, scatteringBack: -0.18,
  lights: [{type: 0, direction: [-0.34, 0.88, 0.31], position: [0, 0, 0], color: [1, 0.95, 0.88], intensity: 1.25}]},
  'temperature', 40, 36, 8 / 30);
fs.writeFileSync(out + '.fields', Buffer.concat(parts.map((p) => Buffer.from(p.buffer))));
fs.writeFileSync(out + '.flags', Buffer.from(flags.buffer));
fs.writeFileSync(out + '.img', Buffer.from(img.buffer));
console.log(JSON.stringify({mode: sim.getScalarAdvectionMode(), runtime: sim.getRuntimeParams(),
                            quality: sim.getQualitySettings(), active: sim.activeBrickCount()}));
sim.dispose();
"""


@pytest.mark.gpu
def test_node_host_matches_the_python_host(tmp_path):
    import ctypes as C
    from fenix_amd import abi
    from fenix_amd.render import RenderLight, VolumeRaymarchPass, camera_struct, OrbitCameraSnapshot
    from fenix_amd.simulation import createCombustionVolumeSimulation
    src = tmp_path / "sources.json"
    src.write_text(json.dumps(js_sources()))
    base = str(tmp_path / "node")
    meta = json.loads(run_node(_GPU_SCRIPT, JS, str(src), base).strip().splitlines()[-1])
    sim = createCombustionVolumeSimulation(0, {"resolution": {"width": 40, "height": 48, "depth": 32}, "worldSize": 40,
                                              "windStrength": 0.3, "wind": (1, 0, 0.2),
                                              "qualitySettings": {"finePostIterations": 2, "coarseIterations": 12}})
    sim.setRuntimeParams({"buoyancy": 4.1, "vorticityStrength": 1.6})
    sim.setScalarAdvectionMode("maccormack")
    sim.updateSources(SOURCES)
    for k in range(8):
        sim.step(None, k / 30, 1 / 30)
    sim.getDebugBuffers()
    sim.step(None, 8 / 30, 1 / 30)
    shape = (32, 48, 40)
    fields = np.fromfile(base + ".fields", dtype=np.float32).reshape(6, *shape)
    for i, name in enumerate(("density", "temperature", "fuel", "reaction", "velocity_magnitude", "pressure")):
        np.testing.assert_array_equal(fields[i], sim.read_field(name), err_msg=name)
    assert fields[1].max() > 0.05 and fields[4].max() > 0
    np.testing.assert_array_equal(np.fromfile(base + ".flags", dtype=np.uint32), sim.read_flags())
    cam = OrbitCameraSnapshot((24, 26, -30), (0, 14, 0), (-0.55, -0.27, 0.79), (-0.82, 0, -0.57), (-0.15, 0.96, 0.22), 1.0, 0.445)
    rp = VolumeRaymarchPass((40, 48, 32))
    rp.setRenderParams(stepCount=64, lights=[RenderLight()])
    img = rp.render(sim, cam, "temperature", 40, 36, 8 / 30)
    np.testing.assert_array_equal(np.fromfile(base + ".img", dtype=np.uint16).reshape(36, 40, 4), img.view(np.uint16))
    assert meta["mode"] == "maccormack" and abs(meta["runtime"]["buoyancy"] - 4.1) < 1e-6
    assert meta["quality"]["coarseIterations"] == 12 and meta["active"] == sim.active_brick_count()
    sim.dispose()


def test_storage_format_constants_match_the_c_abi():
    """The JS host's PRECISION table carries the FX_PRECISION_* values of fenix.h."""
    out = run_node("console.log(JSON.stringify(require(process.argv[1]).PRECISION))", JS)
    table = json.loads(out.strip().splitlines()[-1])
    assert table == {name: getattr(abi, f"FX_PRECISION_{name}") for name in
                     ("FP32", "BF16", "FP16", "BF16_RNE", "FP16_RNE", "U16", "U16_RNE", "UF16", "UF16_RNE")}
