"""The reference's 14 new-file presets through the native step and raymarch against the oracle.

Each preset (tests/golden/presets, transcribed from src/editor/presets/*.ts) is loaded by the
scene loader and applied as the editor's new-file action does (runtime parameters, quality,
advection mode, sources, lights, render parameters, reset); the FP32 oracle gets the same
settings. 16 steps of 1/15 s (dt_eff = 1/15, ORCH:237) cover the presets' opening bursts and
delayed emitters, with per-step parity: brick flags bit-exact, fields within 1e-5 relative L2.
The presets drive hash noise (noiseMix > 0, turbulent velocity, the eruptive burst falloff,
SH/inject-source-common.wgsl.ts:79-121), whose sine both sides take correctly rounded.
The last state is rendered with the preset's lights and step count (64 x 64, within 2e-3).
"""
import json
import os

import numpy as np
import pytest

import scenes
from fenix_amd import abi
from fenix_amd.render import VolumeRaymarchPass, camera_struct, orbit_camera_snapshot
from fenix_amd.scene import graph_from_preset, scene_from_graph
from fenix_amd.simulation import CombustionVolumeSimulation, make_desc
from oracle import Oracle

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
PRESETS = os.path.join(HERE, "golden", "presets")
PERSISTENT = ("density", "temperature", "fuel", "reaction", "velocity_x", "velocity_y", "velocity_z")
GRID = (64, 128, 64)  # the reference's default 128 x 256 x 128 aspect, halved
DT = 1.0 / 15.0
STEPS = 16


def rel_l2(a, b):
    den = float(np.linalg.norm(b))
    num = float(np.linalg.norm(np.asarray(a, np.float64) - np.asarray(b, np.float64)))
    return num / den if den > 0 else num


def preset_ids():
    with open(os.path.join(PRESETS, "index.json")) as fh:
        return json.load(fh)["presets"]


@pytest.mark.parametrize("pid", preset_ids())
def test_preset_matches_oracle(pid):
    import ctypes as C
    with open(os.path.join(PRESETS, pid + ".json")) as fh:
        preset = json.load(fh)
    sc = scene_from_graph(graph_from_preset(preset))
    desc = make_desc(GRID)
    sim = CombustionVolumeSimulation(desc)
    rp = VolumeRaymarchPass(GRID)
    sc.apply(sim, rp)
    rt, q = abi.RuntimeParams(), abi.Quality()
    assert sim._lib.fx_get_runtime_params(sim._h, C.byref(rt)) == abi.FX_OK
    assert sim._lib.fx_get_quality(sim._h, C.byref(q)) == abi.FX_OK
    orc = Oracle(desc, 32)
    if sc.sources:
        orc.update_sources(*scenes.packed(sc.sources))
    else:
        orc.update_sources(np.zeros(0, np.float32), 0)
    orc.set_runtime(rt, abi.FX_RP_ALL)
    orc.set_quality(q)
    orc.set_mode(abi.FX_ADVECT_MACCORMACK if sc.advection_mode == "maccormack" else abi.FX_ADVECT_SEMI_LAGRANGIAN)
    worst, injected = 0.0, False
    for k in range(STEPS):
        if k > 0:
            for name in PERSISTENT:
                sim.write_field(name, orc.read(name))
        sim.step(None, k * DT, DT)
        orc.step(k * DT, DT)
        np.testing.assert_array_equal(sim.read_flags(abi.FX_FLAGS_INJECTED), orc.flags(abi.FX_FLAGS_INJECTED),
                                      err_msg=f"{pid} step {k}: injection flags")
        np.testing.assert_array_equal(sim.read_flags(), orc.flags(), err_msg=f"{pid} step {k}: render flags")
        for name in PERSISTENT + ("pressure",):
            e = rel_l2(sim.read_field(name), orc.read(name))
            worst = max(worst, e)
            assert e <= 1e-5, f"{pid} step {k}: {name} rel L2 {e:.3e}"
        injected |= max(float(orc.read(f).max()) for f in ("density", "temperature", "fuel", "reaction")) > 1e-3
    print(f"{pid}: {len(sc.sources)} sources, worst rel L2 {worst:.2e}, active bricks {int(orc.flags().sum())}", flush=True)
    if sc.sources:
        assert injected, f"{pid}: nothing was injected"
    for name in PERSISTENT:
        sim.write_field(name, orc.read(name))
    cam = orbit_camera_snapshot()
    img = rp.render(sim, cam, "temperature", 64, 64).astype(np.float32)
    ref = orc.raymarch(camera_struct(cam), rp.params_struct("temperature"), 64, 64).astype(np.float16).astype(np.float32)
    e = rel_l2(img, ref)
    assert e < 2e-3, f"{pid}: image rel L2 {e:.3e}"
    sim.dispose()
