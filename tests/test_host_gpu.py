"""GPU checks of host-side behaviour: the non-overlapped render mode and the debug-field request."""
import os
import subprocess
import sys

import numpy as np
import pytest

import scenes
from fenix_amd import abi
from fenix_amd.render import VolumeRaymarchPass, orbit_camera_snapshot
from fenix_amd.simulation import CombustionVolumeSimulation, make_desc
from oracle import Oracle

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

_RENDER_SCRIPT = r"""
import sys
sys.path.insert(0, {root!r}); sys.path.insert(0, {here!r})
import numpy as np
import scenes
from fenix_amd.render import VolumeRaymarchPass, orbit_camera_snapshot
from fenix_amd.simulation import CombustionVolumeSimulation
n = 40
desc = scenes.c2_desc(n)
srcs = scenes.c2_sources(n, radius_vox=4.0)
sim = CombustionVolumeSimulation(desc)
sim.updateSources(srcs)
rp = VolumeRaymarchPass((n, n, n))
rp.setRenderParams(stepCount=64)
imgs = []
for k in range(5):
    sim.step(None, k / 60.0, 1 / 60.0)
    imgs.append(rp.render(sim, orbit_camera_snapshot(), "temperature", 72, 56))
np.save(sys.argv[1], np.stack(imgs))
print("rendered", flush=True)
"""


@pytest.mark.parametrize("overlap", ["1", "0"])
def test_render_without_overlap_matches_overlapped(tmp_path, overlap):
    """FENIX_RENDER_OVERLAP=0 launches the raymarch on the simulation stream; fx_raymarch must
    wait for that stream before copying the image back. Each mode runs in a fresh process with
    the variable set before any GPU call; the frames of the two modes are bit-identical."""
    out = {}
    for mode in ("1", overlap):
        path = str(tmp_path / f"img_{mode}.npy")
        env = dict(os.environ, FENIX_RENDER_OVERLAP=mode)
        r = subprocess.run([sys.executable, "-c", _RENDER_SCRIPT.format(root=ROOT, here=HERE), path], env=env,
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        out[mode] = np.load(path)
    np.testing.assert_array_equal(out["1"], out[overlap])
    assert out[overlap][-1][..., 3].astype(np.float32).max() > 0.05


def test_debug_request_writes_velocity_magnitude_above_the_size_rule():
    """ORCH:193, 298-301, 316: above 1.8M voxels the debug pass runs only on the step after
    getDebugBuffers(). 128x112x128 = 1.835M voxels: without a request velocity_magnitude stays 0;
    after a request the next step writes |velocityCurrent|, equal to the oracle's, and the step
    after that does not write it again."""
    W, H, D = 128, 112, 128
    desc = make_desc((W, H, D), runtime={"worldSize": 128.0})
    srcs = scenes.c1_sources()
    sim = CombustionVolumeSimulation(desc)
    orc = Oracle(desc, 32)
    sim.updateSources(srcs)
    orc.update_sources(*scenes.packed(srcs))
    dt = 1.0 / 60.0
    for k in range(2):
        sim.step(None, k * dt, dt)
        orc.step(k * dt, dt)
    assert np.all(sim.read_field("velocity_magnitude") == 0)
    sim.getDebugBuffers()
    orc.request_debug()
    sim.step(None, 2 * dt, dt)
    orc.step(2 * dt, dt)
    umag, ref = sim.read_field("velocity_magnitude"), orc.read("velocity_magnitude")
    assert ref.max() > 0
    den = float(np.linalg.norm(ref))
    assert float(np.linalg.norm(umag - ref)) / den < 1e-5
    # the request is one-shot: the next step leaves the field as it was
    sim.step(None, 3 * dt, dt)
    orc.step(3 * dt, dt)
    np.testing.assert_array_equal(sim.read_field("velocity_magnitude"), umag)
    assert np.array_equal(orc.read("velocity_magnitude"), ref)
