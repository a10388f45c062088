"""Host-side logic that needs no GPU: emitter packing, camera/bounds math,
render-parameter sanitising, synthetic-input generators."""
import math

import numpy as np
import pytest

import scenes
from fenix_amd.emitters import (BurstEmitter, IgniterEmitter, ScalarEmitter, VelocityEmitter, pack_emitter_sources)
from fenix_amd.render import (RenderLight, VolumeRaymarchPass, default_step_count, get_volume_world_bounds,
                              orbit_camera_snapshot, sanitize_render_lights)


def test_pack_layout_matches_pack_emitter_sources():
    # packEmitterSources.ts:69-121 slot by slot
    srcs = [
        ScalarEmitter((0.1, 0.2, 0.3), 0.05, startTime=1.0, duration=2.0, densityRate=3.0, heatRate=4.0, fuelRate=5.0,
                      noiseScale=6.0, noiseMix=0.7, seed=8.0),
        VelocityEmitter((0.4, 0.5, 0.6), 0.07, mode="turbulent", speed=9.0, direction=(1.0, 2.0, 3.0), falloff=0.25,
                        tightness=0.5, seed=11.0),
        BurstEmitter((0.7, 0.8, 0.9), 0.02, densityAmount=1.5, heatAmount=2.5, fuelAmount=3.5, reactionAmount=4.5,
                     expansionSpeed=5.5, liftSpeed=6.5, turbulenceSpeed=7.5, falloff=0.1, noiseScale=2.0, noiseMix=0.3,
                     seed=12.0),
        IgniterEmitter((0.0, 0.1, 0.2), 0.03, intensity=0.9, seed=13.0),
    ]
    p = pack_emitter_sources(srcs).reshape(4, 32)
    assert p.dtype == np.float32 and p.shape == (4, 32)
    np.testing.assert_allclose(p[0, [0, 1, 2, 3, 4, 5, 8, 9, 10, 20, 21, 28, 29]],
                               [0.1, 0.2, 0.3, 0.05, 1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 0.7, 0, 8.0], rtol=1e-7)
    np.testing.assert_allclose(p[1, [12, 13, 14, 15, 16, 17, 18, 28, 29]], [9.0, 2, 0.25, 0.5, 1, 2, 3, 1, 11.0])
    np.testing.assert_allclose(p[2, [8, 9, 10, 12, 13, 14, 15, 20, 21, 22, 28]],
                               [1.5, 2.5, 3.5, 5.5, 6.5, 0.1, 7.5, 2.0, 0.3, 4.5, 3], rtol=1e-7)
    np.testing.assert_allclose(p[3, [22, 28, 29]], [0.9, 2, 13.0], rtol=1e-7)
    # untouched slots stay zero
    assert np.all(p[0, [11, 12, 13, 14, 15, 16, 17, 18, 19, 22, 23, 24, 25, 26, 27, 30, 31]] == 0)
    assert pack_emitter_sources([]).size == 0


def test_orbit_camera_snapshot_basis():
    cam = orbit_camera_snapshot()
    f, r, u = (np.array(v) for v in (cam.forward, cam.right, cam.up))
    for v in (f, r, u):
        assert abs(np.linalg.norm(v) - 1) < 1e-12
    assert abs(f @ r) < 1e-12 and abs(f @ u) < 1e-12 and abs(r @ u) < 1e-12
    assert abs(np.linalg.norm(np.array(cam.position) - np.array(cam.target)) - 40.0) < 1e-9
    # forward points from the camera to the target
    np.testing.assert_allclose(np.array(cam.position) + 40.0 * f, cam.target, atol=1e-9)
    assert r[1] == 0.0  # right = forward x world-up is horizontal
    assert cam.tanHalfFovY == pytest.approx(math.tan(math.radians(48)))


def test_volume_bounds_and_step_count():
    c, h = get_volume_world_bounds((64, 64, 64))
    assert h == pytest.approx((27.675, 27.675, 27.675))
    assert c == pytest.approx((0.0, 27.425, 0.0))
    c, h = get_volume_world_bounds((128, 256, 128))
    assert h[1] == pytest.approx(2 * h[0])
    assert default_step_count((64, 64, 64)) == 180
    assert default_step_count((256, 256, 256)) == 384
    assert default_step_count((512, 512, 512)) == 720


def test_render_params_clamps():
    rp = VolumeRaymarchPass((32, 32, 32))
    rp.setRenderParams(stepCount=3, scatteringForward=2.0, scatteringBack=-5.0)
    assert rp.stepCount == 32 and rp.scatteringForward == 0.98 and rp.scatteringBack == -0.98
    lights = sanitize_render_lights([RenderLight(direction=(0, 0, 0), color=(-1, 2, 3), intensity=-4)] * 5)
    assert len(lights) == 3
    assert lights[0].direction == (-0.34, 0.88, 0.31)
    assert lights[0].color == (0.0, 2, 3) and lights[0].intensity == 0.0
    p = rp.params_struct("fuel")
    assert p.display_mode == 2 and p.step_count == 32


def test_splitmix64_known_values():
    # Published SplitMix64 outputs for seed 0 (Vigna's reference implementation).
    g = scenes.SplitMix64(0)
    assert g.next_u64() == 0xE220A8397B1DCDAF
    assert g.next_u64() == 0x6E789E6AA1B965F4
    assert g.next_u64() == 0x06C45D188009454F


def test_c2_sources_on_floor_inside_domain():
    srcs = scenes.c2_sources(256)
    assert len(srcs) == 2
    for s in srcs:
        r = s.radius
        assert r == pytest.approx(10 / 256)
        assert r <= s.position[0] <= 1 - r and r <= s.position[2] <= 1 - r
        assert s.position[1] == pytest.approx(10 / 256)
