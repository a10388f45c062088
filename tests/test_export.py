"""Offline export (fenix_amd.export): the flipbook layout on the CPU; on the GPU,
frames are deterministic, equal the hand-driven step + raymarch + composite, and
start from simulation time 0 whatever the first frame index (the step clock is
relative to the first step after a reset, ORCH:230-236)."""
import os

import numpy as np
import pytest

from fenix_amd.export import export_frames, flipbook

HERE = os.path.dirname(os.path.abspath(__file__))


def test_flipbook_layout_and_background():
    frames = [np.full((2, 3, 4), i * 40, dtype=np.uint8) for i in range(5)]
    sheet = flipbook(frames, columns=2)
    assert sheet.shape == (6, 6, 4)
    for i in range(5):
        r, c = divmod(i, 2)
        assert np.all(sheet[r * 2:(r + 1) * 2, c * 3:(c + 1) * 3] == i * 40)
    assert np.all(sheet[4:6, 3:6] == 0)  # the unused cell stays transparent
    f = np.zeros((1, 1, 4), dtype=np.uint8)
    f[0, 0] = (255, 0, 0, 128)
    out = flipbook([f], columns=1, background=(0, 0, 255, 255))
    # source-over on straight alpha: 128/255 of red over blue
    a = 128 / 255
    assert tuple(out[0, 0]) == (round(255 * a), 0, round(255 * (1 - a)), 255)
    with pytest.raises(ValueError):
        flipbook([], 1)


@pytest.mark.gpu
def test_export_frames_are_deterministic_and_match_manual_frames():
    from fenix_amd.render import BloomPass, VolumeRaymarchPass, orbit_camera_snapshot
    from fenix_amd.scene import load_scene
    from fenix_amd.simulation import CombustionVolumeSimulation, make_desc
    with open(os.path.join(HERE, "golden", "scenes", "twin_fires_v3.json")) as fh:
        sc = load_scene(fh.read())
    n = 32
    sim = CombustionVolumeSimulation(make_desc((n, n, n)))
    rp = VolumeRaymarchPass((n, n, n))
    sc.apply(sim, rp)
    cam = orbit_camera_snapshot()
    bloom = BloomPass()
    enabled, thr, strength, radius = sc.bloom()
    bp = bloom.params_struct(enabled, thr, strength, radius)
    a = list(export_frames(sim, rp, cam, 0, 5, 30.0, 48, 40, bloom=bloom, bloom_params=bp))
    b = list(export_frames(sim, rp, cam, 0, 5, 30.0, 48, 40, bloom=bloom, bloom_params=bp))
    assert len(a) == 6 and a[0].shape == (40, 48, 4) and a[0].dtype == np.uint8
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    assert a[-1][..., 3].max() > 0, "nothing rendered"
    # the same frames driven by hand
    sim.reset()
    for i in range(6):
        t = i / 30.0
        sim.step(None, t, 1 / 30.0)
        hdr = rp.render(sim, cam, "temperature", 48, 40, t)
        np.testing.assert_array_equal(bloom.render(sim, hdr, bp), a[i])
    # starting at frame 3: the first exported frame is simulation time 0 again
    c = list(export_frames(sim, rp, cam, 3, 4, 30.0, 48, 40, bloom=bloom, bloom_params=bp))
    np.testing.assert_array_equal(c[0], a[0])


@pytest.mark.gpu
def test_export_frames_match_the_oracle_chain():
    """renderOffscreenFrame (createViewportRuntime.ts:205-232) per exported frame: step at
    (start + i) / rate, raymarch into the HDR image, bloom/composite into the canvas. The GPU
    frames follow the same chain run on the CPU (FP32 oracle step and raymarch, rounded to the
    rgba16float HDR image, then the bloom restatement), free-running from the reset: canvases
    within 2 of 255 everywhere and within 1 on >= 99.9% of the channel values."""
    import scenes
    from oracle import Oracle, bloom_model
    from fenix_amd import abi
    from fenix_amd.render import BloomPass, VolumeRaymarchPass, camera_struct, orbit_camera_snapshot
    from fenix_amd.scene import load_scene
    from fenix_amd.simulation import CombustionVolumeSimulation, make_desc
    import ctypes as C
    with open(os.path.join(HERE, "golden", "scenes", "twin_fires_v3.json")) as fh:
        sc = load_scene(fh.read())
    n, w, h, rate = 32, 48, 40, 30.0
    desc = make_desc((n, n, n))
    sim = CombustionVolumeSimulation(desc)
    rp = VolumeRaymarchPass((n, n, n))
    sc.apply(sim, rp)
    orc = Oracle(desc, 32)
    orc.update_sources(*scenes.packed(sc.sources))
    rt, q = abi.RuntimeParams(), abi.Quality()
    assert sim._lib.fx_get_runtime_params(sim._h, C.byref(rt)) == abi.FX_OK
    assert sim._lib.fx_get_quality(sim._h, C.byref(q)) == abi.FX_OK
    orc.set_runtime(rt, abi.FX_RP_ALL)
    orc.set_quality(q)
    cam = orbit_camera_snapshot()
    bloom, bm = BloomPass(), bloom_model()
    enabled, thr, strength, radius = sc.bloom()
    frames = list(export_frames(sim, rp, cam, 2, 9, rate, w, h, bloom=bloom,
                                bloom_params=bloom.params_struct(enabled, thr, strength, radius)))
    orc.reset()
    lit = 0
    for i, frame in enumerate(frames):
        t = (2 + i) / rate
        orc.step(t, 1.0 / rate)
        hdr = orc.raymarch(camera_struct(cam), rp.params_struct("temperature", t), w, h).astype(np.float16)
        ref = bm.bloom(hdr, enabled=enabled, threshold_=thr, strength=strength, radius=radius)
        diff = np.abs(frame.astype(np.int32) - ref.astype(np.int32))
        assert diff.max() <= 2, (i, diff.max())
        assert np.mean(diff <= 1) >= 0.999, (i, np.mean(diff <= 1))
        lit += int(frame[..., 3].max() > 0)
    assert lit >= 4, "the exported frames show nothing"
