"""The bloom/composite restatement (oracle/bloom.py) pinned by closed forms, and
the HIP passes against it on the GPU (images within 1 LSB of the 8-bit canvas)."""
import numpy as np
import pytest

from oracle import bloom_model


def test_sampler_reads_texel_centres_exactly_and_averages_2x2_on_downsample():
    bm = bloom_model()
    rng = np.random.default_rng(0)
    img = rng.standard_normal((6, 8, 4))
    v, u = np.meshgrid((np.arange(6) + 0.5) / 6, (np.arange(8) + 0.5) / 8, indexing="ij")
    np.testing.assert_array_equal(bm.sample_linear(img, u, v), img)
    v, u = np.meshgrid((np.arange(3) + 0.5) / 3, (np.arange(4) + 0.5) / 4, indexing="ij")
    avg = 0.25 * (img[0::2, 0::2] + img[1::2, 0::2] + img[0::2, 1::2] + img[1::2, 1::2])
    np.testing.assert_allclose(bm.sample_linear(img, u, v), avg, atol=1e-14)
    # clamp to edge: outside the image the border texel repeats
    np.testing.assert_array_equal(bm.sample_linear(img, np.array([-3.0]), np.array([0.5 / 6])), img[0:1, 0])


def test_threshold_knee_limits():
    bm = bloom_model()
    thr = 0.6
    img = np.zeros((2, 2, 4))
    for lum, expect in ((1.0, 1.0), (0.6, 1.0), (0.2, 0.0), (0.3, 0.0)):
        img[...] = [lum, lum, lum, 0.5]
        out = bm.threshold(img, 1, 1, thr)[0, 0]
        assert abs(out[3] - 0.5 * expect) < 1e-3 and abs(out[0] - lum * 0.5 * expect) < 1e-3, (lum, out)
    # inside the knee (lum in (thr - knee, thr)): w = rq^2 / (4 knee^2)
    lum = 0.5
    img[...] = [lum, lum, lum, 1.0]
    knee = thr / 2
    rq = lum - thr + knee
    assert abs(bm.threshold(img, 1, 1, thr)[0, 0, 3] - rq * rq / (4 * knee * knee + 1e-5)) < 1e-3


def test_blur_preserves_constants_and_has_the_gaussian_impulse_response():
    bm = bloom_model()
    assert abs(sum(bm.WEIGHTS) + sum(bm.WEIGHTS[1:]) - 1.0) < 1e-8
    const = np.full((9, 12, 4), 0.75)
    np.testing.assert_allclose(bm.blur(const, 0.7, True), 0.75, atol=5e-4)
    imp = np.zeros((1, 21, 4))
    imp[0, 10] = 1.0
    row = bm.blur(imp, 0.0, True)[0, :, 0]
    expect = np.zeros(21)
    for k, w in enumerate(bm.WEIGHTS):
        expect[10 + k] = expect[10 - k] = w
    np.testing.assert_allclose(row, expect, atol=5e-4)
    # separable: horizontal then vertical of an impulse is the outer product
    imp2 = np.zeros((21, 21, 4))
    imp2[10, 10] = 1.0
    both = bm.blur(bm.blur(imp2, 0.0, True), 0.0, False)[..., 0]
    np.testing.assert_allclose(both, np.outer(expect, expect), atol=1e-3)


def test_disabled_bloom_is_the_gamma_blit():
    bm = bloom_model()
    img = np.zeros((2, 3, 4), dtype=np.float16)
    img[..., 0] = [[0.5, 2.0, -1.0], [0.25, 0.0, 1.0]]
    img[..., 3] = [[0.5, 1.5, -0.2], [1.0, 0.0, 0.3]]
    out = bm.bloom(img, enabled=False)
    expect_r = np.rint(np.clip(np.power(np.maximum(img[..., 0].astype(np.float64), 0), 0.92), 0, 1) * 255)
    np.testing.assert_array_equal(out[..., 0], expect_r)
    np.testing.assert_array_equal(out[..., 3], np.rint(np.clip(img[..., 3].astype(np.float64), 0, 1) * 255))


def _hdr_image(h, w, seed):
    rng = np.random.default_rng(seed)
    img = np.zeros((h, w, 4), dtype=np.float32)
    yy, xx = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    for _ in range(6):
        cx, cy, r = rng.uniform(0, w), rng.uniform(0, h), rng.uniform(3, 12)
        g = np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / (2 * r * r))
        img[..., :3] += g[..., None] * rng.uniform(0.2, 2.5, 3)
        img[..., 3] += g * rng.uniform(0.3, 1.0)
    img[..., 3] = np.clip(img[..., 3], 0, 1)
    return img.astype(np.float16)


@pytest.mark.gpu
@pytest.mark.parametrize("w,h,enabled,radius", [(96, 64, True, 0.4), (97, 51, True, 1.3), (64, 64, False, 0.4),
                                                (1, 1, True, 0.4), (3, 2, True, 0.0)])
def test_hip_bloom_matches_restatement(w, h, enabled, radius):
    from fenix_amd.render import BloomPass
    from fenix_amd.simulation import CombustionVolumeSimulation, make_desc
    bm = bloom_model()
    sim = CombustionVolumeSimulation(make_desc((16, 16, 16)))
    bp = BloomPass()
    src = _hdr_image(h, w, seed=w * 7 + h)
    out = bp.render(sim, src, bp.params_struct(enabled=enabled, threshold=0.6, strength=0.9, radius=radius))
    ref = bm.bloom(src, enabled=enabled, threshold_=0.6, strength=0.9, radius=radius)
    diff = np.abs(out.astype(np.int32) - ref.astype(np.int32))
    assert diff.max() <= 2, diff.max()
    assert np.mean(diff <= 1) >= 0.999, np.mean(diff <= 1)
    if enabled and w > 8:
        # the glow spreads light beyond the sources' alpha: composite differs from the blit somewhere
        blit = bm.bloom(src, enabled=False)
        assert np.any(out != blit)
