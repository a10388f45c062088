"""The FP64 restatement of the differentiable step (oracle/diff_model.py) and its
hand-derived adjoint, pinned on the CPU:

* every transposed operator (trilinear scatter, central difference, curl, source
  upsampling) passes a dot-product test against its forward operator;
* the position derivative of the trilinear sample matches central differences;
* the pressure projection (the oracle's FP64 V-cycle) is self-adjoint with equal
  pre- and post-sweeps, and measurably not with unequal ones (the planted
  asymmetry shows the test can fail);
* the whole rollout's adjoint matches FP64 central differences of the forward
  along random directions, with each stage switched on in turn.
"""
import numpy as np
import pytest

from oracle import OracleProjector, diff_model

SHAPE = (12, 14, 10)


def _interior(shape):
    m = np.zeros(shape, dtype=bool)
    m[1:-1, 1:-1, 1:-1] = True
    return m


def _dot(a, b):
    return float(sum(np.sum(x * y) for x, y in zip(a, b)))


def test_sampler_scatter_is_transpose_of_eval():
    dm = diff_model()
    rng = np.random.default_rng(1)
    D, H, W = SHAPE
    # positions inside, on the faces and beyond them (clamped)
    px = rng.uniform(-2, W + 1, SHAPE)
    py = rng.uniform(-2, H + 1, SHAPE)
    pz = rng.uniform(-2, D + 1, SHAPE)
    px[0, 0, :3] = [0.0, W - 1.0, 3.0]
    t = dm.Sampler(SHAPE, px, py, pz)
    f = rng.standard_normal(SHAPE)
    a = rng.standard_normal(SHAPE)
    out = np.zeros(SHAPE)
    t.scatter(out, a)
    assert abs(np.sum(t.eval(f) * a) - np.sum(f * out)) <= 1e-12 * np.sum(np.abs(f * out))


def test_sampler_position_gradient_matches_central_differences():
    dm = diff_model()
    rng = np.random.default_rng(2)
    D, H, W = SHAPE
    p = [rng.uniform(0.3, W - 1.3, SHAPE), rng.uniform(0.3, H - 1.3, SHAPE), rng.uniform(0.3, D - 1.3, SHAPE)]
    # keep away from integer coordinates, where the sample has a kink
    p = [np.where(np.abs(c - np.round(c)) < 1e-3, c + 0.01, c) for c in p]
    f = rng.standard_normal(SHAPE)
    g = dm.Sampler(SHAPE, *p).grad(f)
    h = 1e-6
    for axis in range(3):
        hi = [c + (h if k == axis else 0.0) for k, c in enumerate(p)]
        lo = [c - (h if k == axis else 0.0) for k, c in enumerate(p)]
        fd = (dm.Sampler(SHAPE, *hi).eval(f) - dm.Sampler(SHAPE, *lo).eval(f)) / (2 * h)
        np.testing.assert_allclose(g[axis], fd, rtol=0, atol=1e-7)
    # clamped axes do not move the sample
    out = [np.full(SHAPE, -0.5), p[1], p[2]]
    assert np.all(dm.Sampler(SHAPE, *out).grad(f)[0] == 0.0)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_central_difference_and_curl_transposes(axis):
    dm = diff_model()
    rng = np.random.default_rng(3 + axis)
    a, b = rng.standard_normal(SHAPE), rng.standard_normal(SHAPE)
    assert abs(np.sum(dm._cdiff(a, axis) * b) - np.sum(a * dm._cdiff_t(b, axis))) < 1e-10
    u = [rng.standard_normal(SHAPE) for _ in range(3)]
    w = [rng.standard_normal(SHAPE) for _ in range(3)]
    assert abs(_dot(dm.curl(u), w) - _dot(u, dm.curl_t(w))) < 1e-10
    # on fields that vanish off the interior the transpose is the curl itself
    I = _interior(SHAPE)
    wi = [np.where(I, c, 0.0) for c in w]
    for x, y in zip(dm.curl_t(wi), dm.curl(wi)):
        np.testing.assert_allclose(x[I], y[I], atol=1e-12)


def test_curl_of_rigid_rotation():
    dm = diff_model()
    D, H, W = SHAPE
    z, y, x = np.meshgrid(np.arange(D), np.arange(H), np.arange(W), indexing="ij")
    # u = omega x r with omega = (0, 0, 1): curl = (0, 0, 2) in the interior
    w = dm.curl([-(y - 5.0), (x - 4.0), np.zeros(SHAPE)])
    I = _interior(SHAPE)
    np.testing.assert_allclose(w[2][I], 2.0)
    assert np.all(w[0] == 0) and np.all(w[1] == 0) and np.all(w[2][~I] == 0)


def test_source_upsampling_transpose_and_constant_control():
    dm = diff_model()
    rng = np.random.default_rng(5)
    m = dm.DiffModel(SHAPE, steps=1, control_n=5, slab=3, dt=1.0, buoyancy=0.0, confinement=0.0, project=None)
    c = rng.standard_normal((4, 5, 5, 5))
    r = [rng.standard_normal(SHAPE) for _ in range(4)]
    assert abs(_dot(m.source_rates(c), r) - float(np.sum(c * m.source_t(r)))) < 1e-10
    # a constant control is reproduced exactly on the slab and nowhere else
    s = m.source_rates(np.full((4, 5, 5, 5), 0.75))
    D, H, W = SHAPE
    slab = np.zeros(SHAPE, dtype=bool)
    slab[1:-1, 1:4, 1:-1] = True
    np.testing.assert_allclose(s[2][slab], 0.75)
    assert np.all(s[2][~slab] == 0.0)


@pytest.mark.parametrize("post,symmetric", [(2, True), (1, False)])
def test_projection_self_adjoint_only_for_symmetric_vcycle(post, symmetric):
    shape = (16, 16, 16)
    proj = OracleProjector(shape, {"levels": 3, "post_sweeps": post})
    rng = np.random.default_rng(7)
    I = _interior(shape)
    a = [np.where(I, rng.standard_normal(shape), 0.0) for _ in range(3)]
    b = [np.where(I, rng.standard_normal(shape), 0.0) for _ in range(3)]
    lhs, rhs = _dot(proj(a), b), _dot(a, proj(b))
    rel = abs(lhs - rhs) / abs(lhs)
    if symmetric:
        assert rel < 1e-12, rel
    else:
        assert rel > 1e-5, rel
    # the projection removes most of the divergence and leaves the boundary at 0
    pa = proj(a)
    assert all(np.all(c[~I] == 0) for c in pa)


def _model(shape, steps, buoyancy, confinement, project):
    dm = diff_model()
    m = dm.DiffModel(shape, steps=steps, control_n=4, slab=3, dt=1.0, buoyancy=buoyancy, confinement=confinement,
                     project=project)
    D, H, W = shape
    z, y, x = np.meshgrid(np.arange(D), np.arange(H), np.arange(W), indexing="ij")
    m.target = np.exp(-((x - W / 2) ** 2 + (y - 0.6 * H) ** 2 + (z - D / 2) ** 2) / (2 * 3.0 ** 2))
    return m


def _control(rng):
    c = np.zeros((4, 4, 4, 4))
    c[1, :, :, 1:3] = 0.5   # rise
    c[3, :, :, 1:3] = 0.5   # density
    return c + 0.1 * rng.standard_normal(c.shape)


@pytest.mark.parametrize("stages", ["advection", "buoyancy", "confinement", "projection", "all"])
def test_adjoint_matches_fp64_central_differences(stages):
    """Directional derivatives of the 3-step rollout loss: adjoint vs central
    differences of the FP64 forward at h = 1e-7 (small enough that no limiter or
    clamp switches branch between the two evaluations)."""
    shape = (16, 16, 16)
    boundary_only = lambda us: [np.pad(c[1:-1, 1:-1, 1:-1], 1) for c in us]
    beta = 0.05 if stages in ("buoyancy", "all") else 0.0
    eps = 0.3 if stages in ("confinement", "all") else 0.0
    proj = OracleProjector(shape, {"levels": 3}) if stages in ("projection", "all") else boundary_only
    m = _model(shape, 3, beta, eps, proj)
    rng = np.random.default_rng(11)
    c = _control(rng)
    loss, grad = m.gradient(c)
    assert loss > 0
    for k in range(3):
        v = np.random.default_rng(100 + k).standard_normal(c.shape)
        ad = float(np.sum(grad * v))
        h = 1e-7
        fd = (m.loss(c + h * v) - m.loss(c - h * v)) / (2 * h)
        assert abs(fd - ad) <= 1e-6 * abs(ad), (stages, k, fd, ad)
