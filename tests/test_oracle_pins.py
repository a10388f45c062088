"""Known-answer pins of the oracle functions that uniform or integer-shift cases leave open (no GPU).

Each test states the closed form it checks and names the planted mistake it catches. Paths are
relative to the reference repository root:
SH = src/engine/simulation/shaders/passes, SC = src/engine/simulation/shaders/common,
RP = src/engine/render/passes/createVolumeRaymarchPass.ts.
"""
import math

import numpy as np
import pytest

import scenes
from fenix_amd import abi
from fenix_amd.simulation import make_desc
from oracle import (PROBE_FIRE_PALETTE, PROBE_LIGHT_ATT, PROBE_LIGHT_DIR, PROBE_LIGHT_POWER, PROBE_PHASE_HG,
                    PROBE_SMOKE_ALBEDO, Oracle, probe)

LA, LB, LDIV = abi.FX_LEVEL_PRESSURE_A, abi.FX_LEVEL_PRESSURE_B, abi.FX_LEVEL_DIVERGENCE
ZERO_Q = {"finePreIterations": 0, "finePostIterations": 0, "midPreIterations": 0, "midPostIterations": 0,
          "coarseIterations": 0}


def grid(shape):
    d, h, w = shape
    return np.meshgrid(np.arange(d), np.arange(h), np.arange(w), indexing="ij")  # z, y, x


def affine(coef, x, y, z):
    a, bx, by, bz = coef
    return a + bx * x + by * y + bz * z


# ----------------------------------------------------------- prolongation --
@pytest.mark.parametrize("solver,shape,level", [
    (abi.FX_SOLVER_REFERENCE, (32, 24, 16), 0),
    (abi.FX_SOLVER_REFERENCE, (32, 24, 16), 1),
    (abi.FX_SOLVER_VCYCLE, (41, 32, 24), 0),   # odd fine width: the last interior column clamps
    (abi.FX_SOLVER_VCYCLE, (48, 32, 40), 2),
])
def test_prolongation_adds_trilinear_of_affine_coarse_field(solver, shape, level):
    """SH/pressure-prolongation.wgsl.ts:36-43: fine += sampleCoarse((id + 0.5) * 0.5 - 0.5) and 0
    on the fine boundary faces; sampleCoarse clamps to [0, dim - 1] (SC/sampling-scalar.wgsl.ts:17).
    Trilinear interpolation reproduces an affine coarse field exactly, so every interior fine voxel
    gains affine(clamp(pos)), with distinct slopes per axis. Planted mistakes this catches: sampling
    at id * 0.5 (half-cell offset: off by slope / 4), assigning instead of adding (off by the prior
    fine value), or swapping the x and y fractions (off where the slopes differ)."""
    W, H, D = shape
    o = Oracle(make_desc((W, H, D), solver=solver, quality=ZERO_Q, vcycle={"levels": 4}), 64)
    cw, ch, cd = o.level_dims(level + 1)
    fw, fh, fd = o.level_dims(level)
    coef = (0.3, 0.11, -0.07, 0.05)
    cz, cy, cx = grid((cd, ch, cw))
    o.write_level(level + 1, LA, affine(coef, cx, cy, cz))
    rng = np.random.default_rng(11)
    prior = rng.uniform(-1, 1, (fd, fh, fw))
    o.write_level(level, LB, prior)
    o.level_op(Oracle.OP_PROLONG, level, src=LA, dst=LB)
    out = o.level(level, LB)
    fz, fy, fx = grid((fd, fh, fw))
    pos = [np.clip((c + 0.5) * 0.5 - 0.5, 0, m - 1) for c, m in ((fx, cw), (fy, ch), (fz, cd))]
    expect = prior + affine(coef, *pos)
    inner = np.zeros((fd, fh, fw), bool)
    inner[1:-1, 1:-1, 1:-1] = True
    np.testing.assert_allclose(out[inner], expect[inner], atol=1e-13)
    assert np.all(out[~inner] == 0)


def test_full_weight_restriction_is_half_the_transposed_prolongation():
    """The V-cycle restriction (4 x the (1,3,3,1)/8 stencil per axis) is P^T / 2, P the trilinear
    prolongation above: per axis the prolongation weights of coarse I over fine 2I-1 .. 2I+2 are
    (1,3,3,1)/4, the restriction's (1,3,3,1)/8, so <R f, c> = 4/8 <f, P c> for f and c vanishing
    near the faces. Planted mistakes this catches: (1,2,2,1)/6 weights, the stencil anchored at 2I
    instead of 2I-1, a missing h^2 factor 4 (the identity then holds with 1/8)."""
    W, H, D = 40, 32, 48
    o = Oracle(make_desc((W, H, D), solver=abi.FX_SOLVER_VCYCLE, vcycle={"levels": 3}), 64)
    cw, ch, cd = o.level_dims(1)
    rng = np.random.default_rng(5)
    f = np.zeros((D, H, W))
    f[3:-3, 3:-3, 3:-3] = rng.uniform(-1, 1, (D - 6, H - 6, W - 6))
    c = np.zeros((cd, ch, cw))
    c[2:-2, 2:-2, 2:-2] = rng.uniform(-1, 1, (cd - 4, ch - 4, cw - 4))
    o.write_level(0, LB, f)
    o.level_op(Oracle.OP_RESTRICT_FULL, 0, src=LB)
    Rf = o.level(1, LDIV)
    o.write_level(1, LA, c)
    o.write_level(0, LA, np.zeros((D, H, W)))
    o.level_op(Oracle.OP_PROLONG, 0, src=LA, dst=LA)
    Pc = o.level(0, LA)
    lhs, rhs = float(np.sum(Rf * c)), 0.5 * float(np.sum(f * Pc))
    assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(rhs)), (lhs, rhs)
    # the mean restriction (reference schedule) is the plain 8-child average
    o.level_op(Oracle.OP_RESTRICT_MEAN, 0, src=LB)
    mean = o.level(1, LDIV)
    blocks = f[:2 * cd, :2 * ch, :2 * cw].reshape(cd, 2, ch, 2, cw, 2).mean(axis=(1, 3, 5))
    inner = np.zeros_like(mean, bool)
    inner[1:-1, 1:-1, 1:-1] = True
    np.testing.assert_allclose(mean[inner], blocks[inner], atol=1e-15)


# ------------------------------------------------------------ the sampler --
def test_trilinear_sampler_is_exact_on_affine_velocity_at_fractional_positions():
    """SC/sampling-velocity.wgsl.ts:3-30 via SH/advect-velocity-semilagrangian.wgsl.ts:23-41.
    With an affine velocity u_c(p) = a_c + B_c p (voxel units) and dt = 1, the backtrace lands at
    p = i - u(i), fractional in all three axes, and trilinear interpolation of an affine field is
    exact there: u'(i) = (a + B (i - u(i))) (1 - 0.025). Planted mistakes this catches: a half-cell
    offset in the sample position, x and y fractions swapped, the upper corner taken as base + 2."""
    W, H, D = 20, 18, 16
    o = Oracle(make_desc((W, H, D)), 64)
    z, y, x = grid((D, H, W))
    a = np.array([0.37, -0.21, 0.44])
    B = np.array([[0.013, -0.021, 0.017], [0.019, 0.011, -0.015], [-0.012, 0.023, 0.009]])
    pos = np.stack([x, y, z]).astype(np.float64)
    u = [a[c] + B[c, 0] * x + B[c, 1] * y + B[c, 2] * z for c in range(3)]
    for name, comp in zip(("velocity_x", "velocity_y", "velocity_z"), u):
        o.write(name, comp)
    o.run_pass("advect_velocity", scenes.sim_params(dt=1.0))
    back = pos - np.stack(u)
    inside = np.all([(back[k] >= 0) & (back[k] <= m - 1) for k, m in enumerate((W, H, D))], axis=0)
    inner = inside & (x > 0) & (y > 0) & (z > 0) & (x < W - 1) & (y < H - 1) & (z < D - 1)
    assert inner.sum() > 2000
    fr = back[:, inner] - np.floor(back[:, inner])
    assert np.all(fr.std(axis=1) > 0.1)  # the fractions vary in every axis
    for c, name in enumerate(("scratch_x", "scratch_y", "scratch_z")):
        expect = (a[c] + B[c, 0] * back[0] + B[c, 1] * back[1] + B[c, 2] * back[2]) * (1 - 0.025)
        np.testing.assert_allclose(o.read(name)[inner], expect[inner], atol=1e-13)


# ------------------------------------------------------ MacCormack limiter --
def test_maccormack_limiter_clamps_to_original_and_predicted():
    """SH/advect-scalar-maccormack.wgsl.ts:35-45 on the fuel channel, dt = 1, velocity along x only.
    Three rows (each through voxel x = 10), with u(9) = 2.5, u(10) = 0.5, so the backtrace of voxel
    10 samples x = 9.5 (predicted = mean(phi9, phi10)), the forward velocity there is 1.5 and the
    recovered sample is phi11:
      row A: phi9..11 = 0.5, 0.6, 0.9 -> corrected 0.40, clamped to [0.55, 0.6]: 0.55
      row B: phi9..11 = 0.6, 0.5, 0.2 -> corrected 0.70, clamped to [0.5, 0.55]: 0.55
      row C: phi9..11 = 0.3, 0.7, 0.5 -> corrected 0.60, inside [0.5, 0.7]:      0.60
    The next row up and the next layer hold values (0.2 in A, 0.8 in B) that widen the 8-corner
    neighbourhood of the backtrace, so clamping to the neighbourhood instead (a common MacCormack
    limiter) would leave 0.40 and 0.70; dropping the correction would give 0.5 in row C. The
    noise-modulated fuel decay depends only on the voxel position (thermalNoise(coord)), so a run
    with zero velocity gives it exactly: limited = out - out(u = 0) + phi10."""
    W, H, D = 16, 12, 12
    rows = {  # (y, z): (phi9, phi10, phi11, neighbourhood value)
        (3, 3): (0.5, 0.6, 0.9, 0.2),
        (6, 3): (0.6, 0.5, 0.2, 0.8),
        (9, 7): (0.3, 0.7, 0.5, 0.45),
    }
    expect = {(3, 3): 0.55, (6, 3): 0.55, (9, 7): 0.60}
    fuel = np.full((D, H, W), 0.4)
    ux = np.zeros((D, H, W))
    for (y0, z0), (p9, p10, p11, nb) in rows.items():
        fuel[z0, y0, 9:12] = (p9, p10, p11)
        fuel[z0, y0 + 1, 9:11] = nb
        fuel[z0 + 1, y0, 9:11] = nb
        fuel[z0 + 1, y0 + 1, 9:11] = nb
        ux[z0, y0, 9], ux[z0, y0, 10] = 2.5, 0.5
    outs = []
    for vel in (ux, np.zeros_like(ux)):
        o = Oracle(make_desc((W, H, D)), 64)
        o.write("temperature", np.full((D, H, W), 0.5))  # every brick active
        o.write("fuel", fuel)
        o.write("scratch_x", vel)
        p = scenes.sim_params(dt=1.0)
        o.run_pass("activate_render", p)
        assert o.flags(abi.FX_FLAGS_INJECTED).all()
        o.run_pass("advect_scalars", p)
        outs.append(o.read("adv_fuel"))
    for (y0, z0), val in expect.items():
        limited = outs[0][z0, y0, 10] - outs[1][z0, y0, 10] + fuel[z0, y0, 10]
        assert abs(limited - val) < 1e-12, ((y0, z0), limited, val)


# ------------------------------------------------------------- combustion --
def _combustion_run(T, F=None, R=None, d0=0.5, dt=0.05):
    n = T.shape[0]
    o = Oracle(make_desc((n, n, n)), 64)
    zero = np.zeros_like(T)
    o.write("adv_density", np.full_like(T, d0))
    o.write("adv_temperature", T)
    o.write("adv_fuel", zero if F is None else F)
    o.write("adv_reaction", zero if R is None else R)
    p = scenes.sim_params(dt=dt)
    o.run_pass("activate_advected", p)
    assert o.flags().all()
    o.run_pass("combustion", p)
    return o.read("reaction"), o.read("temperature")


def test_combustion_radiative_gain_uses_the_six_face_neighbours():
    """SH/combustion-update.wgsl.ts:36-40, 56-68 (read-old). With no fuel and no reaction the
    reaction output is exactly 1.8 x radiative, radiative = max(nh - 0.55 local, 0) s(d) k dt 2.6,
    nh the mean hot-gas of the six face neighbours, hotgas = clamp(1.55 T + 0.26 F + 0.44 R) and k
    a noise factor of the voxel's own position. One hot voxel h in a uniform T_b = 0.2 field:
      - footprint: only h and its six face neighbours change (not the 12 edge or 8 corner ones);
      - at a neighbour, the ratio of radiative for T_h = 0.5 and T_h = 0.4 is exact (k cancels):
        (1.7 L + H_a) / (1.7 L + H_b), L = hotgas(T_b);
      - the same hot-gas through reaction (R_h) or fuel (F_h) instead of temperature gives the same
        neighbour gain (weights 1.55 : 0.26 : 0.44).
    Planted mistakes this catches: a neighbour at x + 2, the 26-neighbourhood, 0.5 for 0.55, a sum
    instead of a mean, swapped hot-gas weights."""
    n = 16
    h = (8, 8, 8)
    Tb = np.full((n, n, n), 0.2)
    base_r, _ = _combustion_run(Tb)
    Ta = Tb.copy(); Ta[h] = 0.5
    Tc = Tb.copy(); Tc[h] = 0.4
    ra, _ = _combustion_run(Ta)
    rc, _ = _combustion_run(Tc)
    changed = np.argwhere(ra != base_r)
    faces = {(8, 8, 8), (7, 8, 8), (9, 8, 8), (8, 7, 8), (8, 9, 8), (8, 8, 7), (8, 8, 9)}
    assert {tuple(c) for c in changed} <= faces
    nbrs = faces - {h}
    assert all(ra[c] > base_r[c] for c in nbrs)
    L, Ha, Hc = 1.55 * 0.2, 1.55 * 0.5, 1.55 * 0.4
    want = (1.7 * L + Ha) / (1.7 * L + Hc)
    for c in nbrs:
        assert abs(ra[c] / rc[c] - want) < 1e-7, (c, ra[c] / rc[c], want)
    # the same hot-gas value through reaction or fuel at h
    Rh = np.zeros_like(Tb); Rh[h] = (Ha - L) / 0.44
    Fh = np.zeros_like(Tb); Fh[h] = (Ha - L) / 0.26
    rr, _ = _combustion_run(Tb, R=Rh)
    rf, _ = _combustion_run(Tb, F=Fh)
    for c in nbrs:
        assert abs(rr[c] - ra[c]) < 1e-12 and abs(rf[c] - ra[c]) < 1e-12, c
    # magnitude band: k in [0.45, 1], s(0.5) = smoothstep(0.04, 0.76, 0.5)
    t = (0.5 - 0.04) / 0.72
    s = t * t * (3 - 2 * t)
    rad = (5 * L + Ha) / 6 - 0.55 * L
    for c in nbrs:
        g = ra[c] / 1.8 / (rad * s * 0.05 * 2.6)
        assert 0.45 - 1e-9 <= g <= 1 + 1e-9, (c, g)


# --------------------------------------------------------------- buoyancy --
def _buoyancy_run(T, den, R=None, dt=0.05, seed=3):
    n = T.shape[0]
    o = Oracle(make_desc((n, n, n)), 64)
    o.write("density", den)
    o.write("temperature", T)
    o.write("fuel", np.zeros_like(T))
    o.write("reaction", np.zeros_like(T) if R is None else R)
    o.write("adv_density", np.full_like(T, 0.5))
    p = scenes.sim_params(dt=dt)
    o.run_pass("activate_advected", p)
    rng = np.random.default_rng(seed)
    for c in "xyz":
        o.write("scratch_" + c, rng.uniform(-0.5, 0.5, T.shape))
    o.run_pass("buoyancy", p)
    return [o.read("scratch_" + c) for c in "xyz"]


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_buoyancy_temperature_gradient_collapse_term(axis):
    """SH/apply-buoyancy.wgsl.ts:55-57, 73-80, 85: tg = central difference of hotGas / 2 (clamped
    reads), collapse = d (1 - smoothstep(0.16, 0.64, hotGas)) and v += tg collapse dt (3.2 + 0.08 |b|).
    On the plane where T = 0.1 (hotGas 0.1, collapse = d) with T varying linearly along one axis,
    only tg depends on the slope s, so two slopes differ by exactly
    (s1 - s2) d dt (3.2 + 0.08 * 3.6) in that component and by 0 in the other two.
    Planted mistakes this catches: a one-sided difference, a missing 1/2, the gradient of density
    instead of hot-gas, swapped components."""
    n, c0, d0, dt = 16, 8, 0.3, 0.05
    zyx = grid((n, n, n))
    coord = zyx[2 - axis]
    den = np.full((n, n, n), d0)
    runs = [_buoyancy_run(0.1 + s * (coord - c0), den, dt=dt) for s in (0.02, -0.01)]
    plane = (coord == c0)
    m = plane & (zyx[2] > 0) & (zyx[2] < n - 1) & (zyx[0] > 0) & (zyx[0] < n - 1) & (zyx[1] > 0) & (zyx[1] < n - 1)
    want = (0.02 - (-0.01)) * d0 * dt * (3.2 + 0.08 * np.float32(3.6))
    for k in range(3):
        diff = runs[0][k][m] - runs[1][k][m]
        np.testing.assert_allclose(diff, want if k == axis else 0.0, atol=1e-13)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_buoyancy_density_gradient_breakup_term(axis):
    """SH/apply-buoyancy.wgsl.ts:62-70, 81: dg = central difference of density / 2,
    surfaceBreakup = smoothstep(0.008, 0.06, |dg|) coldSmoke, v += (-dg / |dg|) surfaceBreakup dt 2.8,
    coldSmoke = d (1 - smoothstep(0.12, 0.72, hotGas)) = d at T = 0. Flipping the sign of a density
    slope of 0.08 (|dg| >= 0.06) along one axis flips only this term (the shred term depends on |dg|):
    v(+s) - v(-s) = -2 d dt 2.8 in that component and 0 in the others.
    Planted mistakes this catches: +dg instead of -dg, an unnormalised direction, a wrong factor."""
    n, c0, d0, dt = 16, 8, 0.3, 0.05
    zyx = grid((n, n, n))
    coord = zyx[2 - axis]
    T = np.zeros((n, n, n))
    runs = []
    for s in (0.08, -0.08):
        den = np.full((n, n, n), d0)
        near = np.abs(coord - c0) <= 1
        den[near] = d0 + s * (coord[near] - c0)
        runs.append(_buoyancy_run(T, den, dt=dt))
    m = (coord == c0) & (zyx[2] > 0) & (zyx[2] < n - 1) & (zyx[0] > 0) & (zyx[0] < n - 1) & (zyx[1] > 0) & (zyx[1] < n - 1)
    for k in range(3):
        diff = runs[0][k][m] - runs[1][k][m]
        np.testing.assert_allclose(diff, -2 * d0 * dt * 2.8 if k == axis else 0.0, atol=1e-13)


# ------------------------------------------------------- shading helpers --
@pytest.mark.parametrize("g", [0.0, 0.32, -0.18, 0.7, -0.5])
def test_henyey_greenstein_phase_is_normalised(g):
    """RP:497-501: the phase function integrates to 1 over the sphere (2 pi int_-1^1 p(c) dc) and
    equals 1 / (4 pi) at g = 0. Planted mistakes: a (1 + g^2 + 2 g c) denominator still integrates
    to 1, so the forward peak p(1) = (1 + g) / (4 pi (1 - g)^2) is checked as well; exponent 1.5."""
    c, wts = np.polynomial.legendre.leggauss(400)
    vals = np.array([probe(PROBE_PHASE_HG, ci, g) for ci in c])
    assert abs(2 * math.pi * float(np.sum(wts * vals)) - 1.0) < 1e-9
    assert abs(probe(PROBE_PHASE_HG, 1.0, g) - (1 + g) / (4 * math.pi * (1 - g) ** 2)) < 1e-12
    if g == 0.0:
        assert abs(probe(PROBE_PHASE_HG, 0.3, 0.0) - 1 / (4 * math.pi)) < 1e-15


def test_fire_palette_and_smoke_albedo_endpoints():
    """RP:477-495. The palette is ember (0.18, 0.018, 0) at T <= 0.01 and white (1, 0.86, 0.56) at
    T >= 0.92; the orange stage (1, 0.42, 0.035) holds between 0.38 and 0.58. Smoke albedo of thin,
    cold smoke is the cool grey (0.26, 0.275, 0.295); fully sooted cold smoke is
    mix(cool, soot (0.04, 0.041, 0.045), 0.86). Planted mistakes: reversed mix order, wrong stage."""
    np.testing.assert_allclose(probe(PROBE_FIRE_PALETTE, 0.0), (0.18, 0.018, 0.0), atol=1e-15)
    np.testing.assert_allclose(probe(PROBE_FIRE_PALETTE, 1.0), (1.0, 0.86, 0.56), atol=1e-15)
    np.testing.assert_allclose(probe(PROBE_FIRE_PALETTE, 0.45), (1.0, 0.42, 0.035), atol=1e-15)
    np.testing.assert_allclose(probe(PROBE_SMOKE_ALBEDO, 0.05, 0.0, 0.0), (0.26, 0.275, 0.295), atol=1e-15)
    soot = 0.86
    want = np.array([0.26, 0.275, 0.295]) * (1 - soot) + np.array([0.04, 0.041, 0.045]) * soot
    np.testing.assert_allclose(probe(PROBE_SMOKE_ALBEDO, 0.95, 0.0, 0.0), want, atol=1e-15)


def test_light_direction_attenuation_and_power():
    """RP:503-524, 655-656: a point light at distance 10 is attenuated by 1 / (1 + 0.9 + 2) and
    points from the sample to the light; a directional light with |dir|^2 <= 1e-4 falls back to
    normalize(-0.34, 0.88, 0.31); power = raw / (1 + 0.18 raw), raw = max(intensity, 0) att.
    Planted mistakes: attenuation squared, direction from light to sample, negative intensity kept."""
    wp, lp = (1.0, 2.0, 3.0), (1.0 + 6.0, 2.0 + 8.0, 3.0)
    assert abs(probe(PROBE_LIGHT_ATT, 1.0, *lp, *wp) - 1 / 3.9) < 1e-15
    assert probe(PROBE_LIGHT_ATT, 0.0, *lp, *wp) == 1.0
    np.testing.assert_allclose(probe(PROBE_LIGHT_DIR, 1.0, 0, 0, 0, *lp, *wp), (0.6, 0.8, 0.0), atol=1e-15)
    d = np.array([-0.34, 0.88, 0.31])
    np.testing.assert_allclose(probe(PROBE_LIGHT_DIR, 0.0, 0.005, 0.0, 0.0, 0, 0, 0, *wp), d / np.linalg.norm(d),
                               atol=1e-15)
    np.testing.assert_allclose(probe(PROBE_LIGHT_DIR, 0.0, 0.0, 2.0, 0.0, 0, 0, 0, *wp), (0, 1, 0), atol=1e-15)
    raw = 2.0 / 3.9
    assert abs(probe(PROBE_LIGHT_POWER, 2.0, 1 / 3.9) - raw / (1 + 0.18 * raw)) < 1e-15
    assert probe(PROBE_LIGHT_POWER, -1.0, 1.0) == 0.0


# --------------------------------------------------- raymarch closed forms --
def _slab_scene(n, rho):
    o = Oracle(make_desc((n, n, n)), 64)
    o.write("density", np.full((n, n, n), rho))
    o.write("adv_density", np.full((n, n, n), 0.5))
    o.run_pass("activate_advected", scenes.sim_params())
    assert o.flags().all()
    return o


def _axis_camera(n):
    """Camera on the volume's vertical centre line looking along +z; the centre pixel of an odd
    image sees NDC (0, 0) exactly."""
    from fenix_amd.render import get_volume_world_bounds
    centre, half = get_volume_world_bounds((n, n, n))
    cam = abi.Camera()
    cam.position[:] = (0.0, centre[1], -60.0)
    cam.target[:] = (0.0, centre[1], 0.0)
    cam.forward[:] = (0.0, 0.0, 1.0)
    cam.right[:] = (1.0, 0.0, 0.0)
    cam.up[:] = (0.0, 1.0, 0.0)
    cam.aspect = 1.0
    cam.tan_half_fov_y = 0.4
    return cam, half


def _params(steps, mode=0, lights=()):
    p = abi.RenderParams()
    p.step_count, p.light_count = steps, len(lights)
    for i, (kind, vec, color, intensity) in enumerate(lights):
        p.lights[i].type = kind
        if kind == 1:
            p.lights[i].position[:] = vec
        else:
            p.lights[i].direction[:] = vec
        p.lights[i].color[:] = color
        p.lights[i].intensity = intensity
    p.scattering_forward, p.scattering_back = 0.32, -0.18
    p.display_mode = mode
    return p


def _slab_march(n, rho, steps, half_z):
    """Per-sample quantities of a uniform slab on the centre ray (RP:573-724): den = rho - 0.004,
    |gradient| floored at 1e-4 (detail 5.2e-4), dd = den^0.96 mix(0.82, 1.42, detail),
    stretch m = mix(1.02, 0.76, 1.35 dd); every sample sees the same values, and the march takes
    min(steps, floor(steps / m) + 1) samples of world length (2 half_z / steps) m."""
    den = rho - 0.004
    dd = den ** 0.96 * (0.82 + 0.6 * 5.2e-4)
    m = 1.02 - 0.26 * (dd * 1.35)
    ns = min(steps, int(math.floor(steps / m)) + 1)
    return den, dd, m, ns, 2 * half_z / steps * m


def _accumulated_alpha(op, ns):
    """alpha after the march: n samples of opacity op, stopping early once alpha > 0.985 (RP:612-614)."""
    alpha = 0.0
    for _ in range(ns):
        if alpha > 0.985:
            break
        alpha += (1 - alpha) * op
    return alpha


@pytest.mark.parametrize("mode", [0, 1])
def test_raymarch_uniform_slab_closed_form(mode):
    """Unlit uniform slab (RP:632-724): opacity per sample 1 - exp(-4.75 dd L) (temperature mode)
    or 1 - exp(-8.6 dd L) (density mode), alpha = 1 - (1 - op)^n, n cut where alpha first exceeds
    0.985 (the density mode reaches it after 23 of 32 samples); the colour is premultiplied:
    temperature mode's unlit colour is smokeAlbedo * ambient = cool grey * (0.045, 0.048, 0.052),
    density mode's mix((0.01, 0.01, 0.012), (0.88, 0.9, 0.94), 1.4 dd). Planted mistakes this catches:
    4.5 for 4.75, the voxel-space step in the opacity, a missing step stretch, straight alpha."""
    n, steps, rho = 16, 32, 0.017
    o = _slab_scene(n, rho)
    cam, half = _axis_camera(n)
    img = o.raymarch(cam, _params(steps, mode), 5, 5)
    den, dd, m, ns, L = _slab_march(n, rho, steps, half[2])
    k = 4.75 if mode == 0 else 8.6
    op = 1 - math.exp(-k * dd * L)
    alpha = _accumulated_alpha(op, ns)
    if mode == 0:
        comp = np.array([0.26 * 0.045, 0.275 * 0.048, 0.295 * 0.052])
    else:
        t = min(dd * 1.4, 1.0)
        comp = np.array([0.01, 0.01, 0.012]) * (1 - t) + np.array([0.88, 0.9, 0.94]) * t
    px = img[2, 2]
    assert abs(px[3] - alpha) < 2e-6, (px[3], alpha)
    np.testing.assert_allclose(px[:3], comp * alpha, rtol=2e-6)


def test_raymarch_thin_smoke_is_skipped_below_the_density_threshold():
    """RP:625-628: samples with den < 0.012 and hot < 0.04 only stretch the step (1.8x) and add
    nothing: rho = 0.0155 (den 0.0115) leaves the pixel discarded (0); rho = 0.0165 (den 0.0125)
    is accumulated. Planted mistake: a 0.01 threshold."""
    n = 16
    cam, _ = _axis_camera(n)
    assert np.all(_slab_scene(n, 0.0155).raymarch(cam, _params(32), 5, 5)[2, 2] == 0)
    assert _slab_scene(n, 0.0165).raymarch(cam, _params(32), 5, 5)[2, 2, 3] > 0.5


@pytest.mark.parametrize("primary", [True, False])
def test_raymarch_lit_slab_closed_form(primary):
    """A directional light straight up (0, 1, 0) over the uniform slab, colour c, intensity I
    (RP:534-549, 558-561, 640-683). On the centre ray dl = (0,1,0), -rayDir = (0,0,-1): the normal
    falls back to (0,1,0) (|gradient| <= 1.1e-4), so nf = 1, surface = 1, rim = 0.5^1.2, cos = 0;
    directional = (0.42 + 0.22 rim) P, scatter = (5.2 HG(0, 0.32) + 1.55 HG(0, -0.18)) P 0.36, with
    P = I / (1 + 0.18 I). As light 0 it is multiplied by the 3-tap primary transmittance
    exp(-2.15 (0.34 dd + (1.15 + 0.72 + 0.4) den)); as light 1 (after a zero-intensity light 0) it
    is not. Then tl = 1 - exp(-0.72 total), lit = albedo (ambient + 0.92 tl) + 0.18 smoothstep(0.018,
    0.22, dd) tl. Planted mistakes this catches: transmittance applied to lights 1-2, swapped tap
    weights, the rim exponent, the tonemap factor, the absorb constant."""
    n, steps, rho, I = 16, 32, 0.017, 1.7
    color = np.array([1.0, 0.9, 0.7])
    light = (0, (0.0, 1.0, 0.0), tuple(color), I)
    lights = (light,) if primary else ((0, (0.0, 1.0, 0.0), (1.0, 1.0, 1.0), 0.0), light)
    o = _slab_scene(n, rho)
    cam, half = _axis_camera(n)
    px = o.raymarch(cam, _params(steps, 0, lights), 5, 5)[2, 2]
    den, dd, m, ns, L = _slab_march(n, rho, steps, half[2])
    P = I / (1 + 0.18 * I)
    hg = lambda g: (1 - g * g) / (4 * math.pi * (1 + g * g) ** 1.5)
    directional = (0.42 + 0.22 * 0.5 ** 1.2) * P
    scatter = (5.2 * hg(0.32) + 1.55 * hg(-0.18)) * P * 0.36
    trans = math.exp(-2.15 * (0.34 * dd + (1.15 + 0.72 + 0.4) * den)) if primary else 1.0
    total = color * trans * (directional + scatter)
    tl = 1 - np.exp(-0.72 * total)
    albedo = np.array([0.26, 0.275, 0.295])
    t = min(max((dd - 0.018) / (0.22 - 0.018), 0.0), 1.0)
    sm = t * t * (3 - 2 * t)
    lit = albedo * np.array([0.045, 0.048, 0.052]) + albedo * tl * 0.92 + tl * sm * 0.18
    op = 1 - math.exp(-4.75 * dd * L)
    alpha = _accumulated_alpha(op, ns)
    assert abs(px[3] - alpha) < 2e-6
    np.testing.assert_allclose(px[:3], lit * alpha, rtol=2e-6)


def test_raymarch_secondary_lights_skip_primary_transmittance():
    """RP:640-683 (trans = 1 for lights 1 and 2) on a non-uniform blob: a light as light 1 or as
    light 2 behind zero-intensity lights gives the same image bit for bit; as light 0 it is
    shadowed, so no pixel is brighter and some are darker; alpha does not depend on lighting."""
    n = 24
    o = Oracle(make_desc((n, n, n)), 64)
    z, y, x = grid((n, n, n))
    r2 = (x - 12.0) ** 2 + (y - 10.0) ** 2 + (z - 12.0) ** 2
    o.write("density", 0.6 * np.exp(-r2 / 30.0))
    o.write("temperature", 0.3 * np.exp(-r2 / 12.0))
    o.write("adv_density", 0.6 * np.exp(-r2 / 30.0))
    o.run_pass("activate_advected", scenes.sim_params())
    from fenix_amd.render import camera_struct, orbit_camera_snapshot
    cam = camera_struct(orbit_camera_snapshot())
    b = (1, (6.0, 30.0, -8.0), (1.0, 0.8, 0.6), 3.0)
    z0 = (0, (-0.3, 0.9, 0.2), (1.0, 1.0, 1.0), 0.0)
    z1 = (0, (0.5, 0.5, 0.1), (1.0, 1.0, 1.0), 0.0)
    i1 = o.raymarch(cam, _params(64, 0, (z0, b)), 24, 24)
    i2 = o.raymarch(cam, _params(64, 0, (z0, z1, b)), 24, 24)
    i0 = o.raymarch(cam, _params(64, 0, (b,)), 24, 24)
    np.testing.assert_array_equal(i1, i2)
    np.testing.assert_array_equal(i0[..., 3], i1[..., 3])
    assert np.all(i0[..., :3] <= i1[..., :3] + 1e-7)
    assert (i1[..., :3] - i0[..., :3]).max() > 1e-3


# ------------------------------------------------------------- hash noise --
def test_hash_noise_uses_the_correctly_rounded_float_sine():
    """SH/inject-source-common.wgsl.ts:102-121: hash31(p) = fract(sin(dot(p, (127.1, 311.7, 74.7)))
    * 43758.5453123) in f32. WGSL leaves the precision of f32 sin at these arguments (hundreds to
    thousands of radians) to the implementation, and the factor 43758 turns a one-ulp difference of
    the sine into ~3e-3 of hash; both the oracle and the kernels take the double sine rounded to
    float. Checked bit for bit against that definition (f32 dot product in WGSL order, no FMA);
    value noise equals the hash at lattice points (u = 0) and stays in [0, 1).
    Planted mistakes: the f32 library sine, the dot product in double, a missing fract."""
    from oracle import PROBE_EMITTER_NOISE, PROBE_HASH31
    rng = np.random.default_rng(12)
    pts = np.round(rng.uniform(-40, 40, (400, 3)), 2).astype(np.float32)
    for p in pts:
        d = np.float32(np.float32(p[0] * np.float32(127.1)) + np.float32(p[1] * np.float32(311.7)))
        d = np.float32(d + np.float32(p[2] * np.float32(74.7)))
        s = np.float32(np.sin(np.float64(d)))
        v = np.float32(s * np.float32(43758.5453123))
        want = np.float32(v - np.floor(v))
        got = probe(PROBE_HASH31, *p, bits=32)
        assert got == float(want), (p, got, want)
    for c in np.floor(pts[:50]):
        n = probe(PROBE_EMITTER_NOISE, *c, 1.0, 0.0, bits=32)  # scale 1, seed 0: the lattice point itself
        assert n == probe(PROBE_HASH31, *c, bits=32) and 0.0 <= n < 1.0


# --------------------------------------------------------- 16-bit storage --
@pytest.mark.parametrize("fmt", ["bf16", "f16"])
def test_stochastic_rounding_is_unbiased_and_brackets_the_value(fmt):
    """The 16-bit stores of FX_PRECISION_BF16/FP16 round up with probability equal to the dropped
    fraction of an ulp, the decision taken from the low bits of a hash (bf16: 16 bits, half: 13
    bits for normal values, the top d bits for subnormal halves with d dropped bits). Over every
    value of those hash bits the mean of the stored values equals the input exactly, each stored
    value is one of the two representable neighbours, and representable inputs are unchanged.
    Planted mistakes: '<=' for '<' (biased up by 1/2^k), rounding down only, a wrong shift."""
    from oracle import PROBE_SR_BF16, PROBE_SR_F16
    fn = PROBE_SR_BF16 if fmt == "bf16" else PROBE_SR_F16
    rng = np.random.default_rng(7)
    vals = list(np.float32(rng.uniform(0, 1, 4))) + [np.float32(3.1e-6), np.float32(4.5e-5)]  # two subnormal halves
    for v in vals:
        v64 = float(v)
        if fmt == "bf16":
            hs = np.arange(65536)
        elif v64 >= 2.0 ** -14:
            hs = np.arange(8192)
        else:
            e = int(np.frexp(v)[1]) - 1 + 127          # biased f32 exponent
            dbits = 126 - e
            hs = np.arange(1 << min(dbits, 16)) << (32 - min(dbits, 16))
        out = np.array([probe(fn, v64, int(h), bits=32) for h in hs])
        lo, hi = out.min(), out.max()
        assert set(np.unique(out)) <= {lo, hi}
        assert lo <= v64 <= hi
        if fmt == "bf16" or v64 >= 2.0 ** -14 or dbits <= 16:
            assert abs(out.mean() - v64) <= 1e-12 * max(v64, 1e-30), (v, out.mean())
    for v in (0.5, 0.25 + 2.0 ** -9, 0.0):  # representable in both formats
        for h in (0, 1, 0xffffffff, 0x12345678):
            assert probe(fn, v, h, bits=32) == v


def test_fixed_point_rounding_pins():
    """FX_PRECISION_U16 stores u * 2^-16 with u in [0, 65535] after clamping to [0, 1]. The
    stochastic store rounds up when the top 24 bits of the hash, as an integer, are below the
    dropped fraction times 2^24: for a value whose fraction of a step is m / 2^k, the hashes
    j << (32 - k) (j < 2^k) round up exactly for j < m, so their mean is the input exactly
    (k is kept so that the input fits a float's 24 bits).
    Nearest-even stores: ties go to the even step; clamping at both ends. Planted mistakes:
    '<=' for '<' (mean biased by 1/4096 of a step), a dropped clamp, round-half-up ties."""
    from oracle import PROBE_RNE_U16, PROBE_SR_U16
    step = 2.0 ** -16
    for n, m, k in ((0, 1, 12), (7, 3000, 12), (4095, 2048, 12), (12345, 128, 8), (65000, 255, 8), (65534, 17, 8)):
        v = (n + m / 2 ** k) * step
        assert float(np.float32(v)) == v
        out = np.array([probe(PROBE_SR_U16, v, j << (32 - k), bits=32) for j in range(1 << k)])
        assert set(np.unique(out)) == {n * step, (n + 1) * step}
        assert int((out > v).sum()) == m
        assert abs(out.mean() - v) <= 1e-15
        half = 1 << (k - 1)
        assert probe(PROBE_RNE_U16, v, bits=32) == (n + (m > half or (m == half and n % 2 == 1))) * step
    for v in (0.0, 0.5, 12345 * step, 65535 * step):  # representable: unchanged for every hash
        for h in (0, 1, 0xffffffff, 0x12345678):
            assert probe(PROBE_SR_U16, v, h, bits=32) == v
        assert probe(PROBE_RNE_U16, v, bits=32) == v
    for v, want in ((-0.25, 0.0), (1.0, 65535 * step), (3.5, 65535 * step), (1.0 - step / 4, 65535 * step)):
        assert probe(PROBE_RNE_U16, v, bits=32) == want
        assert probe(PROBE_SR_U16, v, 0xffffffff, bits=32) == want
    assert probe(PROBE_RNE_U16, 2.5 * step, bits=32) == 2 * step      # tie to even
    assert probe(PROBE_RNE_U16, 3.5 * step, bits=32) == 4 * step


def test_unsigned_float16_rounding_pins():
    """FX_PRECISION_UF16 stores an unsigned float with 4 exponent and 12 mantissa bits: values
    (1 + m/4096) * 2^(E-15) for E = 1..15, multiples of 2^-26 below 2^-14, clamped to [0, 1].
    Pinned by construction from that definition (numpy's frexp/ldexp, not the oracle's bit
    tricks): every input rounds to one of the two neighbours of its representable bracket; over
    all 2048 values of the 11 low hash bits (normal values) or all top-k hash values (small
    values, fraction m / 2^k) the mean is the input exactly; nearest even picks the closer
    neighbour, ties to an even code; representable values, 0 and 1 are unchanged; clamping at
    both ends. Planted mistakes: '<=' for '<', a mantissa of 11 or 13 bits, a wrong subnormal
    step or threshold, an exponent off by one."""
    from oracle import PROBE_RNE_UF16, PROBE_SR_UF16

    def bracket(v):
        if v < 2.0 ** -14:
            step = 2.0 ** -26
        else:
            e = np.frexp(v)[1] - 1          # v in [2^e, 2^(e+1))
            step = 2.0 ** (e - 12)
        lo = np.floor(v / step) * step
        return lo, lo + step, step

    rng = np.random.default_rng(11)
    vals = [float(np.float32(x)) for x in rng.uniform(0, 1, 6)] + [float(np.float32(x)) for x in 10 ** rng.uniform(-4.2, -1, 6)]
    for v in vals:
        lo, hi, step = bracket(v)
        out = np.array([probe(PROBE_SR_UF16, v, h, bits=32) for h in range(2048)])
        assert set(np.unique(out)) <= {lo, hi}, v
        assert abs(out.mean() - v) <= 1e-12 * v, (v, out.mean())
        r = probe(PROBE_RNE_UF16, v, bits=32)
        code_lo = round(lo / step)
        want = lo if (v - lo < hi - v or (v - lo == hi - v and code_lo % 2 == 0)) else hi
        assert r == want, (v, r, want)
    for n, m, k in ((0, 1, 12), (1000, 77, 12), (4095, 2048, 12)):  # below 2^-14: steps of 2^-26
        v = (n + m / 2 ** k) * 2.0 ** -26
        out = np.array([probe(PROBE_SR_UF16, v, j << (32 - k), bits=32) for j in range(1 << k)])
        assert int((out > v).sum()) == m
        assert abs(out.mean() - v) <= 1e-12 * v
    for v in (0.0, 1.0, 0.5, 0.75 + 2.0 ** -13, 2.0 ** -14, 3 * 2.0 ** -26, (4096 + 4095) / 4096 * 2.0 ** -2):
        for h in (0, 1, 0x7ff, 0xffffffff, 0x12345678):
            assert probe(PROBE_SR_UF16, v, h, bits=32) == v
        assert probe(PROBE_RNE_UF16, v, bits=32) == v
    for v, want in ((-0.5, 0.0), (1.5, 1.0), (7.0, 1.0)):
        assert probe(PROBE_RNE_UF16, v, bits=32) == want
        assert probe(PROBE_SR_UF16, v, 0xffffffff, bits=32) == want
    assert probe(PROBE_RNE_UF16, 1.0 - 2.0 ** -15, bits=32) == 1.0            # top binade: step 2^-13, tie -> even
    assert probe(PROBE_RNE_UF16, 0.5 + 2.0 ** -14, bits=32) == 0.5            # tie between codes 0 and 1 of [0.5, 1)
    assert probe(PROBE_RNE_UF16, 0.5 + 3 * 2.0 ** -14, bits=32) == 0.5 + 2.0 ** -12  # codes 1 and 2: to 2
