"""Mixed precision: density, temperature and fuel stored in BF16, FP16 or 16-bit
fixed point (FP32 arithmetic, FP32 reaction/velocity/pressure).

Parity is checked against the FP32 oracle with the same round-to-nearest-even
on store; accuracy is reported against the FP64 oracle (C4-style scene at a
size the FP64 oracle finishes in seconds).
"""
import json
import os

import numpy as np
import pytest

import scenes
from fenix_amd import abi
from fenix_amd.emitters import ScalarEmitter
from fenix_amd.simulation import CombustionVolumeSimulation, make_desc
from oracle import Oracle

pytestmark = pytest.mark.gpu

PERSISTENT = ("density", "temperature", "fuel", "reaction", "velocity_x", "velocity_y", "velocity_z")
# One 16-bit ulp flip in a handful of voxels moves rel L2 by ~1e-5 (bf16: 2^-8 relative).
# Fixed point: one step is 2^-16 of the unit range, below a half ulp for values above 1/64.
TOL = {abi.FX_PRECISION_BF16: 2e-3, abi.FX_PRECISION_FP16: 3e-4, abi.FX_PRECISION_BF16_RNE: 2e-3,
       abi.FX_PRECISION_FP16_RNE: 3e-4, abi.FX_PRECISION_U16: 3e-4, abi.FX_PRECISION_U16_RNE: 3e-4,
       abi.FX_PRECISION_UF16: 3e-4, abi.FX_PRECISION_UF16_RNE: 3e-4}
SIXTEEN = [abi.FX_PRECISION_BF16, abi.FX_PRECISION_FP16, abi.FX_PRECISION_BF16_RNE, abi.FX_PRECISION_FP16_RNE,
           abi.FX_PRECISION_U16, abi.FX_PRECISION_U16_RNE, abi.FX_PRECISION_UF16, abi.FX_PRECISION_UF16_RNE]


def rel_l2(a, b):
    den = float(np.linalg.norm(b))
    num = float(np.linalg.norm(np.asarray(a, np.float64) - np.asarray(b, np.float64)))
    return num / den if den > 0 else num


def c4_sources(n, seed=4, count=4, radius_vox=None):
    rng = scenes.SplitMix64(seed)
    r = (radius_vox or max(3.0, n / 25.6)) / n
    out = []
    for _ in range(count):
        x = r + (1 - 2 * r) * rng.uniform()
        z = r + (1 - 2 * r) * rng.uniform()
        out.append(ScalarEmitter(position=(x, r, z), radius=r, densityRate=1.0, heatRate=2.0, fuelRate=1.5))
    return out


@pytest.mark.parametrize("prec", SIXTEEN)
@pytest.mark.parametrize("pass_name", ["inject_scalar", "inject_igniter", "advect_scalars", "combustion", "buoyancy",
                                       "confinement"])
def test_pass_parity_16bit(prec, pass_name):
    grid = (24, 16, 20)
    desc = make_desc(grid, precision=prec, runtime={"worldSize": 24.0})
    sim = CombustionVolumeSimulation(desc)
    orc = Oracle(desc, 32)
    srcs = scenes.mixed_sources()
    sim.updateSources(srcs)
    orc.update_sources(*scenes.packed(srcs))
    st = scenes.random_state(grid[::-1], seed=21)
    for k, v in st.items():
        sim.write_field(k, v)
        orc.write(k, v)
    # the written 16-bit fields are identical on both sides
    np.testing.assert_array_equal(sim.read_field("density"), orc.read("density").astype(np.float32))
    p = scenes.sim_params(time=0.4, dt=1.0 / 30.0, dx=1.0, world=24.0)
    for name in ("activate_render", "activate_advected"):
        sim.run_pass(name, p)
        orc.run_pass(name, p)
    sim.run_pass(pass_name, p)
    orc.run_pass(pass_name, p)
    for name in abi.FIELDS:
        e = rel_l2(sim.read_field(name), orc.read(name))
        assert e <= TOL[prec], f"{pass_name}/{prec}: {name} rel L2 {e:.3e}"


@pytest.mark.parametrize("prec", SIXTEEN)
def test_advect_scalars_16bit_long_backtraces(prec):
    """Backtraces of up to 3 voxels (velocity +-3, dt 1): the scalar advection's samples mostly
    leave its LDS tile (decoded 16-bit values) and decode their corners from memory instead;
    both against the oracle with the same storage."""
    grid = (24, 16, 20)
    desc = make_desc(grid, precision=prec, runtime={"worldSize": 24.0})
    sim = CombustionVolumeSimulation(desc)
    orc = Oracle(desc, 32)
    st = scenes.random_state(grid[::-1], seed=22)
    for k, v in st.items():
        sim.write_field(k, v)
        orc.write(k, v)
    p = scenes.sim_params(time=0.4, dt=1.0, dx=1.0, world=24.0)
    for name in ("activate_render", "activate_advected", "advect_scalars"):
        sim.run_pass(name, p)
        orc.run_pass(name, p)
    for name in abi.FIELDS:
        e = rel_l2(sim.read_field(name), orc.read(name))
        assert e <= TOL[prec], f"advect_scalars/{prec}: {name} rel L2 {e:.3e}"


@pytest.mark.parametrize("prec", SIXTEEN)
def test_step_parity_16bit_sparse(prec):
    n = 48
    desc = scenes.c2_desc(n, precision=prec)
    srcs = scenes.c2_sources(n, radius_vox=4.0)
    sim = CombustionVolumeSimulation(desc)
    orc = Oracle(desc, 32)
    sim.updateSources(srcs)
    orc.update_sources(*scenes.packed(srcs))
    for k in range(20):
        if k > 0:
            for name in PERSISTENT:
                sim.write_field(name, orc.read(name))
        sim.step(None, k / 60.0, 1 / 60.0)
        orc.step(k / 60.0, 1 / 60.0)
        np.testing.assert_array_equal(sim.read_flags(), orc.flags())
        for name in PERSISTENT:
            e = rel_l2(sim.read_field(name), orc.read(name))
            assert e <= TOL[prec], f"step {k}: {name} rel L2 {e:.3e}"


def test_precision_report_against_fp64():
    """C4-style sparse scene (4 seeded emitters, config mode) for 100 steps:
    per-field relative L2 and max error of FP32 / BF16 / FP16 / fixed-point storage (stochastic
    rounding, and round-to-nearest for comparison) against the FP64 oracle. FP32 and UF16
    storage must stay within 1e-3 relative L2."""
    n, steps = 64, 100
    srcs = c4_sources(n)
    ref = Oracle(scenes.c2_desc(n), 64)
    ref.update_sources(*scenes.packed(srcs))
    for k in range(steps):
        ref.step(k / 60.0, 1 / 60.0)
    truth = {f: ref.read(f) for f in PERSISTENT}
    report = {"scene": f"sparse {n}^3, 4 floor emitters (seed 4), config mode, {steps} steps",
              "active_bricks_fp64": ref.active_count(), "fields": {}}
    for label, prec in (("fp32", abi.FX_PRECISION_FP32), ("bf16", abi.FX_PRECISION_BF16), ("fp16", abi.FX_PRECISION_FP16),
                        ("bf16_rne", abi.FX_PRECISION_BF16_RNE), ("fp16_rne", abi.FX_PRECISION_FP16_RNE),
                        ("u16", abi.FX_PRECISION_U16), ("u16_rne", abi.FX_PRECISION_U16_RNE),
                        ("uf16", abi.FX_PRECISION_UF16), ("uf16_rne", abi.FX_PRECISION_UF16_RNE)):
        sim = CombustionVolumeSimulation(scenes.c2_desc(n, precision=prec))
        sim.updateSources(srcs)
        for k in range(steps):
            sim.step(None, k / 60.0, 1 / 60.0)
        row = {}
        for f in PERSISTENT:
            g = sim.read_field(f).astype(np.float64)
            row[f] = {"rel_l2": rel_l2(g, truth[f]), "max_abs": float(np.max(np.abs(g - truth[f]))),
                      "ref_max": float(np.max(np.abs(truth[f])))}
        report["fields"][label] = row
        report.setdefault("active_bricks", {})[label] = sim.active_brick_count()
        sim.dispose()
    # reports go to FENIX_REPORT_DIR (default: reports/ at the repository root)
    out = os.environ.get("FENIX_REPORT_DIR", os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "reports"))
    os.makedirs(out, exist_ok=True)
    with open(os.path.join(out, "precision_report.json"), "w") as fh:
        json.dump(report, fh, indent=1)
    for f in ("density", "temperature", "fuel", "velocity_y"):
        assert report["fields"]["fp32"][f]["rel_l2"] <= 1e-3, (f, report["fields"]["fp32"][f])
    # the unsigned 4/12-bit float with stochastic rounding meets the bound in every field
    for f in PERSISTENT:
        assert report["fields"]["uf16"][f]["rel_l2"] <= 1e-3, (f, report["fields"]["uf16"][f])
    for label in ("bf16", "fp16", "bf16_rne", "fp16_rne", "u16", "u16_rne", "uf16", "uf16_rne"):
        for f in PERSISTENT:
            assert np.isfinite(report["fields"][label][f]["rel_l2"])
            assert report["fields"][label][f]["rel_l2"] < 0.2, (label, f, report["fields"][label][f])


def test_uf16_small_codes_decode_exactly_on_device():
    """UF16 codes below 2^-14 (multiples of 2^-26) decode on the device through an f32 subnormal
    scaled by 2^112, so they are exact only under IEEE f32 denormals. The threshold rule compares
    the decoded density with a = 2^-20 (code 64): a brick holding code 65 activates (with its 26
    neighbours), a brick holding exactly a does not; flushed denormals would decode both as 0.
    Flags bit-exact against the oracle; the written values read back unchanged."""
    grid = (32, 32, 32)
    a = 2.0 ** -20
    desc = make_desc(grid, precision=abi.FX_PRECISION_UF16, activation_rule=abi.FX_ACTIVATION_THRESHOLD,
                     activation_density=a, runtime={"worldSize": 32.0})
    sim = CombustionVolumeSimulation(desc)
    orc = Oracle(desc, 32)
    dens = np.zeros(grid[::-1], np.float32)
    dens[2, 3, 4] = a + 2.0 ** -26        # brick (0, 0, 0)
    dens[20, 20, 20] = a                  # brick (2, 2, 2)
    dens[28, 12, 4] = 3 * 2.0 ** -26      # brick (0, 1, 3)
    for s in (sim, orc):
        (s.write_field if s is sim else s.write)("density", dens)
    np.testing.assert_array_equal(sim.read_field("density"), dens)
    p = scenes.sim_params(time=0.0, dt=1.0 / 60.0, dx=1.0, world=32.0)
    sim.run_pass("activate_render", p)
    orc.run_pass("activate_render", p)
    got = sim.read_flags(abi.FX_FLAGS_INJECTED)
    np.testing.assert_array_equal(got, orc.flags(abi.FX_FLAGS_INJECTED))
    on = got.reshape(4, 4, 4) != 0
    assert on[0, 0, 0] and on[1, 1, 1] and not on[2, 2, 2] and not on[3, 1, 0]
