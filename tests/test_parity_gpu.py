"""GPU parity: libfenix (HIP, gfx950) against the FP32 oracle on identical inputs.

Bars: brick flags, active list and page table bit-exact; FP32 fields within
1e-5 relative L2 per pass and per step.
"""
import numpy as np
import pytest

import scenes
from fenix_amd import abi
from fenix_amd.render import RenderLight, VolumeRaymarchPass, camera_struct, orbit_camera_snapshot
from fenix_amd.simulation import CombustionVolumeBatch, CombustionVolumeSimulation, make_desc
from oracle import Oracle

pytestmark = pytest.mark.gpu

REL_L2 = 1e-5
PERSISTENT = ("density", "temperature", "fuel", "reaction", "velocity_x", "velocity_y", "velocity_z")


def rel_l2(a, b):
    den = float(np.linalg.norm(b))
    num = float(np.linalg.norm(np.asarray(a, np.float64) - np.asarray(b, np.float64)))
    return num / den if den > 0 else num


def make_pair(desc, sources=None):
    sim = CombustionVolumeSimulation(desc)
    orc = Oracle(desc, 32)
    if sources is not None:
        sim.updateSources(sources)
        orc.update_sources(*scenes.packed(sources))
    return sim, orc


def load_state(sim, orc, st):
    for k, v in st.items():
        sim.write_field(k, v)
        orc.write(k, v)


def compare_all(sim, orc, label, tol=REL_L2, fields=None):
    worst = []
    for name in fields or abi.FIELDS:
        g = sim.read_field(name)
        o = orc.read(name)
        e = rel_l2(g, o)
        worst.append((e, name))
        assert np.all(np.isfinite(g)), f"{label}: non-finite values in {name}"
        assert e <= tol, f"{label}: field {name} rel L2 {e:.3e} > {tol:g}"
    return max(worst)


GRID = (24, 16, 20)  # x not, z not a multiple of 8 in one axis: exercises brick padding


@pytest.mark.parametrize("rule", [abi.FX_ACTIVATION_REFERENCE, abi.FX_ACTIVATION_THRESHOLD])
def test_activation_bit_exact(rule):
    desc = make_desc(GRID, activation_rule=rule, runtime={"worldSize": 24.0})
    sim, orc = make_pair(desc, scenes.mixed_sources())
    st = scenes.random_state(GRID[::-1], seed=3)
    # sparse energy pattern: single hot voxels in some bricks
    load_state(sim, orc, st)
    p = scenes.sim_params(dx=1.0, world=24.0)
    for name, which in (("activate_render", abi.FX_FLAGS_INJECTED), ("activate_advected", abi.FX_FLAGS_RENDER)):
        sim.run_pass(name, p)
        orc.run_pass(name, p)
        np.testing.assert_array_equal(sim.read_flags(which), orc.flags(which))


def _prepare_flags(sim, orc, p):
    for name in ("activate_render", "activate_advected"):
        sim.run_pass(name, p)
        orc.run_pass(name, p)


@pytest.mark.parametrize("pass_name", [
    "inject_scalar", "inject_velocity", "inject_igniter", "advect_velocity", "advect_scalars", "combustion",
    "buoyancy", "vorticity", "confinement", "divergence", "projection", "debug_fields",
])
def test_pass_parity_dense(pass_name):
    desc = make_desc(GRID, runtime={"worldSize": 24.0, "wind": (0.4, 0.0, 0.2), "windStrength": 1.5})
    sim, orc = make_pair(desc, scenes.mixed_sources())
    st = scenes.random_state(GRID[::-1], seed=11)
    load_state(sim, orc, st)
    p = scenes.sim_params(time=0.7, dt=1.0 / 30.0, dx=1.0, world=24.0, wind=(0.4, 0.0, 0.2, 1.5))
    _prepare_flags(sim, orc, p)
    sim.run_pass(pass_name, p)
    orc.run_pass(pass_name, p)
    compare_all(sim, orc, pass_name)


def test_semi_lagrangian_parity():
    desc = make_desc(GRID, advection_mode=abi.FX_ADVECT_SEMI_LAGRANGIAN, runtime={"worldSize": 24.0})
    sim, orc = make_pair(desc, scenes.mixed_sources())
    load_state(sim, orc, scenes.random_state(GRID[::-1], seed=12))
    p = scenes.sim_params(time=0.3, dt=1.0 / 20.0, dx=1.0, world=24.0)
    _prepare_flags(sim, orc, p)
    sim.run_pass("advect_scalars", p)
    orc.run_pass("advect_scalars", p)
    compare_all(sim, orc, "semi-lagrangian")


@pytest.mark.parametrize("mode", [abi.FX_ADVECT_MACCORMACK, abi.FX_ADVECT_SEMI_LAGRANGIAN])
@pytest.mark.parametrize("dt", [0.5, 1.0])
def test_advection_parity_with_long_backtraces(mode, dt):
    """The advection kernels read trilinear corners from LDS tiles of the brick +- 1 voxel and
    gather from the fields in memory where a sample's corners leave the tile. The plume's
    backtraces stay below 0.2 voxels, so the other parity tests never take that path: here the
    velocities (+-3 voxels per unit time) with dt 0.5 and 1 give backtraces of up to 1.5 and 3
    voxels per axis, a mix of tile and memory samples (and clamped ones at the domain faces),
    both advection passes against the oracle."""
    desc = make_desc(GRID, advection_mode=mode, runtime={"worldSize": 24.0})
    sim, orc = make_pair(desc, scenes.mixed_sources())
    st = scenes.random_state(GRID[::-1], seed=13)
    load_state(sim, orc, st)
    # the share of voxels whose velocity backtrace leaves the tile along some axis
    u = np.stack([st[f"velocity_{a}"] for a in "xyz"])
    assert float(np.mean(np.any(np.abs(u) * dt >= 1.0, axis=0))) > 0.5
    p = scenes.sim_params(time=0.4, dt=dt, dx=1.0, world=24.0)
    _prepare_flags(sim, orc, p)
    for name in ("advect_velocity", "advect_scalars"):
        sim.run_pass(name, p)
        orc.run_pass(name, p)
        compare_all(sim, orc, f"{name}, dt {dt}")


@pytest.mark.parametrize("solver,quality,grid,nlev", [
    (abi.FX_SOLVER_REFERENCE, {}, (32, 24, 16), 3),
    (abi.FX_SOLVER_REFERENCE, {"finePreIterations": 3, "midPreIterations": 1, "coarseIterations": 7, "finePostIterations": 0},
     (32, 24, 16), 3),
    (abi.FX_SOLVER_VCYCLE, {}, (32, 24, 16), 3),
    # coarsest level 20x16x16: 3528 active voxels staged in LDS, several slots per thread
    (abi.FX_SOLVER_VCYCLE, {}, (160, 128, 128), 4),
    # coarsest level 32^3: 27000 active voxels exceed the LDS staging, global sweeps
    (abi.FX_SOLVER_VCYCLE, {}, (64, 64, 64), 2),
])
def test_pressure_parity(solver, quality, grid, nlev):
    desc = make_desc(grid, solver=solver, quality=quality, vcycle={"levels": nlev, "cycles": 3})
    sim, orc = make_pair(desc)
    st = scenes.random_state(grid[::-1], seed=5)
    sim.write_field("divergence", st["divergence"])
    orc.write("divergence", st["divergence"])
    p = scenes.sim_params()
    sim.run_pass("pressure", p)
    orc.run_pass("pressure", p)
    compare_all(sim, orc, "pressure", fields=["pressure", "divergence"])
    # The reference schedule fixes which buffer holds what (quirks Q1-Q4); the
    # V-cycle's ping-pong buffers are an implementation detail, only the
    # per-level right-hand sides and the final pressure are defined.
    buffers = (abi.FX_LEVEL_DIVERGENCE,) if solver == abi.FX_SOLVER_VCYCLE else (
        abi.FX_LEVEL_DIVERGENCE, abi.FX_LEVEL_PRESSURE_A, abi.FX_LEVEL_PRESSURE_B)
    for lvl in range(nlev):
        for which in buffers:
            g, o = sim.read_level(lvl, which), orc.level(lvl, which)
            assert rel_l2(g, o) <= REL_L2, f"level {lvl} buffer {which}: {rel_l2(g, o):.3e}"
    if solver == abi.FX_SOLVER_VCYCLE:
        rg, ro = sim.solver_residuals(), orc.residuals()
        assert len(rg) == len(ro) == 3
        np.testing.assert_allclose(rg, ro, rtol=1e-4)
        assert rg[-1] < rg[0]


def _step_parity(desc, sources, steps, init_velocity=None, tol=REL_L2):
    sim, orc = make_pair(desc, sources)
    if init_velocity is not None:
        for c, name in enumerate(("velocity_x", "velocity_y", "velocity_z")):
            sim.write_field(name, init_velocity[c])
            orc.write(name, init_velocity[c])
    dt = 1.0 / 60.0
    worst = 0.0
    for k in range(steps):
        if k > 0:  # per-step parity: start both from the oracle's state
            for name in PERSISTENT:
                sim.write_field(name, orc.read(name))
        sim.step(None, k * dt, dt)
        orc.step(k * dt, dt)
        np.testing.assert_array_equal(sim.read_flags(abi.FX_FLAGS_RENDER), orc.flags(abi.FX_FLAGS_RENDER),
                                      err_msg=f"render flags differ at step {k}")
        e, name = compare_all(sim, orc, f"step {k}", tol=tol, fields=list(PERSISTENT) + ["pressure", "divergence"])
        worst = max(worst, e)
    return sim, orc, worst


def test_step_parity_dense_reference():
    n = 32
    desc = scenes.c1_desc(n)
    srcs = scenes.c1_sources()
    _step_parity(desc, srcs, steps=12, init_velocity=scenes.c1_initial_velocity(n))


def test_step_parity_dense_vcycle():
    n = 32
    desc = scenes.c1_desc(n, config=True)
    _step_parity(desc, scenes.c1_sources(), steps=8)


def test_step_parity_sparse_config():
    n = 48
    desc = scenes.c2_desc(n)
    srcs = scenes.c2_sources(n, radius_vox=4.0)
    sim, orc, _ = _step_parity(desc, srcs, steps=40)
    # active list / page table are exact functions of the flags
    flags = orc.flags(abi.FX_FLAGS_RENDER)
    lst = sim.read_active_list()
    np.testing.assert_array_equal(lst, np.flatnonzero(flags).astype(np.uint32))
    page = sim.read_page_table()
    expect = np.full(flags.size, -1, np.int32)
    expect[lst] = np.arange(lst.size, dtype=np.int32)
    np.testing.assert_array_equal(page, expect)
    assert sim.active_brick_count() == orc.active_count() == lst.size
    # every field is zero outside the active bricks
    d = sim.read_field("density")
    bx = (n + 7) // 8
    mask = flags.reshape(bx, bx, bx).repeat(8, 0).repeat(8, 1).repeat(8, 2)[:n, :n, :n]
    assert np.all(d[mask == 0] == 0)


def test_step_parity_sparse_plume():
    """Sparse config from a developed plume (~15% of bricks active, the benchmark's start)."""
    n = 64
    srcs = scenes.c2_sources(n, radius_vox=2.5)
    st, frac = scenes.c2_plume_state(n, srcs, target=0.15)
    assert 0.1 < frac < 0.2
    desc = scenes.c2_desc(n)
    sim, orc = make_pair(desc, srcs)
    load_state(sim, orc, st)
    dt = 1.0 / 60.0
    for k in range(10):
        if k > 0:
            for name in PERSISTENT:
                sim.write_field(name, orc.read(name))
        sim.step(None, k * dt, dt)
        orc.step(k * dt, dt)
        np.testing.assert_array_equal(sim.read_flags(), orc.flags(), err_msg=f"flags differ at step {k}")
        compare_all(sim, orc, f"plume step {k}", fields=list(PERSISTENT) + ["pressure", "divergence"])
    assert sim.active_brick_count() == orc.active_count() > 0.1 * sim.brick_count


def test_vcycle_plans_bit_identical():
    """Split V-cycle legs (lean one-sweep kernels + residual/restriction and
    prolongation + one sweep, paired sweeps in LDS tiles, or the sweep next to the level
    change fused with the restriction or the prolongation) give bit-identical results to
    the fused LDS legs, and all match the oracle."""
    n = 64
    srcs = scenes.c2_sources(n, radius_vox=2.5)
    st, _ = scenes.c2_plume_state(n, srcs, target=0.15)
    desc = scenes.c2_desc(n)
    out, res = {}, {}
    for plan in ((0, 0), (0xF, 0), (0, 0xF), (0xF, 0xF), (0xF0F, 0xF0F), (0x10F, 0x30F),
                 (0xF000F, 0xF000F), (0x3000F, 0x1020F)):  # fused boundary sweeps (bits 16 + l)
        sim = CombustionVolumeSimulation(desc)
        sim.set_vcycle_plan(*plan)
        assert sim.vcycle_plan() == plan
        sim.updateSources(srcs)
        for k, v in st.items():
            sim.write_field(k, v)
        for k in range(3):
            sim.step(None, k / 60.0, 1 / 60.0)
        out[plan] = {f: sim.read_field(f) for f in PERSISTENT + ("pressure",)}
        res[plan] = np.array(sim.solver_residuals())
        sim.dispose()
    for plan, fields in out.items():
        for name, arr in fields.items():
            np.testing.assert_array_equal(arr, out[(0, 0)][name], err_msg=f"plan {plan}: {name}")
        # FP64 norms are sums of per-block atomics: equal up to summation order
        np.testing.assert_allclose(res[plan], res[(0, 0)], rtol=1e-12)
    orc = Oracle(desc, 32)
    orc.update_sources(*scenes.packed(srcs))
    for k, v in st.items():
        orc.write(k, v)
    for k in range(3):
        orc.step(k / 60.0, 1 / 60.0)
    for name in PERSISTENT + ("pressure",):
        assert rel_l2(out[(0xF, 0xF)][name], orc.read(name)) <= 1e-5, name


@pytest.mark.parametrize("shape,domain", [((64, 64, 64), abi.FX_DOMAIN_SPARSE), ((60, 52, 44), abi.FX_DOMAIN_SPARSE),
                                          ((40, 40, 40), abi.FX_DOMAIN_DENSE)])
def test_vcycle_sweep_tiles_bit_identical(shape, domain):
    """The split sweeps on 4^3 octants and on 16^3 brick groups (plan bits 20 + 2l of the
    up mask), one or two sweeps per launch, give the same bits as the 8^3-brick sweeps and
    the fused legs, on a cube, on a ragged domain (partial bricks, odd brick counts per
    axis, so 2x2x2 groups are cut at the faces) and densely."""
    W, H, D = shape
    n = min(shape)
    srcs = scenes.c2_sources(n, radius_vox=2.5)
    desc = scenes.c2_desc(n, domain_mode=domain)
    desc.width, desc.height, desc.depth = W, H, D
    st = {k: v for k, v in scenes.random_state((D, H, W), 21).items() if k in PERSISTENT}
    if domain == abi.FX_DOMAIN_SPARSE:
        # a sparse start: zero outside a blob so that levels hold partial brick sets
        zz, yy, xx = np.meshgrid(np.arange(D), np.arange(H), np.arange(W), indexing="ij")
        blob = ((xx - W * 0.4) ** 2 + (yy - H * 0.3) ** 2 + (zz - D * 0.5) ** 2) < (0.3 * n) ** 2
        st = {k: np.where(blob, v, 0.0) for k, v in st.items()}
    t4, t16 = sum(1 << (20 + 2 * l) for l in range(3)), sum(2 << (20 + 2 * l) for l in range(3))
    plans = [(0, 0), (0x7, 0x7), (0x7, 0x7 | t4), (0x707, 0x707 | t4), (0x7, 0x7 | t16), (0x707, 0x707 | t16),
             (0x107, 0x7 | (2 << 20) | (1 << 22))]
    out = {}
    for plan in plans:
        sim = CombustionVolumeSimulation(desc)
        sim.set_vcycle_plan(*plan)
        sim.updateSources(srcs)
        for k, v in st.items():
            sim.write_field(k, v)
        for k in range(3):
            sim.step(None, k / 60.0, 1 / 60.0)
        out[plan] = {f: sim.read_field(f) for f in PERSISTENT + ("pressure",)}
        sim.dispose()
    assert float(np.abs(out[(0, 0)]["pressure"]).max()) > 0
    for plan, fields in out.items():
        for name, arr in fields.items():
            np.testing.assert_array_equal(arr, out[(0, 0)][name], err_msg=f"plan {plan}: {name}")


def test_autotune_keeps_results():
    """The autotuner only picks kernels: a run that autotunes mid-way equals a
    run that never does, bit for bit."""
    n = 64
    srcs = scenes.c2_sources(n, radius_vox=2.5)
    st, _ = scenes.c2_plume_state(n, srcs, target=0.15)
    outs = []
    for tune in (False, True):
        sim = CombustionVolumeSimulation(scenes.c2_desc(n))
        sim.updateSources(srcs)
        for k, v in st.items():
            sim.write_field(k, v)
        for k in range(6):
            sim.step(None, k / 60.0, 1 / 60.0)
            if tune and k == 2:
                plan = sim.autotune_vcycle(2)
                assert sim.vcycle_plan() == plan
        outs.append({f: sim.read_field(f) for f in PERSISTENT + ("pressure",)})
        sim.dispose()
    for name in outs[0]:
        np.testing.assert_array_equal(outs[1][name], outs[0][name], err_msg=name)


def _raymarch_scene(n=32):
    """A C1 scene after 31 oracle steps, written into the simulation (flags checked)."""
    desc = scenes.c1_desc(n)
    sim, orc = make_pair(desc, scenes.c1_sources())
    for k in range(30):
        orc.step(k / 60.0, 1 / 60.0)
    for name in PERSISTENT:
        sim.write_field(name, orc.read(name))
    sim.step(None, 30 / 60.0, 1 / 60.0)
    orc.step(30 / 60.0, 1 / 60.0)
    for name in PERSISTENT:
        sim.write_field(name, orc.read(name))
    np.testing.assert_array_equal(sim.read_flags(), orc.flags())
    return sim, orc


def _check_images(sim, orc, rp, modes=("temperature", "density", "fuel")):
    cam = orbit_camera_snapshot()
    for mode in modes:
        img = rp.render(sim, cam, mode, 64, 64).astype(np.float32)
        ref = orc.raymarch(camera_struct(cam), rp.params_struct(mode), 64, 64).astype(np.float16).astype(np.float32)
        assert ref[..., 3].max() > 0.05, "scene renders nothing"
        diff = np.abs(img - ref)
        assert rel_l2(img, ref) < 2e-3, f"{mode}: image rel L2 {rel_l2(img, ref):.3e}"
        assert np.mean(diff.max(axis=-1) < 2e-3) > 0.99


def test_raymarch_parity():
    n = 32
    sim, orc = _raymarch_scene(n)
    rp = VolumeRaymarchPass((n, n, n))
    rp.setRenderParams(stepCount=64)
    _check_images(sim, orc, rp)


def test_raymarch_parity_point_and_several_lights():
    """Point lights (direction to the light and distance attenuation per sample,
    RP:503-520) and lights 1 and 2 (evaluated per sample, not per ray), which the
    default single directional light never reaches."""
    n = 32
    sim, orc = _raymarch_scene(n)
    rp = VolumeRaymarchPass((n, n, n))
    lights = [RenderLight("point", position=(5.0, 20.0, 5.0), color=(1.0, 0.8, 0.6), intensity=4.0),
              RenderLight("directional", direction=(0.2, -1.0, 0.3), color=(0.6, 0.7, 1.0), intensity=1.5),
              RenderLight("point", position=(-12.0, 8.0, 14.0), color=(0.5, 1.0, 0.7), intensity=3.0)]
    cam = orbit_camera_snapshot()
    rp.setRenderParams(stepCount=64)
    default_img = rp.render(sim, cam, "temperature", 64, 64).astype(np.float32)
    rp.setRenderParams(lights=lights)
    _check_images(sim, orc, rp)
    lit = rp.render(sim, cam, "temperature", 64, 64).astype(np.float32)
    assert rel_l2(lit, default_img) > 2e-2, "the lights do not change the image"
    # a point light as light 0 (its terms are per sample too) and no light at all
    rp.setRenderParams(lights=[RenderLight("point", position=(0.0, 30.0, -6.0), intensity=5.0)])
    _check_images(sim, orc, rp, modes=("temperature",))
    rp.setRenderParams(lights=[])
    _check_images(sim, orc, rp, modes=("temperature",))
    # non-default Henyey-Greenstein asymmetries for the forward and back lobes
    rp.setRenderParams(lights=lights, scatteringForward=0.7, scatteringBack=-0.5)
    _check_images(sim, orc, rp, modes=("temperature",))


def test_snapshot_skipped_while_not_rendered():
    """A step packs the render snapshot only when the previous one was rendered (a simulation
    that is not being rendered skips it); a raymarch after unrendered steps repacks every brick
    first. A scene rendered only after its fourth step gives the same image and fields as the
    same scene rendered after every step."""
    n, w, h = 48, 96, 80
    desc = scenes.c2_desc(n)
    srcs = scenes.c2_sources(n, radius_vox=4.0)
    st, _ = scenes.c2_plume_state(n, srcs, target=0.15)
    sims = []
    for _ in range(2):
        sim = CombustionVolumeSimulation(desc)
        sim.updateSources(srcs)
        for k, v in st.items():
            sim.write_field(k, v)
        sims.append(sim)
    rp = VolumeRaymarchPass((n, n, n))
    rp.setRenderParams(stepCount=64)
    cam = orbit_camera_snapshot()
    for k in range(4):
        for j, sim in enumerate(sims):
            sim.step(None, k / 60.0, 1 / 60.0)
            if j == 0:
                rp.render(sim, cam, "temperature", w, h, k / 60.0)
    a = rp.render(sims[0], cam, "temperature", w, h, 3 / 60.0)
    b = rp.render(sims[1], cam, "temperature", w, h, 3 / 60.0)
    np.testing.assert_array_equal(a, b)
    assert float(a[..., 3].astype(np.float32).max()) > 0.05
    for name in ("density", "temperature", "velocity_y", "pressure"):
        np.testing.assert_array_equal(sims[0].read_field(name), sims[1].read_field(name), err_msg=name)
    for sim in sims:
        sim.dispose()


def test_render_overlap_matches_serial():
    """The raymarch runs on its own stream, overlapped with the second half of
    the step; images equal those rendered after a full synchronisation, and a
    following step never overwrites what a pending render reads."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n, w, h = 48, 96, 80
    desc = scenes.c2_desc(n)
    srcs = scenes.c2_sources(n, radius_vox=4.0)
    st, _ = scenes.c2_plume_state(n, srcs, target=0.15)
    sims, bufs = [], []
    for _ in range(2):
        sim = CombustionVolumeSimulation(desc)
        sim.updateSources(srcs)
        for k, v in st.items():
            sim.write_field(k, v)
        ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(ptr), C.c_size_t(w * h * 8)) == 0
        sims.append(sim)
        bufs.append(ptr)
    rp = VolumeRaymarchPass((n, n, n))
    rp.setRenderParams(stepCount=64)
    cam = orbit_camera_snapshot()
    imgs = [[], []]
    for k in range(6):
        for j, sim in enumerate(sims):
            sim.step(None, k / 60.0, 1 / 60.0)
            if j == 1:
                sim.synchronize()
            rp.render_device(sim, cam, bufs[j].value, "temperature", w, h, k / 60.0)
        for j, sim in enumerate(sims):
            sim.synchronize()
            out = np.empty((h, w, 4), np.float16)
            assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), bufs[j], C.c_size_t(out.nbytes), 2) == 0
            imgs[j].append(out)
    for k in range(6):
        np.testing.assert_array_equal(imgs[0][k], imgs[1][k], err_msg=f"frame {k}")
    assert float(imgs[0][-1][..., 3].astype(np.float32).max()) > 0.05
    for ptr in bufs:
        hip.hipFree(ptr)


def test_batch_matches_single_scenes():
    """Scenes stepped concurrently through fx_batch equal the same scenes stepped
    alone, bit for bit (independent streams, no shared state)."""
    specs = scenes.c3_scenes(count=4, seed=11)
    descs = [scenes.c2_desc(sp["n"]) for sp in specs]
    states = [scenes.plume_state_fast(sp["n"], sp["sources"], sp["target"], sp["seed"]) for sp in specs]
    batch = CombustionVolumeBatch(descs)
    for sim, sp, st in zip(batch.scenes, specs, states):
        sim.updateSources(sp["sources"])
        for k, v in st.items():
            sim.write_field(k, v)
    for k in range(5):
        batch.step(k / 60.0, 1 / 60.0)
    batch.synchronize()
    for j, (sp, st) in enumerate(zip(specs, states)):
        sim = CombustionVolumeSimulation(descs[j])
        sim.updateSources(sp["sources"])
        for k, v in st.items():
            sim.write_field(k, v)
        for k in range(5):
            sim.step(None, k / 60.0, 1 / 60.0)
        for name in PERSISTENT + ("pressure",):
            np.testing.assert_array_equal(batch.scenes[j].read_field(name), sim.read_field(name),
                                          err_msg=f"scene {j} ({sp['n']}^3): {name}")
        sim.dispose()
    batch.dispose()


def test_deterministic_repeat():
    n = 32
    desc = scenes.c2_desc(n)
    srcs = scenes.c2_sources(n, radius_vox=3.0)
    outs = []
    for _ in range(2):
        sim = CombustionVolumeSimulation(desc)
        sim.updateSources(srcs)
        for k in range(10):
            sim.step(None, k / 60.0, 1 / 60.0)
        outs.append([sim.read_field(f) for f in PERSISTENT])
        sim.dispose()
    for a, b in zip(*outs):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("sparse", [False, True])
def test_graph_replay_matches_eager(sparse, monkeypatch):
    """The captured step graph, replayed many times with eager readbacks in
    between, gives bit-identical results to the eager launch sequence."""
    n = 32
    if sparse:
        desc, srcs = scenes.c2_desc(n), scenes.c2_sources(n, radius_vox=3.0)
    else:
        desc, srcs = scenes.c1_desc(n, quality={"pressureInterval": 2, "vorticityInterval": 3}), scenes.c1_sources()
    sims = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("FENIX_GRAPH", mode)
        sims[mode] = CombustionVolumeSimulation(desc)
        sims[mode].updateSources(srcs)
    for k in range(14):
        for sim in sims.values():
            sim.step(None, k / 60.0, 1 / 60.0)
        for name in PERSISTENT + ("pressure",):
            np.testing.assert_array_equal(sims["1"].read_field(name), sims["0"].read_field(name), err_msg=f"{name} step {k}")
        np.testing.assert_array_equal(sims["1"].read_flags(), sims["0"].read_flags())


def test_velocity_written_through_device_pointer_reaches_the_advection():
    """The velocity advection gathers a float4 mirror of u. A velocity written
    through fx_field_device_ptr (a borrowed pointer) must be seen by the next
    step exactly like one written with fx_write_field. Dense domain: in a sparse
    one fx_write_field also makes the written bricks activation candidates, which
    a write through the pointer cannot."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = 40
    desc = scenes.c2_desc(n, domain_mode=abi.FX_DOMAIN_DENSE)
    srcs = scenes.c2_sources(n, radius_vox=3.0)
    st, _ = scenes.c2_plume_state(n, srcs, target=0.2)
    other = {k: v for k, v in scenes.random_state((n, n, n), 11).items() if k.startswith("velocity_")}
    sims = []
    for _ in range(3):
        sim = CombustionVolumeSimulation(desc)
        sim.updateSources(srcs)
        for k, v in st.items():
            sim.write_field(k, v)
        sims.append(sim)
    a, b, c = sims
    for k, v in other.items():
        c.write_field(k, v)
    for k in range(2):
        a.step(None, k / 60.0, 1 / 60.0)
        b.step(None, k / 60.0, 1 / 60.0)
    for s in sims:
        s.synchronize()
    nbytes = ((n + 7) // 8) ** 3 * 512 * 4
    for comp in ("velocity_x", "velocity_y", "velocity_z"):
        fid = abi.FIELDS[comp]
        src = c._lib.fx_field_device_ptr(c._h, fid)
        dst = b._lib.fx_field_device_ptr(b._h, fid)
        assert src and dst
        assert hip.hipMemcpy(C.c_void_p(dst), C.c_void_p(src), C.c_size_t(nbytes), 3) == 0  # device to device
        a.write_field(comp, c.read_field(comp))
    assert hip.hipDeviceSynchronize() == 0
    for k in range(2, 4):
        a.step(None, k / 60.0, 1 / 60.0)
        b.step(None, k / 60.0, 1 / 60.0)
    for name in PERSISTENT:
        np.testing.assert_array_equal(b.read_field(name), a.read_field(name), err_msg=name)
    assert float(np.abs(a.read_field("velocity_y")).max()) > 0.1
    for s in sims:
        s.dispose()


@pytest.mark.gpu
def test_scalars_written_through_device_pointer_reach_the_advection():
    """The scalar advection gathers a float4 mirror of density, temperature, fuel
    and reaction that the injection and combustion passes keep. Scalars written
    through fx_field_device_ptr (a borrowed pointer) between steps must reach the
    next step exactly like ones written with fx_write_field. Dense domain, as in
    the velocity case above."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = 40
    desc = scenes.c2_desc(n, domain_mode=abi.FX_DOMAIN_DENSE)
    srcs = scenes.c2_sources(n, radius_vox=3.0)
    st, _ = scenes.c2_plume_state(n, srcs, target=0.2)
    other = {k: np.clip(v, 0.0, 1.0) for k, v in scenes.random_state((n, n, n), 12).items()
             if k in ("density", "temperature", "fuel", "reaction")}
    sims = []
    for _ in range(3):
        sim = CombustionVolumeSimulation(desc)
        sim.updateSources(srcs)
        for k, v in st.items():
            sim.write_field(k, v)
        sims.append(sim)
    a, b, c = sims
    for k, v in other.items():
        c.write_field(k, v)
    for k in range(2):
        a.step(None, k / 60.0, 1 / 60.0)
        b.step(None, k / 60.0, 1 / 60.0)
    for s in sims:
        s.synchronize()
    nbytes = ((n + 7) // 8) ** 3 * 512 * 4
    for name in ("density", "temperature", "fuel", "reaction"):
        fid = abi.FIELDS[name]
        src = c._lib.fx_field_device_ptr(c._h, fid)
        dst = b._lib.fx_field_device_ptr(b._h, fid)
        assert src and dst
        assert hip.hipMemcpy(C.c_void_p(dst), C.c_void_p(src), C.c_size_t(nbytes), 3) == 0  # device to device
        a.write_field(name, c.read_field(name))
    assert hip.hipDeviceSynchronize() == 0
    for k in range(2, 4):
        a.step(None, k / 60.0, 1 / 60.0)
        b.step(None, k / 60.0, 1 / 60.0)
    for name in PERSISTENT:
        np.testing.assert_array_equal(b.read_field(name), a.read_field(name), err_msg=name)
    assert float(np.abs(a.read_field("density")).max()) > 0.1
    for s in sims:
        s.dispose()


def test_coarsest_level_beyond_the_lds_cap():
    """Sparse 208^3 plume with ~70% of the bricks active: the coarsest V-cycle level (26^3) holds
    more active voxels than the one-workgroup LDS smoother (8192). The first step, before the
    host has seen an active count, runs that smoother's global fallback; later steps run the
    coarsest sweeps as multi-workgroup launches (the lagged count exceeds the cap). Both match the
    oracle per step, with flags, list and page table bit-exact."""
    n = 208
    srcs = scenes.c2_sources(n, count=4, radius_vox=12.0)
    st = scenes.plume_state_fast(n, srcs, target=0.7, seed=8)
    desc = scenes.c2_desc(n)
    sim, orc = make_pair(desc, srcs)
    load_state(sim, orc, st)
    dt = 1.0 / 60.0
    for k in range(4):
        if k > 0:
            for name in PERSISTENT:
                sim.write_field(name, orc.read(name))
        sim.step(None, k * dt, dt)
        orc.step(k * dt, dt)
        np.testing.assert_array_equal(sim.read_flags(), orc.flags(), err_msg=f"flags differ at step {k}")
        compare_all(sim, orc, f"step {k}", fields=list(PERSISTENT) + ["pressure", "divergence"])
    # a level-3 voxel J is active iff fine brick J is: the coarsest interior holds more than the cap
    f3 = orc.flags().reshape(26, 26, 26).astype(bool)
    assert f3[1:-1, 1:-1, 1:-1].sum() > 8192
    lst = sim.read_active_list()
    np.testing.assert_array_equal(lst, np.flatnonzero(orc.flags()).astype(np.uint32))


def test_coarsest_sweeps_per_launch_bit_identical():
    """Beyond the one-workgroup smoother the coarsest level's 16 sweeps per cycle run as 16, 8 or 4
    multi-workgroup launches (plan bits 24-25: LDS tiles of halo 1, 2 or 4); the fields are
    bit-identical after steps on the 70%-active 208^3 plume, and the autotuner keeps them so."""
    n = 208
    srcs = scenes.c2_sources(n, count=4, radius_vox=12.0)
    st = scenes.plume_state_fast(n, srcs, target=0.7, seed=8)
    out = []
    for opt in (0, 1, 2):
        sim = CombustionVolumeSimulation(scenes.c2_desc(n))
        sim.updateSources(srcs)
        for k, v in st.items():
            sim.write_field(k, v)
        sim.set_vcycle_plan(0x3 | (opt << 24), 0x7)
        for k in range(3):
            sim.step(None, k / 60.0, 1 / 60.0)
        assert sim.vcycle_plan()[0] >> 24 == opt
        out.append({f: sim.read_field(f) for f in ("pressure", "velocity_x", "velocity_y", "density")})
        if opt == 2:
            down, _ = sim.autotune_vcycle(2)
            assert (down >> 24) in (0, 1, 2)
            sim.step(None, 3 / 60.0, 1 / 60.0)
        sim.dispose()
    for opt in (1, 2):
        for f in out[0]:
            np.testing.assert_array_equal(out[opt][f], out[0][f], err_msg=f"plan option {opt}: {f}")
