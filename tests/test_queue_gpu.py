"""Brick-queue batches (fx_create_scenes) on the GPU: scenes of different sizes, emitters and
runtime parameters share every level's storage, and each pass of a step is one launch over the
active bricks of all scenes. Each scene must come out bit for bit as the same scene simulated
alone (fields, brick flags, active list and page table), in FP32 and with stochastically
rounded BF16 and UF16 storage, through several steps, a V-cycle plan change and a reset; per-scene
state written through the batch joins the next activation like a single scene's."""
import ctypes as C

import numpy as np
import pytest

import scenes
from fenix_amd import abi
from fenix_amd.render import VolumeRaymarchPass, orbit_camera_snapshot
from fenix_amd.simulation import CombustionVolumeQueueBatch, CombustionVolumeSimulation

pytestmark = pytest.mark.gpu

PERSISTENT = ("density", "temperature", "fuel", "reaction", "velocity_x", "velocity_y", "velocity_z")
CHECKED = PERSISTENT + ("pressure", "divergence")
DT = 1.0 / 60.0
# (width, height, depth, plume target, emitters): uneven sizes, some not multiples of 8
SHAPES = [(48, 64, 40, 0.25, 2), (32, 32, 32, 0.3, 1), (72, 56, 64, 0.15, 3), (40, 88, 48, 0.2, 2), (64, 64, 64, 0.1, 4)]


def make_scenes(precision):
    out = []
    for k, (w, h, d, target, ne) in enumerate(SHAPES):
        n = max(w, h, d)
        srcs = scenes.c2_sources(n, seed=40 + k, count=ne, radius_vox=max(2.5, n / 20.0))
        st = scenes.plume_state_fast(n, srcs, target, seed=40 + k)
        st = {f: v[:d, :h, :w].copy() for f, v in st.items()}  # the plume cropped to the scene's box
        desc = scenes.c2_desc(n, precision=precision)
        desc.width, desc.height, desc.depth = w, h, d
        desc.runtime.buoyancy = 3.6 + 0.3 * k  # per-scene runtime parameters
        out.append((desc, srcs, st))
    return out


def assert_same(batch_scene, sim, label):
    for name in CHECKED:
        np.testing.assert_array_equal(batch_scene.read_field(name), sim.read_field(name), err_msg=f"{label}: {name}")
    for which in (abi.FX_FLAGS_RENDER, abi.FX_FLAGS_INJECTED):
        np.testing.assert_array_equal(batch_scene.read_flags(which), sim.read_flags(which), err_msg=f"{label}: flags")
    np.testing.assert_array_equal(batch_scene.read_active_list(), sim.read_active_list(), err_msg=f"{label}: list")
    np.testing.assert_array_equal(batch_scene.read_page_table(), sim.read_page_table(), err_msg=f"{label}: page table")
    assert batch_scene.active_brick_count() == sim.active_brick_count()


@pytest.mark.parametrize("precision", [abi.FX_PRECISION_FP32, abi.FX_PRECISION_BF16, abi.FX_PRECISION_UF16])
def test_queue_batch_equals_scenes_alone(precision):
    spec = make_scenes(precision)
    batch = CombustionVolumeQueueBatch([d for d, _, _ in spec])
    alone = [CombustionVolumeSimulation(d) for d, _, _ in spec]
    for sc, sim, (d, srcs, st) in zip(batch.scenes, alone, spec):
        sc.updateSources(srcs)
        sim.updateSources(srcs)
        for f, v in st.items():
            sc.write_field(f, v)
            sim.write_field(f, v)
        assert sc.getRuntimeParams()["buoyancy"] == pytest.approx(d.runtime.buoyancy)
    total = 0
    for k in range(6):
        if k == 3:  # a plan change mid-run (results never depend on the plan): the batch's split
            # sweeps on 16^3 brick groups two per launch (level 0 down) and 4^3 octants (level 1)
            batch.set_vcycle_plan(0x103, 0x7 | (2 << 20) | (1 << 22))
            for sim in alone:
                sim.set_vcycle_plan(0x3, 0x7)
        batch.step(k * DT, DT)
        for sim in alone:
            sim.step(None, k * DT, DT)
        batch.synchronize()
        for j, (sc, sim) in enumerate(zip(batch.scenes, alone)):
            assert_same(sc, sim, f"scene {j} step {k}")
        total = batch.active_brick_count()
    assert total == sum(sim.active_brick_count() for sim in alone)
    assert min(sim.active_brick_count() for sim in alone) > 0
    print(f"precision {precision}: {len(spec)} scenes, {total} active bricks after 6 steps", flush=True)
    # a reset clears every scene; stepping again from rewritten state still matches
    batch.reset()
    for sim in alone:
        sim.reset()
    for sc, sim, (_, _, st) in zip(batch.scenes, alone, spec):
        for f in ("density", "temperature"):
            sc.write_field(f, st[f])
            sim.write_field(f, st[f])
    batch.step(0.0, DT)
    for sim in alone:
        sim.step(None, 0.0, DT)
    for j, (sc, sim) in enumerate(zip(batch.scenes, alone)):
        assert_same(sc, sim, f"scene {j} after reset")
    batch.dispose()
    for sim in alone:
        sim.dispose()


def test_queue_batch_long_steps_equal_scenes_alone():
    """Steps of dt = 0.5 and 1 carry the plumes' backtraces past the advection kernels' LDS tiles
    (the brick +- 1 voxel), so their samples gather from each scene's fields in memory through the
    batch's per-scene grids: still bit for bit the scenes alone."""
    spec = make_scenes(abi.FX_PRECISION_FP32)[:3]
    batch = CombustionVolumeQueueBatch([d for d, _, _ in spec])
    alone = [CombustionVolumeSimulation(d) for d, _, _ in spec]
    vmax = 0.0
    for sc, sim, (d, srcs, st) in zip(batch.scenes, alone, spec):
        sc.updateSources(srcs)
        sim.updateSources(srcs)
        for f, v in st.items():
            sc.write_field(f, v)
            sim.write_field(f, v)
        vmax = max(vmax, max(float(np.abs(st[f"velocity_{a}"]).max()) for a in "xyz"))
    assert vmax * 0.5 >= 1.0, vmax  # the longest backtraces leave the tile
    for k, dt in enumerate((0.5, 1.0)):
        batch.step(k * 1.0, dt)
        for sim in alone:
            sim.step(None, k * 1.0, dt)
        batch.synchronize()
        for j, (sc, sim) in enumerate(zip(batch.scenes, alone)):
            assert_same(sc, sim, f"scene {j} step {k} dt {dt}")
    batch.dispose()
    for sim in alone:
        sim.dispose()


def test_queue_batch_with_an_empty_scene_and_a_batch_of_one():
    """Edge cases of the queue: a scene with no emitters and no state (no active bricks at any
    level, an empty run in every list) beside a busy one, and a batch of a single scene; both
    equal the scenes simulated alone."""
    (d0, s0, st0), (d1, _, _) = make_scenes(abi.FX_PRECISION_FP32)[:2]
    for descs, busy in (([d0, d1], [True, False]), ([d0], [True])):
        batch = CombustionVolumeQueueBatch(descs)
        alone = [CombustionVolumeSimulation(d) for d in descs]
        for sc, sim, b in zip(batch.scenes, alone, busy):
            if b:
                sc.updateSources(s0)
                sim.updateSources(s0)
                for f, v in st0.items():
                    sc.write_field(f, v)
                    sim.write_field(f, v)
        for k in range(3):
            batch.step(k * DT, DT)
            for sim in alone:
                sim.step(None, k * DT, DT)
        for j, (sc, sim) in enumerate(zip(batch.scenes, alone)):
            assert_same(sc, sim, f"{len(descs)} scenes, scene {j}")
        assert batch.scenes[0].active_brick_count() > 0
        if len(descs) == 2:
            assert batch.scenes[1].active_brick_count() == 0
            assert not batch.scenes[1].read_field("density").any()
        batch.dispose()
        for sim in alone:
            sim.dispose()


def test_queue_batch_scene_images_equal_scenes_alone():
    """Each scene of a batch renders from the batch's shared snapshot with its own grid, world
    bounds and brick flags (fx_scene_raymarch): the images equal those of the scenes alone, bit
    for bit, with every scene rendered after each step and with renders skipped for two steps."""
    spec = make_scenes(abi.FX_PRECISION_FP32)[:3]
    batch = CombustionVolumeQueueBatch([d for d, _, _ in spec])
    alone = [CombustionVolumeSimulation(d) for d, _, _ in spec]
    passes = []
    for sc, sim, (d, srcs, st) in zip(batch.scenes, alone, spec):
        sc.updateSources(srcs)
        sim.updateSources(srcs)
        for f, v in st.items():
            sc.write_field(f, v)
            sim.write_field(f, v)
        rp = VolumeRaymarchPass((d.width, d.height, d.depth))
        rp.setRenderParams(stepCount=64)
        passes.append(rp)
    cam = orbit_camera_snapshot()
    lit = 0.0
    for k in range(5):
        batch.step(k * DT, DT)
        for sim in alone:
            sim.step(None, k * DT, DT)
        if k in (1, 2):  # nobody renders these steps
            continue
        for j, (sc, sim, rp) in enumerate(zip(batch.scenes, alone, passes)):
            a = rp.render(sc, cam, "temperature", 80, 64, k * DT)
            b = rp.render(sim, cam, "temperature", 80, 64, k * DT)
            np.testing.assert_array_equal(a, b, err_msg=f"scene {j} step {k}")
            lit = max(lit, float(a[..., 3].astype(np.float32).max()))
    assert lit > 0.05
    batch.dispose()
    for sim in alone:
        sim.dispose()


def test_queue_batch_rejects_mixed_settings_and_single_scene_calls():
    a = scenes.c2_desc(32)
    b = scenes.c2_desc(40, precision=abi.FX_PRECISION_BF16)
    with pytest.raises(abi.FenixError, match="share precision"):
        CombustionVolumeQueueBatch([a, b])
    dense = scenes.c1_desc(32)
    with pytest.raises(abi.FenixError, match="sparse domain"):
        CombustionVolumeQueueBatch([dense])
    batch = CombustionVolumeQueueBatch([a, scenes.c2_desc(40)])
    sim_view = CombustionVolumeSimulation(a, batch._lib, borrowed_handle=batch.handle.value)
    with pytest.raises(abi.FenixError, match="per scene"):
        sim_view.read_field("density")
    with pytest.raises(abi.FenixError, match="out of range"):
        batch._check(batch._lib.fx_scene_active_brick_count(batch.handle, 2, C.byref(C.c_uint32())))
    batch.dispose()
