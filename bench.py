"""Benchmark: active voxel-updates/s per full step of the combustion-volume
pipeline on one MI355X.

Default workload (sparse 256^3): FP32, 8^3 bricks, threshold activation
(density > 1e-4 or |u| > 1e-3, one-brick dilation), two seeded floor emitters
of radius 10 voxels, burn rate 0.5, ignition T = 0.6, 4-level V-cycle (2/2
sweeps, omega 2/3, 4 cycles), vorticity eps = 0.3, and a 1024x1024 raymarch with
128 samples per step. The scene starts from a seeded, developed rising plume
holding ~15% of the bricks active (`c2-cold` starts from an empty volume).
`--warmup` untimed steps (default 100) run first, then `--steps` steps
(default 100) are timed: a "step" is fx_step + the raymarch, with every input
already resident in HBM.

After the timed window the same scene is stepped `--steps` more times without its
raymarch (`sim_only`), and the default run then times the batched + mixed-precision
path, the C3 batch of 64 scenes in one brick-queue simulation with UF16 storage
(`batched_mixed`, with its ratio to this plain FP32 single-scene path with and
without the raymarch; `--no-batch` skips it).

Other workloads: `c1` (dense 64^3, reference semantics), `c3` and `--batch K`
(batched scenes against one at a time), `c4` (sparse 512^3 precision runs),
`c5` (loss + gradient of the differentiable step).

Prints one JSON line (see the keys below).
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from fenix_amd import abi  # noqa: E402
from fenix_amd.render import BloomPass, VolumeRaymarchPass, orbit_camera_snapshot  # noqa: E402
from fenix_amd.simulation import CombustionVolumeBatch, CombustionVolumeQueueBatch, CombustionVolumeSimulation  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E

# Algorithmic bytes per active voxel per launch (FP32 storage; velocity/vorticity
# counted as 3 x FP32; each stage reads its inputs once and writes its outputs once).
STAGE_BYTES = {
    "advect_velocity": 24.0,
    "advect_scalars": 44.0,
    "combustion": 32.0,
    "buoyancy": 36.0,
    "vorticity": 28.0,
    "confinement": 52.0,
    "divergence": 16.0,
    "projection": 28.0,
    "jacobi_fine": 12.0,
}
# Whole-step algorithmic bytes per active voxel (config V-cycle, FP32): simulation
# 602.2 + raymarch 16 B/voxel (+ 8 B per pixel, added per workload).
STEP_BYTES_SIM = 602.2
# The same for the reference's parity pressure schedule (plain C1): 351.6 B (318.6 with 16-bit d/T/f).
STEP_BYTES_SIM_REF = {0: 351.6, 2: 318.6}
RAY_BYTES_PER_VOXEL = 16.0
# Accesses of the density/temperature/fuel planes per active voxel, per stage: with
# 16-bit storage each one moves 2 bytes fewer than the FP32 figure above.
SCALAR_ACCESSES = {"advect_scalars": 6, "combustion": 6, "buoyancy": 2, "confinement": 2}
# Whole-step simulation bytes with 16-bit d/T/f (the stages above plus the two
# activation passes, which read d and T at 65/512 of the voxels).
STEP_BYTES_SIM_16 = 569.2
RAY_SCALAR_ACCESSES = 3
# HBM bytes per active voxel per launch from rocprofv3 counter runs of this
# benchmark: FETCH_SIZE x 2.0 (a 256 MiB device copy reads 128 MiB on that counter)
# + WRITE_SIZE x 1.0 (the copy's writes read exactly). Summaries with the commands
# are in profiles/.
TRAFFIC_PER_VOXEL = {"jacobi_fine": 12.61, "advect_scalars": 64.9}
PRECISIONS = {"fp32": (abi.FX_PRECISION_FP32, 4), "bf16": (abi.FX_PRECISION_BF16, 2), "fp16": (abi.FX_PRECISION_FP16, 2),
              "u16": (abi.FX_PRECISION_U16, 2), "uf16": (abi.FX_PRECISION_UF16, 2)}


class HipBuffer:
    """Device allocation through the HIP runtime (no torch dependency)."""

    def __init__(self, nbytes):
        self.hip = C.CDLL("libamdhip64.so")
        self.ptr = C.c_void_p()
        rc = self.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(nbytes))
        if rc != 0:
            raise RuntimeError(f"hipMalloc failed ({rc})")

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


def workload(name, precision="fp32"):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import scenes
    prec = PRECISIONS[precision][0]
    tag = precision.upper()
    if name in ("c2", "c2-cold"):
        n = 256
        srcs = scenes.c2_sources(n)
        wl = dict(n=n, desc=scenes.c2_desc(n, precision=prec), sources=srcs, render=(1024, 1024, 128), state=None,
                  label=f"sparse 256^3 {tag} storage, 8^3 bricks, 2 floor emitters r=10, threshold activation, "
                        "4-level V-cycle, 1024x1024x128 raymarch")
        if name == "c2":
            # The config's ~15% active set: start from the developed plume (seeded).
            wl["state"], frac = scenes.c2_plume_state(n, srcs, target=0.15)
            wl["label"] += f", seeded plume start ({frac:.1%} of bricks)"
        else:
            wl["label"] += ", empty start"
        return wl
    if name == "c1":
        n = 64
        return dict(n=n, desc=scenes.c1_desc(n, precision=prec), sources=scenes.c1_sources(), render=(256, 256, 64), state=None,
                    label=f"dense 64^3 {tag} storage, reference semantics, 256x256x64 raymarch")
    raise SystemExit(f"unknown workload {name}")


# The V-cycle plan the autotuner picks for the C3 batch (split legs on levels 0-2):
# runs that pin plans (profiled runs, where the autotuner's repeated graph captures
# are not run) use it for the batched leg.
C3_PLAN = (0x7, 0x7)


def batched_mixed(precision="uf16", warmup=10, steps=20, autotune=True):
    """The batched + mixed-precision path on the C3 batch: its 64 sparse scenes
    (edges 32-192, 5-40% active) in one brick-queue simulation with density,
    temperature and fuel stored in `precision`, V-cycle plan autotuned after the
    first step (or C3_PLAN pinned), simulation only. Returns the timed window's figures."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import scenes
    specs = scenes.c3_scenes(64)
    qb = CombustionVolumeQueueBatch([scenes.c2_desc(sp["n"], precision=PRECISIONS[precision][0]) for sp in specs])
    for sim, sp in zip(qb.scenes, specs):
        sim.updateSources(sp["sources"])
        for name, values in scenes.plume_state_fast(sp["n"], sp["sources"], sp["target"], sp["seed"]).items():
            sim.write_field(name, values)
    dt = 1.0 / 60.0
    if not autotune:
        qb.set_vcycle_plan(*C3_PLAN)
    qb.step(0.0, dt)
    qb.synchronize()
    plan = qb.autotune_vcycle(4) if autotune else qb.vcycle_plan()
    for k in range(1, warmup):
        qb.step(k * dt, dt)
    qb.synchronize()
    a0 = qb.active_brick_count()
    t0 = time.perf_counter()
    for k in range(warmup, warmup + steps):
        qb.step(k * dt, dt)
    qb.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    active = 0.5 * (a0 + qb.active_brick_count())
    qb.dispose()
    return {"workload": "C3 batch of 64 sparse scenes, edge 32-192, 5-40% target active, brick queue "
                        "(one launch per pass over every scene's bricks), simulation only",
            "storage": precision, "steps": steps, "warmup": warmup, "ms_per_batch_step": ms,
            "active_bricks": active, "value": active * 512 / (ms * 1e-3),
            "vcycle_plan": {"down_split": plan[0], "up_split": plan[1], "autotuned": autotune}}


def cpu_baseline(wl, sim, elapsed0, steps):
    """Times the FP32 oracle (a C++ port, OpenMP) on a bounded sample of the same
    workload, started from the GPU's state."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import scenes
    from oracle import Oracle
    from fenix_amd.render import camera_struct

    orc = Oracle(wl["desc"], 32)
    orc.update_sources(*scenes.packed(wl["sources"]))
    for name in ("density", "temperature", "fuel", "reaction", "velocity_x", "velocity_y", "velocity_z"):
        orc.write(name, sim.read_field(name))
    w, h, sc = wl["render"]
    rp = VolumeRaymarchPass((wl["n"],) * 3)
    rp.setRenderParams(stepCount=sc)
    cam = camera_struct(orbit_camera_snapshot())
    params = rp.params_struct("temperature")
    dt = 1.0 / 60.0
    active = 0
    t0 = time.perf_counter()
    for k in range(steps):
        orc.step(elapsed0 + k * dt, dt)
        active += orc.active_count()
        orc.raymarch(cam, params, w, h)
    t1 = time.perf_counter()
    per_step = (t1 - t0) / steps
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": (active / steps) * 512 / per_step, "unit": "active voxel-updates/s", "cores": cores, "kind": "port",
            "sample": f"{steps} full steps (sim + {w}x{h}x{sc} raymarch) of the same workload from the GPU state "
                      f"after the timed window; {per_step:.2f} s/step"}


def run_batch(args):
    """Batched scenes: aggregate active voxel-updates/s of one batch step (every scene steps
    once; with --workload c2 every scene also renders) as one brick-queue batch
    (fx_create_scenes, the default) or as independent scenes on their own streams (fx_batch_*),
    against the plain path (the same scenes stepped one at a time, each alone on the GPU). A
    queue run also reports the streams batch."""
    # Each scene runs on two streams; give them hardware queues of their own
    # (read by the HIP runtime at initialisation, before any device call).
    os.environ.setdefault("GPU_MAX_HW_QUEUES", str(args.hw_queues))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import scenes
    prec = PRECISIONS[args.precision][0]
    render = None
    if args.workload == "c3":
        specs = scenes.c3_scenes(args.batch or 64)
        label = (f"C3 batch of {len(specs)} sparse scenes, edge 32-192 (log-uniform), 5-40% target active, "
                 f"1-4 floor emitters, {args.precision.upper()} storage, simulation only")
    elif args.workload == "c2":
        k = args.batch
        specs = [dict(n=256, target=0.15, seed=2 + i, sources=scenes.c2_sources(256, seed=2 + i)) for i in range(k)]
        render = (1024, 1024, 128)
        label = (f"{k} sparse 256^3 plume scenes (seeds 2..{1 + k}), {args.precision.upper()} storage, "
                 "each with a 1024x1024x128 raymarch")
    else:
        raise SystemExit("--batch applies to the c2 and c3 workloads")
    descs = [scenes.c2_desc(sp["n"], precision=prec) for sp in specs]
    states = []
    for sp in specs:
        if sp["n"] == 256:
            st, _ = scenes.c2_plume_state(256, sp["sources"], target=sp["target"], seed=sp["seed"])
        else:
            st = scenes.plume_state_fast(sp["n"], sp["sources"], sp["target"], sp["seed"])
        states.append(st)
    queue = args.batch_mode != "streams"
    dt = 1.0 / 60.0
    cam = orbit_camera_snapshot()
    imgs = [HipBuffer(render[0] * render[1] * 8) for _ in specs] if render else []

    def render_all(sims, k):
        # every scene's raymarch into its own image (after the batch step that produced it)
        if render:
            for j, sim in enumerate(sims):
                rp[j].render_device(sim, cam, imgs[j].ptr.value, "temperature", render[0], render[1], k * dt)

    rp = []
    if render:
        for sp in specs:
            rp.append(VolumeRaymarchPass((sp["n"],) * 3))
            rp[-1].setRenderParams(stepCount=render[2])

    def load(sims):
        for sim, sp, st in zip(sims, specs, states):
            sim.updateSources(sp["sources"])
            for name, values in st.items():
                sim.write_field(name, values)

    def timed(step_all, sync, count):
        # warm-up, then the timed window; returns (ms per batch step, mean active bricks)
        for k in range(1, args.warmup):
            step_all(k)
        sync()
        a0 = count()
        t0 = time.perf_counter()
        for k in range(args.warmup, args.warmup + args.steps):
            step_all(k)
        sync()
        t1 = time.perf_counter()
        return (t1 - t0) * 1e3 / args.steps, 0.5 * (a0 + count())

    queue_plan = None
    pinned = tuple(int(v, 0) for v in args.vcycle_plan.split(",")) if args.vcycle_plan else None
    if pinned:  # a pinned plan replaces the autotuner (profiled runs)
        args.autotune = False
    if queue:
        qb = CombustionVolumeQueueBatch(descs)
        if pinned:
            qb.set_vcycle_plan(*pinned)
        load(qb.scenes)
        qb.step(0.0, dt)
        render_all(qb.scenes, 0)
        qb.synchronize()
        queue_plan = qb.autotune_vcycle(4) if args.autotune else qb.vcycle_plan()

        def queue_step(k):
            qb.step(k * dt, dt)
            render_all(qb.scenes, k)

        ms, active = timed(queue_step, qb.synchronize, qb.active_brick_count)
        qb.dispose()
    if queue and args.no_compare:
        out = {
            "metric": "aggregate active voxel-updates/s per batch step (brick queue)",
            "value": active * 512 / (ms * 1e-3), "unit": "active voxel-updates/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic",
            "config": {"workload": label, "scenes": len(specs), "storage": args.precision, "active_bricks": active,
                       "vcycle_plan": {"down_split": queue_plan[0], "up_split": queue_plan[1], "autotuned": args.autotune}},
        }
        for b in imgs:
            b.free()
        print(json.dumps(out), flush=True)
        return
    batch = CombustionVolumeBatch(descs)
    if pinned:
        for sim in batch.scenes:
            sim.set_vcycle_plan(*pinned)
    load(batch.scenes)
    del states

    def batch_step(k):
        batch.step(k * dt, dt)
        render_all(batch.scenes, k)

    def scene_step(j, k):
        sim = batch.scenes[j]
        sim.step(None, k * dt, dt)
        if render:
            rp[j].render_device(sim, cam, imgs[j].ptr.value, "temperature", render[0], render[1], k * dt)

    batch_step(0)
    batch.synchronize()
    if args.autotune:
        for sim in batch.scenes:
            sim.autotune_vcycle(4)
    streams_ms, streams_active = timed(batch_step, batch.synchronize,
                                       lambda: sum(sim.active_brick_count() for sim in batch.scenes))
    if not queue:
        ms, active = streams_ms, streams_active
    voxels = active * 512
    # plain: each scene alone for the same number of steps, one after the other
    base = args.warmup + args.steps
    plain_s = 0.0
    for j in range(len(batch.scenes)):
        t0 = time.perf_counter()
        for k in range(base, base + args.steps):
            scene_step(j, k)
        batch.scenes[j].synchronize()
        plain_s += time.perf_counter() - t0
    plain_ms = plain_s * 1e3 / args.steps
    out = {
        "metric": "aggregate active voxel-updates/s per batch step (batched vs plain single-scene path)",
        "value": voxels / (ms * 1e-3),
        "unit": "active voxel-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": label, "scenes": len(specs), "storage": args.precision,
                   "batch": ("brick queue: one launch per pass over the bricks of all scenes (fx_create_scenes)"
                             if queue else "independent scenes on their own streams (fx_batch_*)"),
                   "edges": sorted(sp["n"] for sp in specs), "active_bricks": active,
                   "voxels": sum(sp["n"] ** 3 for sp in specs)},
        "plain": {"ms_per_batch_step": plain_ms, "value": voxels / (plain_ms * 1e-3)},
        "vs_plain": plain_ms / ms,
    }
    if queue:
        out["config"]["vcycle_plan"] = {"down_split": queue_plan[0], "up_split": queue_plan[1], "autotuned": args.autotune}
        out["streams"] = {"ms_per_batch_step": streams_ms, "value": streams_active * 512 / (streams_ms * 1e-3)}
    for b in imgs:
        b.free()
    batch.dispose()
    print(json.dumps(out), flush=True)


def run_c4(args):
    """C4 precision at full size: sparse 512^3, 4 floor emitters (seed 4) under a
    seeded developed plume of ~8% of the bricks, 100 steps (or --steps) in config
    mode, once per storage format of density, temperature and fuel (FP32, BF16, FP16,
    U16 linear fixed point and UF16 unsigned 4/12-bit float; FP32 arithmetic). Reports each run's step time and throughput, and the
    16-bit runs' relative L2 and max error against the FP32 run after the last
    step. The FP64 oracle cannot run 512^3, so accuracy against FP64 is pinned at
    64^3 by the precision tests, where FP32 stays within ~4e-6 of FP64."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import scenes
    n, nsteps = 512, max(1, args.steps)
    srcs = scenes.c2_sources(n, seed=4, count=4, radius_vox=20.0)
    print("c4: building the 512^3 start state", file=sys.stderr, flush=True)
    state = scenes.plume_state_fast(n, srcs, target=0.08, seed=4)
    fields = ("density", "temperature", "fuel", "reaction", "velocity_x", "velocity_y", "velocity_z")
    dt = 1.0 / 60.0
    runs, ref = {}, None
    for name in ("fp32", "bf16", "fp16", "u16", "uf16"):
        sim = CombustionVolumeSimulation(scenes.c2_desc(n, precision=PRECISIONS[name][0]))
        sim.updateSources(srcs)
        for k, v in state.items():
            sim.write_field(k, v)
        if args.vcycle_plan:
            sim.set_vcycle_plan(*(int(v, 0) for v in args.vcycle_plan.split(",")))
        sim.step(None, 0.0, dt)
        sim.synchronize()
        plan = sim.vcycle_plan() if args.vcycle_plan or not args.autotune else sim.autotune_vcycle(4)
        active0 = sim.active_brick_count()
        t0 = time.perf_counter()
        for k in range(1, nsteps + 1):
            sim.step(None, k * dt, dt)
        sim.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / nsteps
        active = 0.5 * (active0 + sim.active_brick_count())
        out = {f: sim.read_field(f) for f in fields}
        sim.dispose()
        run = {"ms_per_step": ms, "value": active * 512 / (ms * 1e-3), "active_bricks": active,
               "vcycle_plan": {"down_split": plan[0], "up_split": plan[1],
                               "autotuned": bool(args.autotune and not args.vcycle_plan)}}
        if ref is None:
            ref = out
        else:
            err = {}
            for f in fields:
                a, b = out[f].astype(np.float64), ref[f].astype(np.float64)
                nb = float(np.linalg.norm(b))
                err[f] = {"rel_l2": float(np.linalg.norm(a - b)) / nb if nb > 0 else 0.0, "max_abs": float(np.abs(a - b).max())}
            run["vs_fp32"] = err
        runs[name] = run
        print(f"c4 {name}: {ms:.3f} ms/step, {active:.0f} active bricks", file=sys.stderr, flush=True)
    sel = runs[args.precision]
    nb = ((n + 7) // 8) ** 3
    print(json.dumps({
        "metric": "active voxel-updates/s per step @512^3 sparse (C4 precision runs)",
        "value": sel["value"], "unit": "active voxel-updates/s", "n_gpus": 1, "steps": nsteps, "warmup": 1,
        "ms_per_step": sel["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "sparse 512^3, 8^3 bricks, 4 floor emitters r=20 (seed 4), threshold activation, "
                               "4-level V-cycle, seeded plume start (~8% of bricks), no render",
                   "storage": args.precision, "active_fraction": sel["active_bricks"] / nb},
        "runs": runs,
    }))


def run_c5(args):
    """Inflow-control gradients (fx_diff): dense 128^3 FP32, 64 steps, checkpoint 8,
    control 32^3 x 4 into the floor slab, target 3 Gaussian blobs (sigma 12, seed 5).
    A step is one loss-and-gradient evaluation (rollout with checkpoints, then
    per-segment recomputation and the adjoint pass); Adam (lr 1e-2) then runs
    --adam-iters iterations and the loss history is reported."""
    from fenix_amd.differentiable import SmokeRollout, diff_desc, gaussian_blobs, optimise_inflow, plume_control
    n, nsteps = 128, 64
    rollout = SmokeRollout(diff_desc(resolution=(n, n, n), steps=nsteps, checkpoint_every=8, control_n=32, slab=4))
    rng = np.random.default_rng(5)
    centres = [(rng.uniform(0.3, 0.7) * n, rng.uniform(0.3, 0.55) * n, rng.uniform(0.3, 0.7) * n) for _ in range(3)]
    rollout.set_target(gaussian_blobs((n, n, n), centres, 12.0, amplitude=0.5))
    ctrl = plume_control(32)
    warm, iters = max(1, min(args.warmup, 3)), max(1, min(args.steps, 20))
    for _ in range(warm):
        rollout.gradient(ctrl)
    t0 = time.perf_counter()
    for _ in range(iters):
        loss0, _ = rollout.gradient(ctrl)
    grad_ms = (time.perf_counter() - t0) * 1e3 / iters
    t0 = time.perf_counter()
    for _ in range(iters):
        rollout.forward(ctrl)
    fwd_ms = (time.perf_counter() - t0) * 1e3 / iters
    t0 = time.perf_counter()
    _, losses = optimise_inflow(rollout, ctrl, args.adam_iters, lr=1e-2)
    adam_s = time.perf_counter() - t0
    voxel_steps = float(n ** 3 * nsteps)
    out = {
        "metric": "voxel-steps/s of loss + gradient (checkpointed adjoint)",
        "value": voxel_steps / (grad_ms * 1e-3),
        "unit": "voxel-steps/s",
        "n_gpus": 1,
        "steps": iters,
        "warmup": warm,
        "ms_per_step": grad_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "inflow control: dense 128^3 FP32, 64 steps, checkpoint 8, control 32^3x4 into slab "
                               "y in [1,4], dt 1, buoyancy 0.02, confinement 0.02, 4-level V-cycle, target 3 Gaussian "
                               "blobs sigma 12 (seed 5)"},
        "forward_ms": fwd_ms,
        "gradient_over_forward": grad_ms / fwd_ms,
        "adam": {"lr": 1e-2, "iterations": args.adam_iters, "seconds": adam_s, "loss_first": losses[0] if losses else None,
                 "loss_last": losses[-1] if losses else None, "losses": losses},
    }
    rollout.dispose()
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", default="c2", choices=["c2", "c2-cold", "c1", "c3", "c4", "c5"])
    ap.add_argument("--adam-iters", type=int, default=16, help="c5: Adam iterations after the timed gradients")
    ap.add_argument("--timing-steps", type=int, default=10, help="extra steps timed per stage with HIP events")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="keep the V-cycle kernel plan (default or --vcycle-plan) instead of timing the candidates")
    ap.add_argument("--vcycle-plan", default=None, metavar="DOWN,UP",
                    help="per-level split masks of the V-cycle legs, e.g. 1,3 (implies --no-autotune); used to profile "
                         "the plan an autotuned run chose, since profiler overhead changes what the autotuner picks")
    ap.add_argument("--batch", type=int, default=0, metavar="K",
                    help="batched scenes: K 256^3 plume scenes (c2) or the C3 batch (c3, default 64 scenes)")
    ap.add_argument("--batch-mode", choices=["streams", "queue"], default="queue",
                    help="batches: one brick queue over all scenes (fx_create_scenes) or independent scenes on their "
                         "own streams (fx_batch_*); a queue run also times the streams batch ('streams'); both time "
                         "every scene alone ('plain')")
    ap.add_argument("--no-compare", action="store_true",
                    help="brick-queue batches: skip the streams batch and the plain path (a run of one simulation)")
    ap.add_argument("--hw-queues", type=int, default=4, choices=range(1, 33), metavar="Q",
                    help="hardware queues per process for batched runs (GPU_MAX_HW_QUEUES, at most 32)")
    ap.add_argument("--bloom", action="store_true",
                    help="a step also runs the bloom/composite passes into an RGBA8 canvas image (the reference's frame)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="also run one device-to-device copy of this many MiB (a known-size reference for the "
                         "HBM traffic counters)")
    ap.add_argument("--no-batch", action="store_true",
                    help="c2: skip the batched + mixed-precision comparison run (C3 brick queue, UF16 storage)")
    ap.add_argument("--precision", choices=sorted(PRECISIONS), default="fp32",
                    help="storage of density/temperature/fuel (arithmetic is FP32 throughout)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on exactly one GPU: there is no multi-GPU path (the sharding plan is design-only).",
              file=sys.stderr)
        sys.exit(2)
    if args.batch or args.workload == "c3":
        return run_batch(args)
    if args.workload == "c5":
        return run_c5(args)
    if args.workload == "c4":
        return run_c4(args)

    wl = workload(args.workload, args.precision)
    saved = 4 - PRECISIONS[args.precision][1]
    sim = CombustionVolumeSimulation(wl["desc"])
    sim.updateSources(wl["sources"])
    for name, values in (wl["state"] or {}).items():
        sim.write_field(name, values)
    wl["state"] = None
    w, h, sc = wl["render"]
    rp = VolumeRaymarchPass((wl["n"],) * 3)
    rp.setRenderParams(stepCount=sc)
    cam = orbit_camera_snapshot()
    img = HipBuffer(w * h * 8)
    canvas = HipBuffer(w * h * 4) if args.bloom else None
    bloom = BloomPass() if args.bloom else None
    dt = 1.0 / 60.0

    def full_step(k):
        sim.step(None, k * dt, dt)
        rp.render_device(sim, cam, img.ptr.value, "temperature", w, h, k * dt)
        if bloom:
            bloom.render_device(sim, img.ptr.value, canvas.ptr.value, w, h)

    # The V-cycle kernel plan is tuned on the first step's active set (results do
    # not depend on the plan), so every warm-up and timed step runs the final plan.
    if args.vcycle_plan:
        down, up = (int(v, 0) for v in args.vcycle_plan.split(","))
        sim.set_vcycle_plan(down, up)
        args.autotune = False
    full_step(0)
    sim.synchronize()
    vcycle = wl["desc"].solver == abi.FX_SOLVER_VCYCLE
    plan = None
    if vcycle:
        plan = sim.autotune_vcycle(4) if args.autotune else sim.vcycle_plan()
    for k in range(1, args.warmup):
        full_step(k)
    sim.synchronize()
    active0 = sim.active_brick_count()
    t0 = time.perf_counter()
    for k in range(args.warmup, args.warmup + args.steps):
        full_step(k)
    sim.synchronize()
    t1 = time.perf_counter()
    active1 = sim.active_brick_count()
    ms = (t1 - t0) * 1e3 / args.steps
    active = 0.5 * (active0 + active1)
    voxels = active * 512
    value = voxels / (ms * 1e-3)

    # Per-stage device time (HIP events on the simulation's stream).
    sim.set_stage_timing(True)
    base = args.warmup + args.steps
    for k in range(base, base + args.timing_steps):
        full_step(k)
    stages = sim.stage_times()
    active_t = sim.active_brick_count() * 512
    sim.set_stage_timing(False)
    stage_ms = {k: v[0] / v[1] for k, v in stages.items()}
    # Dominant kernel: the single-kernel stage with the most device time per step on
    # the simulation stream (the raymarch runs overlapped on its own stream). The
    # fine-level sweep launches ~15 times per step; its mean launch time is that of
    # the production sweeps inside the stage-timed steps, each timed by HIP events
    # that take the kernel's own start and end timestamps (hipExtLaunchKernelGGL), so
    # it is the duration rocprof reports for the same kernel; other stages use their
    # per-stage event pairs.
    totals = {k: stages[k][0] for k in stages if k in STAGE_BYTES}
    if not totals:
        raise SystemExit("--timing-steps must be >= 1 to find the dominant kernel")
    dom = max(totals, key=totals.get)
    avg_ms, vox_launch = stage_ms[dom], active_t
    ray_bytes = RAY_BYTES_PER_VOXEL - saved * RAY_SCALAR_ACCESSES
    sim_bytes = (STEP_BYTES_SIM if saved == 0 else STEP_BYTES_SIM_16) if vcycle else STEP_BYTES_SIM_REF[saved]
    step_bytes_voxel = sim_bytes + ray_bytes
    bytes_launch = (STAGE_BYTES[dom] - saved * SCALAR_ACCESSES.get(dom, 0)) * vox_launch
    achieved = bytes_launch / (avg_ms * 1e-3) / 1e9
    traffic = TRAFFIC_PER_VOXEL[dom] * vox_launch if dom in TRAFFIC_PER_VOXEL and saved == 0 else None
    step_bytes = step_bytes_voxel * voxels + 8.0 * w * h
    step_gbs = step_bytes / (ms * 1e-3) / 1e9

    out = {
        "metric": "active voxel-updates/s per full step @256³ sparse; % HBM roofline vs plain",
        "value": value,
        "unit": "active voxel-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": wl["label"], "storage": args.precision, "active_bricks": active,
                   "vcycle_plan": ({"down_split": plan[0], "up_split": plan[1], "autotuned": args.autotune} if plan
                                   else "reference pressure schedule"),
                   "active_fraction": active / sim.brick_count,
                   "render": f"{w}x{h}x{sc}" + (" + bloom/composite" if args.bloom else "")},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "bytes_per_launch": bytes_launch, "avg_launch_ms": avg_ms, "voxels_per_launch": vox_launch},
        "step_hbm": {"algorithmic_gbs": step_gbs, "frac": step_gbs / HBM_PEAK_GBS,
                     "bytes_per_active_voxel": step_bytes_voxel},
        "stage_ms": stage_ms,
    }
    if args.copy_probe > 0:
        nbytes = args.copy_probe << 20
        src, dst = HipBuffer(nbytes), HipBuffer(nbytes)
        if src.hip.hipMemcpy(dst.ptr, src.ptr, C.c_size_t(nbytes), 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("probe copy failed")
        src.hip.hipDeviceSynchronize()
        src.free()
        dst.free()
        out["copy_probe_bytes"] = nbytes
    # the same scene stepped without its raymarch (the simulation alone)
    t_sim = base + args.timing_steps
    sim.synchronize()
    a0 = sim.active_brick_count()
    t0 = time.perf_counter()
    for k in range(t_sim, t_sim + args.steps):
        sim.step(None, k * dt, dt)
    sim.synchronize()
    ms_sim = (time.perf_counter() - t0) * 1e3 / args.steps
    sim_value = 0.5 * (a0 + sim.active_brick_count()) * 512 / (ms_sim * 1e-3)
    out["sim_only"] = {"ms_per_step": ms_sim, "value": sim_value}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(wl, sim, (t_sim + args.steps) * dt, args.cpu_steps)
    img.free()
    if canvas:
        canvas.free()
    sim.dispose()
    if args.workload == "c2" and not args.no_batch and args.precision == "fp32":
        # batched + mixed precision against this plain FP32 single-scene path
        bm = batched_mixed(autotune=args.autotune and not args.vcycle_plan)
        bm["vs_plain"] = bm["value"] / value
        bm["vs_plain_sim_only"] = bm["value"] / sim_value
        out["batched_mixed"] = bm
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
