"""Writes tests/golden/c5_fp64_gradient.json: the FP64 reference gradient of the inflow-control
problem at its full size (128^3 dense, 64 steps, control 32^3 x 4, slab y in [1, 4], dt 1,
buoyancy 0.02, confinement 0.02, V-cycle 4 levels 2/2 x 4 with 16 coarse sweeps, target = 3
Gaussian blobs of sigma 12 voxels with centres from seed 5, control = the plume control of
tests/test_diff_gpu.py::c5_problem).

The gradient comes from oracle/diff_model.py (the FP64 restatement of the differentiable step
and its hand-derived adjoint, pinned against FP64 central differences at h = 1e-7) with the FP64
oracle's divergence -> V-cycle -> projection passes. The fixture holds the loss, the gradient's
L2 norm, and the gradient at 32 control entries drawn with seed 6 and 256 more with seed 7
among the entries the slab samples (non-zero gradient). Runs on the CPU (about 10 GB of memory,
~15-30 minutes); the GPU test compares fx_diff_gradient with it.

Usage: python tests/golden/make_c5_fp64_gradient.py
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from fenix_amd.differentiable import gaussian_blobs, plume_control  # noqa: E402
from oracle import OracleProjector, diff_model  # noqa: E402

C5 = dict(n=128, steps=64, control_n=32, slab=4, dt=1.0, buoyancy=0.02, confinement=0.02)


def c5_problem():
    """The same target and control as tests/test_diff_gpu.py::c5_problem."""
    n = C5["n"]
    rng = np.random.default_rng(5)
    centres = [(rng.uniform(0.3, 0.7) * n, rng.uniform(0.3, 0.55) * n, rng.uniform(0.3, 0.7) * n) for _ in range(3)]
    target = gaussian_blobs((n, n, n), centres, 12.0, amplitude=0.5)
    ctrl = plume_control(C5["control_n"], rise=0.05, density=0.1, radius=0.3)
    return target, ctrl


def main():
    n = C5["n"]
    shape = (n, n, n)
    target, ctrl = c5_problem()
    model = diff_model().DiffModel(shape, steps=C5["steps"], control_n=C5["control_n"], slab=C5["slab"], dt=C5["dt"],
                                   buoyancy=C5["buoyancy"], confinement=C5["confinement"],
                                   project=OracleProjector(shape))
    model.target = target.astype(np.float64)
    t0 = time.time()
    loss, grad = model.gradient(ctrl.astype(np.float64))
    flat = np.asarray(grad, np.float64).reshape(-1)
    live = np.nonzero(flat != 0)[0]
    pick = np.random.default_rng(6).choice(live, size=32, replace=False)
    extra = np.random.default_rng(7).choice(np.setdiff1d(live, pick), size=256, replace=False)
    out = {
        "config": "128^3 dense, 64 steps, control 32^3x4 (u_x, u_y, u_z, rho) into the slab y in [1, 4], dt 1, "
                  "buoyancy 0.02, confinement 0.02, V-cycle 4 levels 2/2 x4, 16 coarse sweeps, omega 2/3; "
                  "target 3 Gaussian blobs sigma 12 (centres seed 5); control plume_control(32, rise=0.05, "
                  "density=0.1, radius=0.3)",
        "source": "oracle/diff_model.py (FP64) with the FP64 oracle's V-cycle projection",
        "loss": float(loss),
        "grad_norm": float(np.linalg.norm(flat)),
        "live_entries": int(live.size),
        "entries": [int(i) for i in pick],
        "grad": [float(flat[i]) for i in pick],
        "extra_entries": [int(i) for i in extra],
        "extra_grad": [float(flat[i]) for i in extra],
        "seconds": time.time() - t0,
    }
    path = os.path.join(HERE, "c5_fp64_gradient.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
        fh.write("\n")
    print(f"loss {loss:.6g}, |grad| {out['grad_norm']:.6g}, {live.size} live entries, {out['seconds']:.0f} s -> {path}")


if __name__ == "__main__":
    main()
