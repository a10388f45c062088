"""Differentiable smoke step over libfenix (fx_diff_* in include/fenix.h).

`SmokeRollout` owns one fx_diff handle: a dense grid, a rollout length, the
checkpoint interval and a control field of shape (4, n, n, n) with channels
(u_x, u_y, u_z, rho), trilinearly upsampled into the floor source slab.
`RolloutLoss` is the forward/backward Function pair: its forward runs the
rollout (with the adjoint pass when the control needs a gradient, keeping the
gradient for backward), its backward scales that gradient by the incoming one.
The reference has no gradients; this is the inflow-control optimisation path.
"""
import ctypes as C
from typing import Optional

import numpy as np

from . import abi


def diff_desc(resolution=(64, 64, 64), steps=64, checkpoint_every=8, control_n=32, slab=4, dt=1.0, buoyancy=0.02,
              confinement=0.02, vcycle=None, device=0, lib=None) -> abi.DiffDesc:
    lib = lib or abi.load()
    d = abi.DiffDesc()
    lib.fx_diff_default_desc(C.byref(d))
    d.width, d.height, d.depth = resolution
    d.steps, d.checkpoint_every, d.control_n, d.slab = steps, checkpoint_every, control_n, slab
    d.dt, d.buoyancy, d.confinement = dt, buoyancy, confinement
    for k, v in (vcycle or {}).items():
        setattr(d.vcycle, k, v)
    d.device = device
    return d


class SmokeRollout:
    """One differentiable rollout configuration on the GPU."""

    def __init__(self, desc: abi.DiffDesc, lib=None):
        self._lib = lib or abi.load()
        self.desc = desc
        self._h = C.c_void_p()
        rc = self._lib.fx_diff_create(C.byref(desc), C.byref(self._h))
        if rc != abi.FX_OK:
            msg = self._lib.fx_diff_last_error(None)
            raise abi.FenixError(rc, msg.decode() if msg else "")
        self.shape = (desc.depth, desc.height, desc.width)
        self.control_shape = (4, desc.control_n, desc.control_n, desc.control_n)

    def _check(self, rc):
        if rc != abi.FX_OK:
            msg = self._lib.fx_diff_last_error(self._h)
            raise abi.FenixError(rc, msg.decode() if msg else "")

    def _ctrl(self, control):
        arr = np.ascontiguousarray(np.asarray(control, dtype=np.float32))
        if arr.size != int(np.prod(self.control_shape)):
            raise ValueError(f"control must hold {self.control_shape} floats, got {arr.shape}")
        return arr

    def set_target(self, density):
        arr = np.ascontiguousarray(np.asarray(density, dtype=np.float32).reshape(self.shape))
        self._check(self._lib.fx_diff_set_target(self._h, arr.ctypes.data_as(C.POINTER(C.c_float)), arr.nbytes))

    def forward(self, control) -> float:
        arr = self._ctrl(control)
        loss = C.c_double()
        self._check(self._lib.fx_diff_forward(self._h, arr.ctypes.data_as(C.POINTER(C.c_float)), C.byref(loss)))
        return loss.value

    def gradient(self, control):
        """(loss, d loss / d control) with the control's shape."""
        arr = self._ctrl(control)
        grad = np.empty(self.control_shape, dtype=np.float32)
        loss = C.c_double()
        self._check(self._lib.fx_diff_gradient(self._h, arr.ctypes.data_as(C.POINTER(C.c_float)), C.byref(loss),
                                               grad.ctypes.data_as(C.POINTER(C.c_float))))
        return loss.value, grad

    def read_field(self, field=0):
        """Final state of the last rollout: 0 density, 1..3 velocity x, y, z."""
        out = np.empty(self.shape, dtype=np.float32)
        self._check(self._lib.fx_diff_read_field(self._h, int(field), out.ctypes.data_as(C.POINTER(C.c_float)), out.nbytes))
        return out

    def dispose(self):
        if self._h:
            self._lib.fx_diff_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.dispose()
        except Exception:
            pass


def _function():
    import torch

    class RolloutLoss(torch.autograd.Function):
        """loss = RolloutLoss.apply(control, rollout): forward runs the rollout; when
        the control requires grad it also runs the checkpointed adjoint and saves
        the gradient for backward."""

        @staticmethod
        def forward(ctx, control, rollout: SmokeRollout):
            c = control.detach().to("cpu", torch.float32).contiguous().numpy()
            if ctx.needs_input_grad[0]:
                loss, grad = rollout.gradient(c)
                ctx.save_for_backward(torch.from_numpy(grad).to(control.device))
            else:
                loss = rollout.forward(c)
            return torch.tensor(loss, dtype=torch.float64, device=control.device)

        @staticmethod
        def backward(ctx, grad_out):
            (grad,) = ctx.saved_tensors
            return (grad_out.to(grad.dtype) * grad, None)

    return RolloutLoss


_RL = None


def rollout_loss(control, rollout: SmokeRollout):
    """Differentiable loss of one rollout (torch tensor in, scalar tensor out)."""
    global _RL
    if _RL is None:
        _RL = _function()
    return _RL.apply(control, rollout)


def gaussian_blobs(shape, centres, sigma, amplitude=1.0):
    """Sum of isotropic Gaussian blobs on a (D, H, W) grid; centres as (x, y, z) voxels."""
    D, H, W = shape
    z, y, x = np.meshgrid(np.arange(D), np.arange(H), np.arange(W), indexing="ij")
    out = np.zeros(shape, dtype=np.float64)
    for cx, cy, cz in centres:
        out += amplitude * np.exp(-((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2) / (2.0 * sigma * sigma))
    return out.astype(np.float32)


def plume_control(control_n, rise=0.05, density=0.1, radius=0.3):
    """A starting control: a disc of upward velocity and density at the slab centre."""
    n = control_n
    c = (np.arange(n) + 0.5) / n - 0.5
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    disc = (x * x + z * z <= radius * radius).astype(np.float32)
    out = np.zeros((4, n, n, n), dtype=np.float32)
    out[1] = rise * disc
    out[3] = density * disc
    return out


def optimise_inflow(rollout: SmokeRollout, control, iterations, lr=1e-2, callback: Optional[callable] = None):
    """Adam (torch.optim) on the control; returns (control, losses)."""
    import torch

    param = torch.nn.Parameter(torch.as_tensor(np.asarray(control, dtype=np.float32)).clone())
    opt = torch.optim.Adam([param], lr=lr)
    losses = []
    for it in range(iterations):
        opt.zero_grad()
        loss = rollout_loss(param, rollout)
        loss.backward()
        opt.step()
        losses.append(float(loss))
        if callback:
            callback(it, float(loss))
    return param.detach().numpy().copy(), losses
