"""FP64 numpy restatement of the differentiable smoke step (fx_diff_* in
include/fenix.h) and of its adjoint.

Test infrastructure only: this is the oracle of the HIP forward and adjoint
kernels, never a path the library runs. The reference has no gradients; the
forward stages restate the library's own model, whose advection and
confinement follow the reference's formulas:

* velocity advection, SH/advect-velocity-semilagrangian.wgsl.ts:23-41 (sample at
  x - u(x) dt, boundary 0; without the 0.025 damping);
* MacCormack with the min/max limiter, SH/advect-scalar-maccormack.wgsl.ts:33-101
  (without the decay terms);
* vorticity, SH/compute-vorticity.wgsl.ts:23-50, and the confinement direction
  grad|w| / max(|grad|w||, 1e-4), SH/apply-vorticity-confinement.wgsl.ts:30-66
  (without masks and the speed clamp);
* the projection is injected by the caller (`project`): tests pass the FP64
  oracle's divergence -> V-cycle -> projection passes.

Arrays are row-major [z][y][x]. The adjoint is hand-derived, stage by stage;
tests pin it with central finite differences of this FP64 forward, and pin the
projection's self-adjointness with a dot-product test.
"""
import numpy as np

CONFINE_FLOOR = 1e-4


class Sampler:
    """Clamped trilinear sample positions (tri() of the HIP kernels)."""

    def __init__(self, shape, px, py, pz):
        D, H, W = shape
        self.shape = shape
        self.p = (px, py, pz)
        lims = (W - 1, H - 1, D - 1)
        self.i0, self.i1, self.f, self.free = [], [], [], []
        for p, m in zip((px, py, pz), lims):
            c = np.clip(p, 0.0, m)
            i0 = np.floor(c).astype(np.int64)
            self.i0.append(i0)
            self.i1.append(np.minimum(i0 + 1, m))
            self.f.append(c - i0)
            self.free.append((p > 0.0) & (p < m))

    def corners(self):
        """Flat indices and weights of the 8 corners (bit 0 x, bit 1 y, bit 2 z)."""
        D, H, W = self.shape
        out = []
        for k in range(8):
            ix = self.i1[0] if k & 1 else self.i0[0]
            iy = self.i1[1] if k & 2 else self.i0[1]
            iz = self.i1[2] if k & 4 else self.i0[2]
            wx = self.f[0] if k & 1 else 1.0 - self.f[0]
            wy = self.f[1] if k & 2 else 1.0 - self.f[1]
            wz = self.f[2] if k & 4 else 1.0 - self.f[2]
            out.append(((iz * H + iy) * W + ix, wx * wy * wz))
        return out

    def eval(self, field):
        flat = field.reshape(-1)
        return sum(w * flat[i] for i, w in self.corners())

    def grad(self, field):
        """d(sample)/d(position), 0 along clamped axes."""
        flat = field.reshape(-1)
        c = [flat[i] for i, _ in self.corners()]
        fx, fy, fz = self.f
        ax, ay, az = 1 - fx, 1 - fy, 1 - fz
        gx = ((c[1] - c[0]) * ay + (c[3] - c[2]) * fy) * az + ((c[5] - c[4]) * ay + (c[7] - c[6]) * fy) * fz
        gy = ((c[2] - c[0]) * ax + (c[3] - c[1]) * fx) * az + ((c[6] - c[4]) * ax + (c[7] - c[5]) * fx) * fz
        gz = ((c[4] - c[0]) * ax + (c[5] - c[1]) * fx) * ay + ((c[6] - c[2]) * ax + (c[7] - c[3]) * fx) * fy
        return [np.where(self.free[0], gx, 0.0), np.where(self.free[1], gy, 0.0), np.where(self.free[2], gz, 0.0)]

    def scatter(self, out, a):
        """out[corner] += weight * a (transposed gather)."""
        n = out.size
        flat = out.reshape(-1)
        for i, w in self.corners():
            flat += np.bincount(i.reshape(-1), weights=(w * a).reshape(-1), minlength=n)


def _cdiff(a, axis):
    """0.5 (a[i+e] - a[i-e]) at interior voxels, 0 elsewhere; axis 0 = z, 1 = y, 2 = x."""
    out = np.zeros_like(a)
    s_in = [slice(1, -1)] * 3
    s_hi = list(s_in)
    s_lo = list(s_in)
    s_hi[axis] = slice(2, None)
    s_lo[axis] = slice(None, -2)
    out[tuple(s_in)] = 0.5 * (a[tuple(s_hi)] - a[tuple(s_lo)])
    return out


def _cdiff_t(h, axis):
    """Transpose of _cdiff."""
    out = np.zeros_like(h)
    s_in = [slice(1, -1)] * 3
    s_hi = list(s_in)
    s_lo = list(s_in)
    s_hi[axis] = slice(2, None)
    s_lo[axis] = slice(None, -2)
    hi = h[tuple(s_in)]
    out[tuple(s_hi)] += 0.5 * hi
    out[tuple(s_lo)] -= 0.5 * hi
    return out


X_AX, Y_AX, Z_AX = 2, 1, 0


def curl(u):
    ux, uy, uz = u
    return [_cdiff(uz, Y_AX) - _cdiff(uy, Z_AX), _cdiff(ux, Z_AX) - _cdiff(uz, X_AX), _cdiff(uy, X_AX) - _cdiff(ux, Y_AX)]


def curl_t(wb):
    wx, wy, wz = wb
    return [_cdiff_t(wy, Z_AX) - _cdiff_t(wz, Y_AX), -_cdiff_t(wx, Z_AX) + _cdiff_t(wz, X_AX),
            _cdiff_t(wx, Y_AX) - _cdiff_t(wy, X_AX)]


class DiffModel:
    """The differentiable step on a (D, H, W) grid in FP64.

    project(us) -> u must be the pressure projection (velocity 0 on the boundary)."""

    def __init__(self, shape, steps, control_n, slab, dt, buoyancy, confinement, project):
        self.shape = tuple(shape)
        self.steps, self.cn, self.slab = steps, control_n, slab
        self.dt, self.beta, self.eps = float(dt), float(buoyancy), float(confinement)
        self.project = project
        D, H, W = self.shape
        self.Z, self.Y, self.X = np.meshgrid(np.arange(D, dtype=np.float64), np.arange(H, dtype=np.float64),
                                             np.arange(W, dtype=np.float64), indexing="ij")
        I = np.zeros(self.shape, dtype=bool)
        I[1:-1, 1:-1, 1:-1] = True
        self.interior = I
        self.slab_mask = I & (self.Y >= 1) & (self.Y <= slab)
        self._taps()
        self.target = np.zeros(self.shape)

    # -- source ---------------------------------------------------------------
    def _taps(self):
        """Control cells and weights of every slab voxel (cell-centred trilinear)."""
        D, H, W = self.shape
        cn, m = self.cn, self.cn - 1
        vox = np.nonzero(self.slab_mask.reshape(-1))[0]
        z, rem = np.divmod(vox, H * W)
        y, x = np.divmod(rem, W)
        px = np.clip((x + 0.5) / W * cn - 0.5, 0, m)
        py = np.clip((y - 1 + 0.5) / self.slab * cn - 0.5, 0, m)
        pz = np.clip((z + 0.5) / D * cn - 0.5, 0, m)
        i0 = [np.floor(p).astype(np.int64) for p in (px, py, pz)]
        i1 = [np.minimum(a + 1, m) for a in i0]
        fr = [p - a for p, a in zip((px, py, pz), i0)]
        self.tap_vox = vox
        self.tap_idx, self.tap_w = [], []
        for k in range(8):
            ix = i1[0] if k & 1 else i0[0]
            iy = i1[1] if k & 2 else i0[1]
            iz = i1[2] if k & 4 else i0[2]
            w = (fr[0] if k & 1 else 1 - fr[0]) * (fr[1] if k & 2 else 1 - fr[1]) * (fr[2] if k & 4 else 1 - fr[2])
            self.tap_idx.append((iz * cn + iy) * cn + ix)
            self.tap_w.append(w)

    def source_rates(self, ctrl):
        """S per channel (u_x, u_y, u_z, rho) on the grid."""
        ctrl = np.asarray(ctrl, dtype=np.float64).reshape(4, -1)
        out = []
        for ch in range(4):
            s = np.zeros(int(np.prod(self.shape)))
            s[self.tap_vox] = sum(w * ctrl[ch][i] for i, w in zip(self.tap_idx, self.tap_w))
            out.append(s.reshape(self.shape))
        return out

    def source_t(self, rates_bar):
        """Transpose of source_rates."""
        g = np.zeros((4, self.cn ** 3))
        for ch in range(4):
            a = rates_bar[ch].reshape(-1)[self.tap_vox]
            for i, w in zip(self.tap_idx, self.tap_w):
                g[ch] += np.bincount(i, weights=w * a, minlength=self.cn ** 3)
        return g.reshape(4, self.cn, self.cn, self.cn)

    # -- forward --------------------------------------------------------------
    def step(self, d, u, rates):
        dt = self.dt
        ds = d + dt * rates[3]
        us = [u[c] + dt * rates[c] for c in range(3)]
        t = Sampler(self.shape, self.X - us[0] * dt, self.Y - us[1] * dt, self.Z - us[2] * dt)
        ua = [np.where(self.interior, t.eval(us[c]), 0.0) for c in range(3)]
        dn = self._maccormack(ds, ua)
        ub = [ua[0].copy(), ua[1] + np.where(self.interior, self.beta * dt * dn, 0.0), ua[2].copy()]
        w = curl(ub)
        m = np.sqrt(w[0] ** 2 + w[1] ** 2 + w[2] ** 2)
        g = [_cdiff(m, X_AX), _cdiff(m, Y_AX), _cdiff(m, Z_AX)]
        n = np.sqrt(g[0] ** 2 + g[1] ** 2 + g[2] ** 2)
        den = np.maximum(n, CONFINE_FLOOR)
        N = [gc / den for gc in g]
        f = [N[1] * w[2] - N[2] * w[1], N[2] * w[0] - N[0] * w[2], N[0] * w[1] - N[1] * w[0]]
        uc = [ub[c] + np.where(self.interior, self.eps * dt * f[c], 0.0) for c in range(3)]
        un = self.project(uc)
        return dn, un, (ds, us, ua, ub)

    def _mac_parts(self, ds, ua):
        dt = self.dt
        tp = Sampler(self.shape, self.X - ua[0] * dt, self.Y - ua[1] * dt, self.Z - ua[2] * dt)
        fwd = [tp.eval(ua[c]) for c in range(3)]
        sr = [tp.p[c] + fwd[c] * dt for c in range(3)]
        tr = Sampler(self.shape, *sr)
        return tp, tr, tp.eval(ds), tr.eval(ds)

    def _maccormack(self, ds, ua):
        _, _, pred, rec = self._mac_parts(ds, ua)
        corr = pred + 0.5 * (ds - rec)
        return np.clip(corr, np.minimum(ds, pred), np.maximum(ds, pred))

    def rollout(self, ctrl, keep=False):
        rates = self.source_rates(ctrl)
        d = np.zeros(self.shape)
        u = [np.zeros(self.shape) for _ in range(3)]
        tapes = []
        for _ in range(self.steps):
            d, u, tape = self.step(d, u, rates)
            if keep:
                tapes.append(tape)
        return d, u, tapes

    def loss(self, ctrl):
        d, _, _ = self.rollout(ctrl)
        return 0.5 * float(np.sum((d - self.target) ** 2))

    # -- adjoint --------------------------------------------------------------
    def gradient(self, ctrl):
        d, _, tapes = self.rollout(ctrl, keep=True)
        loss = 0.5 * float(np.sum((d - self.target) ** 2))
        db = d - self.target
        ub = [np.zeros(self.shape) for _ in range(3)]
        rates_bar = [np.zeros(self.shape) for _ in range(4)]
        for tape in reversed(tapes):
            db, ub, rb = self.step_adjoint(tape, db, ub)
            for c in range(4):
                rates_bar[c] += rb[c]
        return loss, self.source_t(rates_bar)

    def step_adjoint(self, tape, dnb, unb):
        ds, us, ua, ubt = tape
        dt, I = self.dt, self.interior
        # projection: self-adjoint on fields that vanish on the boundary
        a = self.project([np.where(I, c, 0.0) for c in unb])
        # confinement
        w = curl(ubt)
        m = np.sqrt(w[0] ** 2 + w[1] ** 2 + w[2] ** 2)
        g = [_cdiff(m, X_AX), _cdiff(m, Y_AX), _cdiff(m, Z_AX)]
        n = np.sqrt(g[0] ** 2 + g[1] ** 2 + g[2] ** 2)
        big = n > CONFINE_FLOOR
        den = np.where(big, n, CONFINE_FLOOR)
        N = [gc / den for gc in g]
        fb = [np.where(I, self.eps * dt * a[c], 0.0) for c in range(3)]
        Nb = [w[1] * fb[2] - w[2] * fb[1], w[2] * fb[0] - w[0] * fb[2], w[0] * fb[1] - w[1] * fb[0]]  # w x fbar
        wbar = [fb[1] * N[2] - fb[2] * N[1], fb[2] * N[0] - fb[0] * N[2], fb[0] * N[1] - fb[1] * N[0]]  # fbar x N
        ndot = N[0] * Nb[0] + N[1] * Nb[1] + N[2] * Nb[2]
        gbar = [np.where(I, np.where(big, (Nb[c] - N[c] * ndot) / den, Nb[c] / CONFINE_FLOOR), 0.0) for c in range(3)]
        mbar = _cdiff_t(gbar[0], X_AX) + _cdiff_t(gbar[1], Y_AX) + _cdiff_t(gbar[2], Z_AX)
        safe = np.where(m > 0, m, 1.0)
        for c in range(3):
            wbar[c] = np.where(I, wbar[c] + np.where(m > 0, mbar / safe, 0.0) * w[c], 0.0)
        ct = curl_t(wbar)
        abar = [a[c] + ct[c] for c in range(3)]
        # buoyancy
        dnb = dnb + np.where(I, self.beta * dt * abar[1], 0.0)
        # MacCormack
        tp, tr, pred, rec = self._mac_parts(ds, ua)
        corr = pred + 0.5 * (ds - rec)
        lo, hi = np.minimum(ds, pred), np.maximum(ds, pred)
        below, above = corr < lo, corr > hi
        inside = ~below & ~above
        cb = np.where(inside, dnb, 0.0)
        origb = np.where(below & (ds <= pred), dnb, 0.0) + np.where(above & (ds >= pred), dnb, 0.0) + 0.5 * cb
        predb = np.where(below & (ds > pred), dnb, 0.0) + np.where(above & (ds < pred), dnb, 0.0) + cb
        recb = -0.5 * cb
        dsb = origb.copy()
        tp.scatter(dsb, predb)
        tr.scatter(dsb, recb)
        srb = [gc * recb for gc in tr.grad(ds)]
        gp = tp.grad(ds)
        spb = [srb[c] + gp[c] * predb for c in range(3)]
        fwdb = [srb[c] * dt for c in range(3)]
        for c in range(3):
            tp.scatter(abar[c], fwdb[c])
        for c in range(3):
            gc = tp.grad(ua[c])
            for k in range(3):
                spb[k] = spb[k] + gc[k] * fwdb[c]
        for k in range(3):
            abar[k] = abar[k] - dt * spb[k]
        # velocity advection (interior voxels; the boundary of u_a is the constant 0)
        ab = [np.where(I, abar[c], 0.0) for c in range(3)]
        t = Sampler(self.shape, self.X - us[0] * dt, self.Y - us[1] * dt, self.Z - us[2] * dt)
        usb = [np.zeros(self.shape) for _ in range(3)]
        for c in range(3):
            t.scatter(usb[c], ab[c])
        spb = [np.zeros(self.shape) for _ in range(3)]
        for c in range(3):
            gc = t.grad(us[c])
            for k in range(3):
                spb[k] = spb[k] + gc[k] * ab[c]
        for k in range(3):
            usb[k] = usb[k] - dt * spb[k]
        # source: d_s = d + dt S, u_s = u + dt S (identity for the state)
        rates_bar = [dt * usb[0], dt * usb[1], dt * usb[2], dt * dsb]
        return dsb, usb, rates_bar
