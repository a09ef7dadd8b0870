"""A sequence of related Poisson solves: warm starts and the learned correction.

Workload (learned smoother and warm start): solves s = 0, 1, ... of
-lap(u) = f_s on the unit cube (vertex-centred n^3 grid, u = 0 Dirichlet),
where f_s is the 8-Gaussian source whose centres drift with seeded
velocities of magnitude ``speed`` per solve (reflected inside [0.2, 0.8]^3
so the source stays in the box).  Every solve runs to
||r_k|| <= rtol * ||r(u = 0)|| (RMS norms of the residual hook), so warm
starts are measured against the same reference as cold solves.

* warm starts: "proj" (default) -- the history projection of
  pmg_history_warm_start: u_0 = sum_k c_k u_{s-k} over the last K solves,
  c minimising ||f_s - sum_k c_k f_{s-k}|| (the residual of u_0 up to the
  earlier solves' own residuals, L being linear); "extrap2" -- second-order
  extrapolation u_0 = 2 u_{s-1} - u_{s-2} (u_{s-1} for s = 1); "zero" --
  u_0 = 0, the cold solve;
* training data: (residual, error) crops from plain FP64 V-cycles of the
  first solves, the error taken against a tightly converged FP64 solution;
  inputs are scaled by 1/||r||, targets by 1/(||r|| h^2) -- the scaling the
  solver's learned step (pmg_solve_rel mode 2) applies;
* the network is trained on the GPU by pmg_net_train_step (BF16 MFMA
  forward/backward, Adam).

Everything here is host orchestration of the C ABI; sources are built on the
device with torch (workload synthesis only).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import Net, Solver, gaussian_params, lib, tptr, uniforms, PMG_F_RHO, PMG_F_RES, PMG_F_U


class DriftingSource:
    """8 Gaussians whose centres move by `speed` per solve (seeded directions)."""

    def __init__(self, seed=1234, m=8, speed=0.002, vseed=4321):
        self.c0, self.w, self.A = gaussian_params(seed, m)
        u = uniforms(vseed, 2 * m).reshape(m, 2)
        ct = 2.0 * u[:, 0] - 1.0
        st = np.sqrt(np.maximum(0.0, 1.0 - ct * ct))
        ph = 2.0 * math.pi * u[:, 1]
        self.v = speed * np.stack([st * np.cos(ph), st * np.sin(ph), ct], axis=1)

    def centres(self, s):
        x = np.mod(self.c0 + s * self.v - 0.2, 1.2)
        x = np.where(x > 0.6, 1.2 - x, x)
        return x + 0.2

    def rho(self, n, s, torch):
        """-f_s on the device (float64, (n, n, n) = (z, y, x)), boundary zero."""
        c = self.centres(s)
        x = torch.arange(n, dtype=torch.float64, device="cuda") / (n - 1)
        f = torch.zeros((n, n, n), dtype=torch.float64, device="cuda")
        for k in range(len(self.A)):
            w2 = float(self.w[k] * self.w[k])
            gx = torch.exp(-(x - float(c[k, 0])) ** 2 / w2)
            gy = torch.exp(-(x - float(c[k, 1])) ** 2 / w2)
            gz = torch.exp(-(x - float(c[k, 2])) ** 2 / w2)
            f += float(self.A[k]) * gz[:, None, None] * gy[None, :, None] * gx[None, None, :]
        f[0], f[-1] = 0, 0
        f[:, 0], f[:, -1] = 0, 0
        f[:, :, 0], f[:, :, -1] = 0, 0
        return -f


def _copy_grid(dst_ptr, src_ptr, n, stream=None):
    """Whole-grid device copy (pmg_dmcopy3d) on `stream`, then a device-wide
    synchronisation: torch work and solver work run on different streams and
    meet only at these hand-off points."""
    import torch
    torch.cuda.synchronize()
    rc = lib().pmg_dmcopy3d(src_ptr, dst_ptr, n, n, n, n, n, n, 1, 1, 1, 1, 1, 1, n, n, n, 1, 1, 1, 1, 1, 1, stream)
    if rc != 0:
        raise RuntimeError(f"pmg_dmcopy3d returned {rc}")
    torch.cuda.synchronize()


class Sequence:
    """Runs the drifting sequence on one Solver, keeping the last two solutions."""

    def __init__(self, n=257, seed=1234, speed=0.002, torch=None, hist_k=8):
        import torch as _t
        self.torch = torch or _t
        self.n = n
        self.src = DriftingSource(seed=seed, speed=speed)
        self.S = Solver(n)
        self.S.set_params(pre=2, pst=2, max_citer=5000, crtol=1e-3, rop_u=3, rop_res=3, rop_src=3)
        self.Lf = self.S.nlev
        t = self.torch
        self.hist_k = max(2, hist_k)
        self.hist = [t.zeros((n, n, n), dtype=t.float64, device="cuda") for _ in range(self.hist_k)]
        self.hist_f = [t.zeros((n, n, n), dtype=t.float64, device="cuda") for _ in range(self.hist_k)]
        self.nhist = 0
        self.h = 1.0 / (n - 1)
        self.rho_cur = None
        self.last_misfit = None

    def load_source(self, s):
        """Load f_s into the solver (finest rho) and restrict it; returns
        ||r(u = 0)|| (the reference norm of the relative test)."""
        t = self.torch
        rho = self.src.rho(self.n, s, t)
        _copy_grid(self.S.field_ptr(self.Lf, PMG_F_RHO), tptr(rho), self.n)
        self.rho_cur = rho
        rref = float(t.sqrt(t.mean(rho[1:-1, 1:-1, 1:-1] ** 2)))
        self.S.init()
        t.cuda.synchronize()
        return rref

    def warm_start(self, method="proj", misfit=True):
        """Set the initial guess on the finest level (see the module text);
        for "proj" with misfit, last_misfit = ||f_s - sum c_k f_{s-k}|| /
        ||f_s|| (one more pass over the sources)."""
        L = lib()
        self.last_misfit = None
        if method == "zero" or self.nhist == 0:
            self.S.zero_solution()
        elif method == "proj":
            K = min(self.nhist, self.hist_k)
            self.last_c, self.last_misfit = self.S.history_warm_start(
                [tptr(f) for f in self.hist_f[:K]], [tptr(u) for u in self.hist[:K]], misfit=misfit)
        elif method == "extrap2" and self.nhist == 1:
            L.pmg_set_solution_lincomb(self.S.h, tptr(self.hist[0]), None, 1.0, 0.0)
        elif method == "extrap2":
            L.pmg_set_solution_lincomb(self.S.h, tptr(self.hist[0]), tptr(self.hist[1]), 2.0, -1.0)
        else:
            raise ValueError(method)

    def prepare(self, s, warm=True, method="proj"):
        """load_source(s) and the initial guess; returns ||r(u = 0)||."""
        rref = self.load_source(s)
        self.warm_start(method if warm else "zero")
        self.torch.cuda.synchronize()
        return rref

    def push_solution(self):
        """(f_s, u_s) becomes the newest history entry."""
        t = self.torch
        u, f = self.hist.pop(), self.hist_f.pop()
        _copy_grid(tptr(u), self.S.field_ptr(self.Lf, PMG_F_U), self.n)
        f.copy_(self.rho_cur)
        t.cuda.synchronize()
        self.hist.insert(0, u)
        self.hist_f.insert(0, f)
        self.nhist = min(self.hist_k, self.nhist + 1)

    def solve(self, mode, rtol, rref, max_iter=100):
        return self.S.solve_rel(mode, rtol, max_iter, r_ref=rref)

    # ------------------------------------------------------------------
    def collect(self, nsolves, pairs, side=16, cycles=8, rtol=1e-10, seed=7):
        """(inputs, targets): device float32 [pairs, side^3] crops of scaled
        residuals and errors from plain FP64 V-cycles of solves 0..nsolves-1."""
        t = self.torch
        n = self.n
        rng = np.random.default_rng(seed)
        per = max(1, pairs // (nsolves * cycles))
        X = t.empty((nsolves * cycles * per, side ** 3), dtype=t.float32, device="cuda")
        Y = t.empty_like(X)
        ustar = t.empty((n, n, n), dtype=t.float64, device="cuda")
        u0 = t.empty_like(ustar)
        L = lib()
        stream = self.S.stream()
        row = 0
        for s in range(nsolves):
            rref = self.prepare(s)
            _copy_grid(tptr(u0), self.S.field_ptr(self.Lf, PMG_F_U), n)
            self.solve(0, 1e-13, rref, max_iter=60)  # the reference solution
            _copy_grid(tptr(ustar), self.S.field_ptr(self.Lf, PMG_F_U), n)
            L.pmg_set_solution_lincomb(self.S.h, tptr(u0), None, 1.0, 0.0)
            self.S.init()
            t.cuda.synchronize()
            for k in range(cycles):
                r = self.S.residual()  # writes f_res on the finest level
                org = rng.integers(0, n - side + 1, size=(per, 3)).astype(np.int32)
                op = org.ctypes.data_as(C.POINTER(C.c_int))
                rc = L.pmg_gather_crops(self.S.field_ptr(self.Lf, PMG_F_RES), None, n, n, n, op, per, side,
                                        1.0 / r, tptr(X[row:row + per]), stream)
                rc |= L.pmg_gather_crops(tptr(ustar), self.S.field_ptr(self.Lf, PMG_F_U), n, n, n, op, per, side,
                                         1.0 / (r * self.h * self.h), tptr(Y[row:row + per]), stream)
                if rc:
                    raise RuntimeError("pmg_gather_crops failed")
                row += per
                self.S.vcycle()
            # the sequence continues from the converged solution
            L.pmg_set_solution_lincomb(self.S.h, tptr(ustar), None, 1.0, 0.0)
            t.cuda.synchronize()
            self.push_solution()
        return X[:row], Y[:row]


def train(net: Net, X, Y, side=16, steps=2000, batch=32, lr=2e-3, margin=3, seed=11, log=None):
    """Adam on random minibatches of the crops; returns the loss history."""
    import torch as t
    g = t.Generator(device="cpu").manual_seed(seed)
    losses = []
    for it in range(steps):
        idx = t.randint(0, X.shape[0], (batch,), generator=g).to(X.device)
        xb = X.index_select(0, idx).contiguous()
        yb = Y.index_select(0, idx).contiguous()
        cur = lr * (0.5 * (1 + math.cos(math.pi * it / steps)))
        losses.append(net.train_step(xb, yb, (side, side, side), batch, cur, margin))
        if log and it % max(1, steps // 10) == 0:
            log(f"train step {it}: loss {losses[-1]:.4e}")
    return losses
