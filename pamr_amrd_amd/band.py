"""Training of the learned band correction for excised problems.

With excision the reference's transfers (renormalised full-weight
restriction, excised interpolation and clear_ex; src/m_util_r8.f:835-949,
368-421, src/transfer.c:930-935) leave the coarse-grid correction short in
a thin layer of fine points next to each excised region, and the V-cycle
contracts slowly (0.5-0.8 per iteration, growing with the grid).  The
network (3x3x3 convolutions 1 -> 16 -> 16 -> 1, GELU, BF16 MFMA with FP32
accumulation) learns the error in that layer from the residual around it:
its correction u += ts ||r|| N(r / ||r||) is applied at the band points only
(pmg_solve_rel mode 3 and the FGMRES preconditioner, pmg_set_band_net).

Training data are (residual, error) crops taken from the plain
mixed-precision iteration (FP32 V-cycles in FP64 defect correction) of a few
training right-hand sides, the error against a tightly converged solution;
crops are centred on random band points, and the loss runs over the band
points of each crop's core (points at least 3 from its faces).  Inputs are
scaled by 1/||r||, targets by 1/(||r|| ts) with ts the RMS of e/||r|| over
the training band points.  Everything runs on the GPU (pmg_band_sample,
pmg_net_train_step_w); this module only orchestrates the C ABI, with torch
holding the crop tensors.
"""
from __future__ import annotations

import math
import time

from . import Net, PMG_F_RHO, PMG_F_U, lib, tptr


def set_state(S, rho, u0):
    """Finest source = rho, u = u0 (device FP64 tensors of the finest shape;
    u0 None: zero), then pmg_init."""
    from .sequence import _copy_grid
    n = S.shape(S.nlev)[0]
    _copy_grid(S.field_ptr(S.nlev, PMG_F_RHO), tptr(rho).value, n)
    if u0 is None:
        S.zero_solution()
    elif lib().pmg_set_solution_lincomb(S.h, tptr(u0), None, 1.0, 0.0) != 0:
        raise RuntimeError("pmg_set_solution_lincomb failed")
    S.init()


def collect_pairs(S, rhos, ustars, u0, torch, iters=10, crops=48, seed=77):
    """(x, t, w) crop tensors from `iters` plain mixed iterations of each
    training right-hand side (device tensors rhos[k], converged solutions
    ustars[k], initial guess u0); t is e/||r|| (not yet divided by ts)."""
    _, ncore, side = S.band_info()
    if side <= 0:
        raise RuntimeError("band_setup first")
    total = len(rhos) * iters * crops
    x = torch.empty((total, side, side, side), dtype=torch.float32, device="cuda")
    t = torch.empty_like(x)
    w = torch.empty_like(x)
    at = 0
    for k, (rp, up) in enumerate(zip(rhos, ustars)):
        set_state(S, rp, u0)
        for it in range(iters):
            _, h = S.solve_rel(1, 0.0, 1)
            r = float(h[-1])
            torch.cuda.synchronize()
            S.band_sample(tptr(up), -1.0 / r, 1.0 / r, crops, seed + 1000 * k + it, tptr(x[at:]),
                          tptr(t[at:]), tptr(w[at:]))
            at += crops
    torch.cuda.synchronize()
    return x, t, w


def collect_pairs_inner(S, rhos, ustars, u0, torch, iters=10, crops=48, seed=91):
    """(x, t, w) crop tensors for the in-cycle placement: inside each of
    `iters` mixed iterations of every training right-hand side, after the
    finest coarse-grid correction of the FP32 cycle (pmg_band_sample_inner);
    t is the error of the cycle's correction at that moment over the outer
    defect's RMS (not yet divided by ts)."""
    _, ncore, side = S.band_info()
    if side <= 0:
        raise RuntimeError("band_setup first")
    total = len(rhos) * iters * crops
    x = torch.empty((total, side, side, side), dtype=torch.float32, device="cuda")
    t = torch.empty_like(x)
    w = torch.empty_like(x)
    at = 0
    for k, (rp, up) in enumerate(zip(rhos, ustars)):
        set_state(S, rp, u0)
        S.solve_rel(1, 0.0, 0)  # the outer defect of u0 as the first FP32 source
        for it in range(iters):
            torch.cuda.synchronize()
            S.band_sample_inner(tptr(up), crops, seed + 1000 * k + it, tptr(x[at:]), tptr(t[at:]), tptr(w[at:]))
            at += crops
    torch.cuda.synchronize()
    return x, t, w


def train(x, t, w, torch, steps=4000, batch=16, lr=2e-3, seed=5, margin=3, net=None, log=None):
    """Train a network on crop tensors (targets scaled here by 1/ts, ts the
    RMS of t over the weighted core points), Adam with the learning rate
    decaying as a half cosine to zero over the steps (the last weights are
    then a converged state, not a sample of the constant-rate noise).
    Returns (net, ts, info)."""
    side = x.shape[-1]
    core = torch.zeros_like(w[0])
    core[margin:side - margin, margin:side - margin, margin:side - margin] = 1.0
    wc = w * core
    ts = float(torch.sqrt((t * t * wc).sum() / wc.sum()))
    tn = t / ts
    if net is None:
        net = Net(seed=seed)
    g = torch.Generator(device="cuda").manual_seed(seed)
    t0 = time.perf_counter()
    losses = []
    for s in range(steps):
        idx = torch.randint(0, x.shape[0], (batch,), generator=g, device="cuda")
        xb, tb, wb = x[idx].contiguous(), tn[idx].contiguous(), w[idx].contiguous()
        torch.cuda.synchronize()
        lr_s = 0.5 * lr * (1.0 + math.cos(math.pi * s / steps))
        loss = net.train_step(xb, tb, (side, side, side), batch, lr_s, margin=margin, w_dev=wb)
        losses.append(loss)
        if log and (s % 250 == 0 or s == steps - 1):
            log(f"band net step {s}: loss {loss:.4e}")
    torch.cuda.synchronize()
    info = {"steps": steps, "batch": batch, "lr": lr, "pairs": int(x.shape[0]), "crop": side,
            "train_s": round(time.perf_counter() - t0, 2), "ts": ts,
            "loss_first_50": float(sum(losses[:50]) / max(1, len(losses[:50]))),
            "loss_last_50": float(sum(losses[-50:]) / max(1, len(losses[-50:])))}
    return net, ts, info


def converged_solution(S, rho, u0, torch, rtol=1e-11, m=8, max_iter=400):
    """A tightly converged solution (flexible GMRES(m) with the FP32 V-cycle)
    of one right-hand side, copied into a new device tensor."""
    from .sequence import _copy_grid
    n = S.shape(S.nlev)[0]
    set_state(S, rho, u0)
    r0, h = S.solve_fgmres(m, rtol, max_iter)
    if len(h) == 0 or h[-1] > 10 * rtol * r0:
        raise RuntimeError(f"training solve did not converge: {h[-3:] if len(h) else h}")
    u = torch.empty((n, n, n), dtype=torch.float64, device="cuda")
    _copy_grid(tptr(u).value, S.field_ptr(S.nlev, PMG_F_U), n)
    return u
