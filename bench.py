#!/usr/bin/env python3
"""bench.py -- time to a 1e-10 relative FP64 residual per 257^3 Poisson solve:
learned + FP32 against plain FP64.

Problem: -lap(u) = f on the unit cube, vertex-centred 257^3 grid, u = 0
Dirichlet, MG hierarchy 257 -> 5 (7 levels), FAS V(2,2) red-black
Gauss-Seidel, full-weight restriction, trilinear prolongation; every solve
runs until ||r_k|| <= 1e-10 ||r(u = 0)|| (RMS norms), inputs resident in HBM.

``value`` (one step = one solve): the drifting sequence -- f_s is a sum of 8
Gaussians whose centres move 0.002 per solve; the network is trained on the
GPU from solves 0-49, and solves 50..49+K (K = --steps, default 150: solves
50-199) are timed after an untimed warm-up of every mode, each including its warm
start (the history projection over the last 8 solves), with FP32 inner
V-cycles in FP64 defect correction and the learned correction applied when
it pays.  The same solves are also timed with FP32 or plain FP64 cycles
from the same warm start, from second-order extrapolation, and plain FP64
from u = 0 (the reference's cold solve); ``single_solve`` times one solve
from u = 0 in plain FP64 (the reference method, amrd/mg.c) and mixed.

Extra fields: the roofline of the dominant kernel (the finest-level fused
red-black sweep, timed live with HIP events on the solver's stream, HBM
traffic from the committed PMC summary), the CPU baseline (the oracle's C
restatement of the reference CPU path, single core, bounded sample), the
batched AMR hierarchy leg and the large excised leg.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "time to 1e-10 rel. residual per 257^3 solve; learned+FP32 vs plain FP64"
HBM_PEAK_GBS = 8000.0


# tile variants of the fused sweep: (TX, TY, threads, u planes in flight, LDS ring planes)
SWEEP_SHAPES = {0: (64, 8, 256, 1, 4), 1: (64, 16, 512, 1, 4), 2: (32, 16, 512, 1, 4), 3: (128, 8, 512, 1, 4),
                4: (64, 32, 512, 1, 4), 5: (128, 16, 1024, 1, 4), 6: (128, 16, 1024, 2, 4), 7: (64, 16, 512, 2, 4),
                8: (128, 8, 512, 2, 4), 9: (128, 16, 1024, 2, 5), 10: (128, 16, 1024, 1, 5),
                11: (128, 8, 512, 1, 5)}


def measured_traffic(n, variant, prec="double"):
    """HBM bytes per launch of the finest-level sweep from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (copy-kernel calibrated), when
    it covers this grid size and tile shape; else None."""
    path = os.path.join(ROOT, "profiles", f"traffic_finest_{n}.json")
    if not os.path.exists(path) or variant not in SWEEP_SHAPES:
        return None
    tx, ty, nt, pf, ring = SWEEP_SHAPES[variant]
    name = f"pmg::k_sweep<{prec}, 0, false, true, true, {tx}, {ty}, {nt}, {pf}, {ring}>"
    with open(path) as f:
        entry = json.load(f).get("kernels", {}).get(name)
    return None if entry is None else entry["traffic_bytes"]


def measured_resid_restrict_traffic(n, name):
    """HBM bytes per launch of the fused finest residual + injection from the
    committed rocprofv3 summary, or None."""
    path = os.path.join(ROOT, "profiles", f"traffic_finest_{n}.json")
    if not os.path.exists(path):
        return None
    kern = {"f64": "pmg::k_resid_restrict_s<double, 0, true, 16, 8, 256, 1>",
            "f32": "pmg::k_resid_restrict_s<float, 0, true, 32, 8, 256, 1>"}[name]
    with open(path) as f:
        entry = json.load(f).get("kernels", {}).get(kern)
    return None if entry is None else entry["traffic_bytes"]


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(rho, n, cycles_needed, sample_cycles):
    """Oracle (C restatement of amrd/mg.c + m_util_r8 + nps hooks, -O2, one
    thread) on the same input: hierarchy build + `sample_cycles` V-cycles
    timed, extrapolated to the cycle count the FP64 solve needs."""
    import oracle as O
    t0 = time.perf_counter()
    m = O.UniformMG(n)
    m.set_params(pre=2, pst=2, max_citer=5000, crtol=1e-3)
    m.field(m.nlev, O.F_RHO)[:] = rho
    m.init()
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(sample_cycles):
        m.vcycle()
    t_cyc = (time.perf_counter() - t0) / sample_cycles
    del m
    ms = 1e3 * (t_build + cycles_needed * t_cyc)
    return {
        "value": round(ms, 2),
        "unit": "ms/solve",
        "cores": 1,
        "kind": "port",
        "sample": (f"{n}^3: MG hierarchy build + {sample_cycles} V(2,2) cycles timed on 1 core "
                   f"({1e3 * t_cyc:.0f} ms/cycle), extrapolated to the {cycles_needed} cycles the "
                   f"FP64 solve needs; host {cpu_model()}, nproc {os.cpu_count()}"),
    }


def sequence_leg(args, pm, torch):
    """Learned + FP32 on the drifting sequence of 257^3 solves.

    Training (not timed): (residual, error) crops from plain FP64 V-cycles of
    solves 0 .. T-1, Adam on the GPU.  Evaluation: solves T .. T+K-1, each
    solved in seven ways from its own initial guess -- learned + FP32 (mode
    2), FP32 inner cycles (mode 1), flexible GMRES(8) preconditioned by the
    FP32 cycle and plain FP64 from the history projection warm start (the last
    --seq-hist solves), FP32 and plain FP64 from second-order extrapolation of
    the two previous solutions, and plain FP64 from u = 0 (the reference's
    cold solve).  Timed per solve: computing and setting the initial guess,
    and the solve to ||r|| <= rtol ||r(u=0)||."""
    from pamr_amrd_amd import sequence as SQ
    n = args.n
    T, K = args.seq_train, args.steps
    log = lambda m: print(f"bench: {m}", file=sys.stderr, flush=True)
    seq = SQ.Sequence(n=n, seed=1234, speed=0.002, hist_k=args.seq_hist)
    t0 = time.perf_counter()
    X, Y = seq.collect(nsolves=T, pairs=args.seq_pairs, side=16, cycles=8)
    t_collect = time.perf_counter() - t0
    log(f"sequence: {X.shape[0]} training pairs from {T} solves in {t_collect:.1f} s")
    net = pm.Net(seed=3)

    def fixed_crops_loss():
        # the loss on the same 256 training crops before and after training
        # (single minibatch losses are too noisy to compare)
        tot = 0.0
        for q in range(8):
            xb, yb = X[32 * q:32 * (q + 1)].contiguous(), Y[32 * q:32 * (q + 1)].contiguous()
            tot += net.loss_grad(xb, yb, (16, 16, 16), 32, margin=3)[0] / 8
        return tot

    loss_before = fixed_crops_loss()
    t0 = time.perf_counter()
    losses = SQ.train(net, X, Y, side=16, steps=args.seq_train_steps, batch=32, lr=2e-3, margin=3)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    loss_after = fixed_crops_loss()
    log(f"sequence: trained {args.seq_train_steps} Adam steps in {t_train:.1f} s, "
        f"loss {sum(losses[:20]) / 20:.3e} -> {sum(losses[-20:]) / 20:.3e}")
    S = seq.S
    S.set_net(net)
    # mode -1: flexible GMRES(8) with the FP32 V-cycle as preconditioner
    modes = (("learned_fp32", 2, "proj"), ("fp32", 1, "proj"), ("fgmres_fp32", -1, "proj"),
             ("fp64_warm", 0, "proj"), ("fp32_extrap2", 1, "extrap2"), ("fp64_extrap2", 0, "extrap2"),
             ("fp64_cold", 0, "zero"))
    times = {m[0]: [] for m in modes}
    iters = {m[0]: [] for m in modes}
    misfit = []
    # untimed warm-up (graph capture of every mode): the last training source
    # solved from u = 0 in each mode, the history left as it is
    rref = seq.load_source(T - 1)
    for _ in range(max(1, args.warmup)):
        for name, mode, method in modes:
            seq.warm_start("zero", misfit=False)
            if mode < 0:
                seq.S.solve_fgmres(8, args.rtol, 100, r_ref=rref)
            else:
                seq.solve(mode, args.rtol, rref)
    torch.cuda.synchronize()
    for s in range(T, T + K):
        rref = seq.load_source(s)
        for name, mode, method in modes:
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            seq.warm_start(method, misfit=False)
            if mode < 0:
                _, h = seq.S.solve_fgmres(8, args.rtol, 100, r_ref=rref)
            else:
                _, h = seq.solve(mode, args.rtol, rref)
            torch.cuda.synchronize()
            dt = 1e3 * (time.perf_counter() - t1)
            # no iteration: the initial guess already met the test
            if len(h) and h[-1] > args.rtol * rref:
                raise RuntimeError(f"sequence solve {s} mode {name} did not converge: {h}")
            times[name].append(dt)
            iters[name].append(len(h))
        if seq.nhist:  # misfit of the projection (untimed)
            seq.warm_start("proj", misfit=True)
            misfit.append(seq.last_misfit)
            _, h = seq.solve(0, args.rtol, rref)
        seq.push_solution()
    mean = {k: sum(v) / len(v) for k, v in times.items()}
    # the projection warm start alone (its cost is inside every timed solve above)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(10):
        seq.warm_start("proj", misfit=False)
    torch.cuda.synchronize()
    ms_ws = 1e3 * (time.perf_counter() - t1) / 10
    return {
        "ms_per_solve": {k: round(v, 4) for k, v in mean.items()},
        "iterations": {k: round(sum(v) / len(v), 2) for k, v in iters.items()},
        "speedup_learned_fp32_vs_fp64_cold": round(mean["fp64_cold"] / mean["learned_fp32"], 3),
        "speedup_learned_fp32_vs_fp64_warm": round(mean["fp64_warm"] / mean["learned_fp32"], 3),
        "speedup_learned_fp32_vs_fp64_extrap2": round(mean["fp64_extrap2"] / mean["learned_fp32"], 3),
        "speedup_fgmres_fp32_vs_fp64_cold": round(mean["fp64_cold"] / mean["fgmres_fp32"], 3),
        "warm_start": {"method": f"history projection over the last {seq.hist_k} solves",
                       "mean_relative_misfit": float(f"{sum(misfit) / len(misfit):.4e}"),
                       "max_relative_misfit": float(f"{max(misfit):.4e}"),
                       "ms_per_warm_start": round(ms_ws, 4)},
        "network": S.net_stats(),
        "training": {"solves": T, "pairs": int(X.shape[0]), "crop": 16, "adam_steps": args.seq_train_steps,
                     "collect_s": round(t_collect, 2), "train_s": round(t_train, 2),
                     "loss_first": sum(losses[:20]) / 20, "loss_last": sum(losses[-20:]) / 20,
                     "loss_256_crops_before": loss_before, "loss_256_crops_after": loss_after},
        "solves_timed": [T, T + K - 1],
        "ms_per_solve_first_10": round(sum(times["learned_fp32"][:10]) / min(10, K), 4),
        "ms_per_solve_max": round(max(times["learned_fp32"]), 4),
    }


def amr_leg(args, pm, torch):
    """Batched AMR hierarchy: base 129^3, refinement ratio 2, two refined
    levels of 24-48 seeded boxes with odd sides in 17..97 clipped to their
    parents, div(a grad u) + c u (a lognormal, sigma 0.5, correlation 0.1;
    c uniform in [0,1]), FP64 V(2,2) cycles, one launch per level operation,
    each cycle one hipGraph."""
    import numpy as np
    from pamr_amrd_amd import hierarchy as H
    levels = H.random_boxes(base_n=129, nref=2, nmin=24, nmax=48, smin=17, smax=97, seed=7)
    A = pm.Amr(129, levels, op=1)
    npts = {}
    for L in range(A.base, A.nlev + 1):
        tot = 0
        for g in range(A.ngrids(L)):
            shape, bb = A.grid(L, g)
            A.upload(L, g, pm.PMG_F_RHO, -H.gaussian_on(shape, bb))
            A.upload(L, g, pm.PMG_F_A, H.lognormal_on_device(shape, bb, torch).cpu().numpy())
            A.upload(L, g, pm.PMG_F_C, H.uniform_on(shape, bb))
            tot += int(np.prod(shape))
        npts[L] = tot
    A.set_params(pre=2, pst=2, max_citer=5000, crtol=1e-3)
    A.init()
    tuned = {}
    if not args.amr_no_tune:
        from pamr_amrd_amd import autotune as AT
        tuned = AT.autotune_amr(A, nrep=5)
    hist = []
    for _ in range(2):
        hist.append(A.vcycle()[0])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ncyc = args.amr_cycles
    for _ in range(ncyc):
        hist.append(A.vcycle()[0])
    torch.cuda.synchronize()
    ms_cycle = 1e3 * (time.perf_counter() - t0) / ncyc
    # FP32 inner cycles in FP64 defect correction (defect after u is injected
    # down the levels, FP32 V-cycle on the error, u += e): a fixed number of
    # iterations from the current u
    A.set_mixed_switch(1.0)
    A.solve_rel(1, 0.0, 2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, hmix = A.solve_rel(1, 0.0, ncyc)
    torch.cuda.synchronize()
    ms_mixed = 1e3 * (time.perf_counter() - t0) / len(hmix)
    # whole solves from u = 0 to ||r|| <= rtol ||r0||: FP64 V-cycles, and
    # mixed precision with FP64 cycles down to 1e-2 ||r0|| (the default)
    solves = {}
    for name, mode, switch in (("fp64", 0, 1.0), ("mixed", 1, 1e-2)):
        for L in range(A.base, A.nlev + 1):
            for g in range(A.ngrids(L)):
                A.upload(L, g, pm.PMG_F_U, np.zeros(A.grid(L, g)[0]))
        A.init()
        A.set_mixed_switch(switch)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r0, h = A.solve_rel(mode, args.rtol, 400)
        torch.cuda.synchronize()
        dt = 1e3 * (time.perf_counter() - t0)
        if len(h) == 0 or h[-1] > args.rtol * r0:
            raise RuntimeError(f"amr {name} solve did not reach {args.rtol}: {len(h)} iterations")
        solves[name] = {"ms": round(dt, 2), "iterations": len(h), "final_rel_residual": float(h[-1] / r0)}
    Lf = A.nlev
    t_sw = A.time_sweep(Lf, 10)
    interior = sum(int(np.prod([s - 2 for s in A.grid(Lf, g)[0]])) for g in range(A.ngrids(Lf)))
    # finest AMR level sweep: read u, rho, a, c, write u (rhs = 0 there)
    bytes_sw = 40 * interior
    stats = {L: A.plan_stats(L) for L in range(1, A.nlev + 1)}
    out = {
        "workload": "amr129_2lev_24to48boxes_varcoef",
        "levels": A.nlev, "grids_per_level": [A.ngrids(L) for L in range(1, A.nlev + 1)],
        "points_per_amr_level": npts,
        "ms_per_vcycle": round(ms_cycle, 4), "vcycles_per_s": round(1e3 / ms_cycle, 1),
        "ms_per_mixed_iteration": round(ms_mixed, 4),
        "solve_to_rtol": solves, "rtol": args.rtol,
        "residual_history": [float(f"{h:.6e}") for h in hist],
        "finest_batched_sweep": {"grids": A.ngrids(Lf), "interior_points": interior, "avg_ms": round(t_sw, 5),
                                 "algorithmic_bytes": bytes_sw,
                                 "achieved_GBps": round(bytes_sw / (t_sw * 1e-3) / 1e9, 1)},
        "transfer_boxes_per_level": {L: {"sync0": v[0], "sync1": v[1], "inject": v[2], "interp": v[3]}
                                     for L, v in stats.items()},
        "sweep_tiling_per_level": {L: {"variant": v, "zchunk": z, "tuned_ms": round(ms, 5),
                                       "default_ms": round(ms0, 5)} for L, (v, z, ms, ms0) in tuned.items()},
        "dtype": "f64",
    }
    A.close()
    return out


def excised_leg(args, pm, torch):
    """Large excised problem: div(a grad u) + c u on n^3 (default 513^3), a
    lognormal, c uniform in [0,1], 3 seeded excised spheres (chr on every
    level, Dirichlet data g held on the excised points), 16 seeded Gaussian
    right-hand sides, each solved to ||r|| <= rtol ||r(u0)||.

    Untimed set-up: the band plan of the learned correction (updatable points
    within 3 grid points of an excised point, in cores of 26^3) and its
    network, trained on (residual, error) crops of the plain mixed iteration
    of c5_band_train extra right-hand sides.  Timed per right-hand side:
    flexible GMRES(8) with the FP32 V-cycle as preconditioner, and FP32 inner
    V-cycles in FP64 defect correction with the learned band correction
    (pmg_solve_rel mode 3, every iteration checked against the plain one);
    the first c5_mixed_rhs right-hand sides also by plain defect correction
    and by FGMRES with the correction in its preconditioner, the first in
    plain FP64 V-cycles (capped).  Then all right-hand sides again,
    concurrently, by the faster of the two methods, spread over c5_streams
    solver instances (one stream and host thread each)."""
    import numpy as np
    from pamr_amrd_amd import band as BD
    from pamr_amrd_amd import hierarchy as H
    from pamr_amrd_amd import sequence as SQ
    log = lambda m: print(f"bench: excised: {m}", file=sys.stderr, flush=True)
    n = args.c5_n
    box = (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)
    spheres = H.excision_spheres(seed=21)
    shape = (n, n, n)
    a = H.lognormal_on_device(shape, box, torch)
    gen = torch.Generator(device="cuda").manual_seed(5)
    c = torch.rand(shape, generator=gen, dtype=torch.float64, device="cuda")
    chrs = {}

    def make_solver():
        S = pm.Solver(n, op=1, do_ex=1, ex=1.0)
        S.set_params(pre=2, pst=2, max_citer=5000, crtol=1e-3, rop_u=3, rop_res=3, rop_src=3)
        for lev in range(1, S.nlev + 1):
            if lev not in chrs:
                chrs[lev] = H.chr_on(S.shape(lev), box, spheres)
            S.set_chr(lev, chrs[lev])
        SQ._copy_grid(S.field_ptr(S.nlev, pm.PMG_F_A), pm.tptr(a), n)
        SQ._copy_grid(S.field_ptr(S.nlev, pm.PMG_F_C), pm.tptr(c), n)
        return S

    S = make_solver()
    Lf = S.nlev
    ch = chrs[Lf]
    excised = int((ch == 1.0).sum())
    u0 = torch.from_numpy(np.where(ch == 1.0, H.excision_data(shape, box), 0.0)).cuda()
    del ch
    L = pm.lib()

    # ---- learned band correction: plan and training (not timed) ----
    t0 = time.perf_counter()
    ncore = S.band_setup(args.c5_band_depth, 26)
    npts, _, side = S.band_info()
    t_plan = time.perf_counter() - t0
    log(f"band: {npts} points in {ncore} cores of 26^3 ({t_plan:.1f} s)")
    t0 = time.perf_counter()
    train_rhos = [SQ.DriftingSource(seed=900 + k, speed=0.0).rho(n, 0, torch) for k in range(args.c5_band_train)]
    ustars = [BD.converged_solution(S, r, u0, torch, rtol=args.rtol) for r in train_rhos]
    t_sol = time.perf_counter() - t0
    # two placements, each with its own network: after the V-cycle (on the
    # FP64 defect) and inside the FP32 V-cycle after the finest coarse-grid
    # correction
    t0 = time.perf_counter()
    x, t, w = BD.collect_pairs(S, train_rhos, ustars, u0, torch, iters=args.c5_band_iters,
                               crops=args.c5_band_crops)
    t_collect = time.perf_counter() - t0
    net, ts, tinfo = BD.train(x, t, w, torch, steps=args.c5_band_steps, log=log)
    del x, t, w
    t0 = time.perf_counter()
    x, t, w = BD.collect_pairs_inner(S, train_rhos, ustars, u0, torch, iters=args.c5_band_iters,
                                     crops=args.c5_band_crops)
    t_collect_in = time.perf_counter() - t0
    net_in, ts_in, tinfo_in = BD.train(x, t, w, torch, steps=args.c5_band_steps, log=log)
    del x, t, w, ustars, train_rhos
    torch.cuda.empty_cache()
    tinfo["collect_s"] = round(t_collect, 2)
    tinfo["solutions_s"] = round(t_sol, 2)
    tinfo_in["collect_s"] = round(t_collect_in, 2)
    S.set_band_net(net, ts, use_dc=True, use_fgmres=False)
    S.set_band_net_inner(net_in, ts_in)
    # the network's inference on the crops of one correction, timed alone
    xin = torch.randn((ncore, side, side, side), dtype=torch.float32, device="cuda")
    yout = torch.empty_like(xin)
    net.infer(xin, yout, (side, side, side), ncore)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(10):
        net.infer(xin, yout, (side, side, side), ncore)
    e1.record()
    e1.synchronize()
    t_inf = e0.elapsed_time(e1) / 10
    del xin, yout
    flops = 15552.0 * ncore * side ** 3
    roof_net = {"bound": "mfma", "achieved": round(flops / (t_inf * 1e-3) / 1e12, 2), "peak": 2500.0,
                "unit": "TFLOP/s", "frac": round(flops / (t_inf * 1e-3) / 1e12 / 2500.0, 4),
                "flop_per_application": flops, "ms_per_application": round(t_inf, 4),
                "crop_points": ncore * side ** 3, "band_points": npts,
                "note": "3x3x3 convs 1->16->16->1 over the crops of one correction (15552 flop per crop point), "
                        "BF16 MFMA with FP32 accumulation; pmg_net_infer timed with HIP events"}

    times = {"fgmres": [], "dc_band": [], "dc_band_in": [], "fgmres_band_in": [], "dc": [], "fgmres_band": [],
             "fgmres2": [], "fgmres2_band": []}
    iters = {k: [] for k in times}
    plain = None
    for r in range(args.c5_rhs):
        src = SQ.DriftingSource(seed=1234 + r, speed=0.0)
        rho = src.rho(n, 0, torch)
        SQ._copy_grid(S.field_ptr(Lf, pm.PMG_F_RHO), pm.tptr(rho), n)
        del rho
        modes = ["fgmres", "dc_band"]
        if r < args.c5_mixed_rhs:
            modes += ["dc", "fgmres_band", "dc_band_in", "fgmres_band_in", "fgmres2", "fgmres2_band"]
        if r == 0:
            modes += ["fp64"]
        for mode in modes:
            S.band_set_inner(mode.endswith("_in"))
            S.band_enable(True, 2 if mode.startswith(("fgmres_band", "fgmres2_band")) else False)
            # fgmres2*: two FP32 V-cycles per preconditioner application, the band
            # correction (when on) between them
            S.set_fgmres_prec_cycles(2 if mode.startswith("fgmres2") else 1)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            L.pmg_set_solution_lincomb(S.h, pm.tptr(u0), None, 1.0, 0.0)
            S.init()
            if mode.startswith("fgmres"):
                r0, h = S.solve_fgmres(args.c5_restart, args.rtol, args.c5_max_iter)
            elif mode == "fp64":
                r0, h = S.solve_rel(0, args.rtol, args.c5_plain_cycles)
            else:
                r0, h = S.solve_rel(3 if mode.startswith("dc_band") else 1, args.rtol, args.c5_max_iter)
            torch.cuda.synchronize()
            dt = 1e3 * (time.perf_counter() - t0)
            if mode == "fp64":
                # the reference's excision transfers make plain FP64 V-cycles
                # contract slowly here; report where the capped run ended
                plain = {"ms": round(dt, 2), "cycles": len(h), "final_rel_residual": float(h[-1] / r0),
                         "converged": bool(h[-1] <= args.rtol * r0),
                         "mean_contraction": float((h[-1] / r0) ** (1.0 / len(h)))}
                continue
            if len(h) == 0 or h[-1] > args.rtol * r0:
                raise RuntimeError(f"excised rhs {r} ({mode}) did not converge in {len(h)} steps: {h[-3:]}")
            times[mode].append(dt)
            iters[mode].append(len(h))
        log(f"rhs {r}: " + ", ".join(f"{m} {times[m][-1]:.1f} ms/{iters[m][-1]}" for m in modes if m != "fp64"))
    S.band_enable(True, False)
    S.band_set_inner(False)
    S.set_fgmres_prec_cycles(1)
    mean = {k: (sum(v) / len(v) if v else None) for k, v in times.items()}
    best = min(("fgmres", "dc_band"), key=lambda k: mean[k])
    # the same right-hand sides concurrently by the faster method: one solver
    # per stream, one host thread each, every solve the single-RHS solve
    conc = None
    if args.c5_streams > 1:
        from pamr_amrd_amd import multirhs as MR
        solvers = [S] + [make_solver() for _ in range(args.c5_streams - 1)]
        nets = []
        inner = best.endswith("_in")
        S.band_set_inner(inner)
        S.band_enable(True, 2 if best.startswith("fgmres_band") else False)
        if best != "fgmres":
            src = net_in if inner else net
            wts = src.get_weights()
            for X in solvers[1:]:
                X.band_setup(args.c5_band_depth, 26)
                nx_ = pm.Net(seed=1)
                nx_.set_weights(wts)
                if inner:
                    X.set_band_net_inner(nx_, ts_in)
                    X.set_band_net(None, 1.0, use_dc=True, use_fgmres=False)
                else:
                    X.set_band_net(nx_, ts, use_dc=True, use_fgmres=False)
                X.band_set_inner(inner)
                X.band_enable(True, 2 if best.startswith("fgmres_band") else False)
                nets.append(nx_)
        rhs = [SQ.DriftingSource(seed=1234 + r, speed=0.0).rho(n, 0, torch) for r in range(args.c5_rhs)]

        def solve(X):
            if best.startswith("dc_band"):
                return X.solve_rel(3, args.rtol, args.c5_max_iter)
            return X.solve_fgmres(args.c5_restart, args.rtol, args.c5_max_iter)

        for X in solvers:  # allocations (FP32 shadows, Krylov basis, graphs) before the timed region
            L.pmg_set_solution_lincomb(X.h, pm.tptr(u0), None, 1.0, 0.0)
            X.init()
            if best.startswith("dc_band"):
                X.solve_rel(3, args.rtol, 3)
            else:
                X.solve_fgmres(args.c5_restart, args.rtol, 2)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = MR.solve_concurrent(solvers, [pm.tptr(r) for r in rhs], pm.tptr(u0), solve, args.rtol)
        torch.cuda.synchronize()
        dt = 1e3 * (time.perf_counter() - t0)
        conc = {"method": best, "streams": len(solvers), "rhs": len(rhs), "total_ms": round(dt, 1),
                "ms_per_rhs": round(dt / len(rhs), 2), "iterations": [len(h) for _, h in res]}
        del rhs
        for X in solvers[1:]:
            X.close()
        for nx_ in nets:
            nx_.close()
    bstats = S.band_stats()
    S.close()
    net.close()
    net_in.close()
    del a, c, u0
    rd = lambda v: None if v is None else round(v, 2)
    return {"workload": f"excised{n}_3spheres_varcoef_16rhs", "n": n, "excised_points": excised,
            "rhs": args.c5_rhs, "rtol": args.rtol, "reference_norm": "||r(u0)||",
            "ms_per_rhs": {k: rd(v) for k, v in mean.items()},
            "iterations": {k: v for k, v in iters.items()},
            "speedup_dc_band_vs_dc": (round(mean["dc"] / mean["dc_band"], 3) if mean["dc"] else None),
            "speedup_dc_band_vs_fgmres": round(mean["fgmres"] / mean["dc_band"], 3),
            "fastest": best,
            "fgmres_restart": args.c5_restart,
            "band": {"depth": args.c5_band_depth, "core": 26, "crop": side, "cores": ncore, "points": npts,
                     "plan_s": round(t_plan, 2), "training": tinfo, "training_in_cycle": tinfo_in,
                     "stats": bstats,
                     "roofline_mfma": roof_net},
            "plain_fp64_first_rhs": plain, "concurrent": conc}


def probe_kernels(S, n, pm, torch):
    """Launch each finest-level kernel of interest 4 times, nothing else
    large: the counter runs attribute FETCH_SIZE / WRITE_SIZE per dispatch."""
    Lf = S.nlev
    for name, which, prec in (("sweep f64", 0, 0), ("residual f64", 1, 0), ("sweep f32", 0, 1),
                              ("resid_restrict f64", 2, 0), ("resid_restrict f32", 2, 1)):
        print(f"probe {name}: {S.time_kernel(which, Lf, prec, 4):.5f} ms", flush=True)
    # calibration: pmg_dmcopy3d over the whole grid = plain copy, 8 B read + 8 B written per point
    src = S.field_ptr(Lf, pm.PMG_F_RHO)
    dst = torch.empty((n, n, n), dtype=torch.float64, device="cuda")
    L = pm.lib()
    for _ in range(4):
        rc = L.pmg_dmcopy3d(src, pm.tptr(dst), n, n, n, n, n, n, 1, 1, 1, 1, 1, 1, n, n, n,
                            1, 1, 1, 1, 1, 1, None)
        if rc != 0:
            raise RuntimeError(f"pmg_dmcopy3d returned {rc}")
    torch.cuda.synchronize()
    print(f"probe copy: {8 * n ** 3} bytes read and written per launch", flush=True)


def measured_copy_gbps(S, n, pm, torch, reps=20):
    """HBM rate of a plain device copy of the n^3 FP64 grid (pmg_dmcopy3d,
    8 B read + 8 B written per point), timed with events on the stream it runs
    on: the bandwidth a streaming kernel can reach on this device."""
    src = S.field_ptr(S.nlev, pm.PMG_F_RHO)
    dst = torch.empty((n, n, n), dtype=torch.float64, device="cuda")
    L = pm.lib()

    def copy():
        if L.pmg_dmcopy3d(src, pm.tptr(dst), n, n, n, n, n, n, 1, 1, 1, 1, 1, 1, n, n, n,
                          1, 1, 1, 1, 1, 1, None) != 0:
            raise RuntimeError("pmg_dmcopy3d failed")
    for _ in range(3):
        copy()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        copy()
    e1.record()
    e1.synchronize()
    return 16.0 * n ** 3 / (e0.elapsed_time(e1) / reps * 1e-3) / 1e9


def probe_net(n, pm, torch):
    """Eight inference passes of the correction network over n^3 points
    (random weights and input; the cost does not depend on them), for a
    kernel trace of the inference kernels."""
    net = pm.Net(seed=1)
    x = torch.randn((n, n, n), dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    net.infer(x, y, (n, n, n), 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(8):
        net.infer(x, y, (n, n, n), 1)
    torch.cuda.synchronize()
    print(f"probe net: {1e3 * (time.perf_counter() - t0) / 8:.3f} ms per {n}^3 inference", flush=True)
    net.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=150,
                    help="timed solves: of the drifting sequence (solves 50..49+K) and of the single solve")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=257)
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-cycles", type=int, default=3)
    ap.add_argument("--variant", type=int, default=-1, help="sweep tile variant (-1: tuned)")
    ap.add_argument("--zchunk", type=int, default=-1)
    ap.add_argument("--seq-train", type=int, default=50, help="training solves of the sequence")
    ap.add_argument("--seq-pairs", type=int, default=4000)
    ap.add_argument("--seq-train-steps", type=int, default=1500)
    ap.add_argument("--seq-hist", type=int, default=8, help="solves kept for the history warm start (<= 8)")
    ap.add_argument("--no-seq", action="store_true", help="skip the learned sequence leg")
    ap.add_argument("--no-amr", action="store_true", help="skip the batched AMR hierarchy leg")
    ap.add_argument("--amr-cycles", type=int, default=10)
    ap.add_argument("--amr-no-tune", action="store_true", help="keep the default AMR sweep tiling")
    ap.add_argument("--no-excised", action="store_true", help="skip the large excised leg")
    ap.add_argument("--c5-n", type=int, default=513)
    ap.add_argument("--c5-rhs", type=int, default=16)
    ap.add_argument("--c5-plain-cycles", type=int, default=60)
    ap.add_argument("--c5-max-iter", type=int, default=300)
    ap.add_argument("--c5-restart", type=int, default=8, help="FGMRES restart length of the excised leg")
    ap.add_argument("--c5-mixed-rhs", type=int, default=2, help="right-hand sides also solved by defect correction")
    ap.add_argument("--c5-band-train", type=int, default=2,
                    help="extra right-hand sides whose plain iterations train the band correction")
    ap.add_argument("--c5-band-steps", type=int, default=4000, help="Adam steps of the band correction")
    ap.add_argument("--c5-band-iters", type=int, default=10, help="plain iterations sampled per training RHS")
    ap.add_argument("--c5-band-crops", type=int, default=48, help="training crops per sampled iteration")
    ap.add_argument("--c5-band-depth", type=int, default=3, help="band: points within this distance of the excision")
    ap.add_argument("--c5-streams", type=int, default=4,
                    help="solver instances (streams, host threads) of the concurrent multi-RHS run; 1 skips it")
    ap.add_argument("--probe-net", action="store_true",
                    help="only time the correction network's inference on the n^3 grid (for a kernel trace)")
    ap.add_argument("--probe-kernels", action="store_true",
                    help="only launch the finest-level kernels (for rocprofv3 --pmc runs): "
                         "FP64 sweep, FP64 residual, FP32 sweep and a plain 8 B/point copy "
                         "used to calibrate the counters, 4 launches each")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus N with N > 1 is not supported "
              "(the multi-GPU sharding is a design, not an implementation).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import pamr_amrd_amd as pm

    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    n = args.n
    if args.probe_net:
        probe_net(n, pm, torch)
        return
    print(f"bench: generating {n}^3 source", file=sys.stderr, flush=True)
    rho = -pm.gaussian_source(n, seed=1234)
    S = pm.Solver(n)
    S.set_params(pre=2, pst=2, max_citer=5000, crtol=1e-3, rop_u=3, rop_res=3, rop_src=3)
    variant, zchunk = args.variant, args.zchunk
    if variant >= 0 and zchunk >= 0:
        S.set_level_sweep_config(S.nlev, variant, zchunk)
        variant32, zchunk32 = variant, zchunk
    else:
        variant, zchunk = pm.tuned_sweep_config(n)
        variant32, zchunk32 = pm.tuned_sweep_config(n, "f32")
    S.upload(S.nlev, pm.PMG_F_RHO, rho)
    S.init()
    Lf = S.nlev
    if args.probe_kernels:
        probe_kernels(S, n, pm, torch)
        S.close()
        return

    def solve(mode):
        S.zero_solution()
        r0, hist = S.solve_rel(mode, args.rtol, 100)
        if len(hist) == 0 or hist[-1] > args.rtol * r0:
            raise RuntimeError(f"mode {mode} did not reach rtol: r0={r0} hist={hist}")
        return r0, hist

    print("bench: warmup", file=sys.stderr, flush=True)
    for _ in range(max(1, args.warmup)):
        r0p, hp = solve(0)
        r0m, hm = solve(1)

    def timed(mode):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            r0, h = solve(mode)
        torch.cuda.synchronize()
        return 1e3 * (time.perf_counter() - t0) / args.steps, r0, h

    print("bench: timing plain FP64", file=sys.stderr, flush=True)
    ms_plain, r0p, hp = timed(0)
    print("bench: timing mixed FP32/FP64", file=sys.stderr, flush=True)
    ms_mixed, r0m, hm = timed(1)
    # the mixed solution's error against a 40-cycle FP64 solve of the same problem
    u_mixed = S.download(Lf, pm.PMG_F_U)
    S.zero_solution()
    S.init()
    S.solve(40, 40)
    u_ref40 = S.download(Lf, pm.PMG_F_U)
    err_mixed = float(np.abs(u_mixed - u_ref40).max() / np.abs(u_ref40).max())

    # dominant kernel: finest-level fused red-black sweep (rhs == 0 there):
    # reads u and rho, writes u -> 24 B (FP64) / 12 B (FP32) per interior point
    npts = (n - 2) ** 3
    # the FP64 sweep as it runs in the solve: HIP events around the finest
    # level's sweep launches of 20 ungraphed V-cycles (80 launches); the
    # same kernel back to back on one grid (40 launches) for comparison
    t64 = S.time_kernel(3, Lf, 0, 20)
    t64_b2b = S.time_kernel(0, Lf, 0, 40)
    t32 = S.time_kernel(0, Lf, 1, 40)
    bytes64 = 24 * npts
    bytes32 = 12 * npts
    ach64 = bytes64 / (t64 * 1e-3) / 1e9
    ach32 = bytes32 / (t32 * 1e-3) / 1e9
    copy_gbps = measured_copy_gbps(S, n, pm, torch)
    roof = {"bound": "hbm", "achieved": round(ach64, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach64 / HBM_PEAK_GBS, 4), "traffic": measured_traffic(n, variant),
            "kernel": "k_sweep<double> finest level (fused red-black sweep, 255^3 interior)",
            "algorithmic_bytes": bytes64, "avg_ms": round(t64, 5),
            "timing": "HIP events around the finest sweeps of 20 ungraphed V(2,2) cycles (80 launches)",
            "avg_ms_back_to_back": round(t64_b2b, 5),
            "copy_GBps": round(copy_gbps, 1), "frac_of_copy": round(ach64 / copy_gbps, 4)}
    roof32 = {"bound": "hbm", "achieved": round(ach32, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
              "frac": round(ach32 / HBM_PEAK_GBS, 4),
              "traffic": measured_traffic(n, variant32, "float"), "sweep_variant": variant32, "zchunk": zchunk32,
              "algorithmic_bytes": bytes32,
              "avg_ms": round(t32, 5)}
    # the finest residual fused with the full-weight injection of u and f_res:
    # reads u and rho at every fine point, writes the coarse u and f_res
    nc = (n - 1) // 2 + 1
    roof_rr = {}
    for prec, name, b in ((0, "f64", 16), (1, "f32", 8)):
        t = S.time_kernel(2, Lf, prec, 40)
        by = b * n ** 3 + b * nc ** 3
        ach = by / (t * 1e-3) / 1e9
        roof_rr[name] = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": measured_resid_restrict_traffic(n, name),
                         "algorithmic_bytes": by, "avg_ms": round(t, 5)}
    cpu = None
    if not args.no_cpu:
        print("bench: CPU baseline (oracle, 1 core)", file=sys.stderr, flush=True)
        cpu = cpu_baseline(rho, n, len(hp), args.cpu_cycles)
    single = {"workload": f"poisson{n}_gauss8_from_zero", "plain_fp64_ms_per_solve": round(ms_plain, 4),
              "mixed_fp32_fp64_ms_per_solve": round(ms_mixed, 4),
              "speedup_mixed_vs_plain_fp64": round(ms_plain / ms_mixed, 3),
              "plain_cycles": len(hp), "mixed_outer_iters": len(hm),
              "final_rel_residual": {"plain": hp[-1] / r0p, "mixed": hm[-1] / r0m},
              "mixed_error_vs_fp64_40_cycles": err_mixed}
    amrres = None
    if not args.no_amr:
        print("bench: batched AMR hierarchy leg", file=sys.stderr, flush=True)
        amrres = amr_leg(args, pm, torch)
    exres = None
    if not args.no_excised:
        print("bench: large excised leg", file=sys.stderr, flush=True)
        exres = excised_leg(args, pm, torch)
    seqres = None
    if not args.no_seq:
        print("bench: learned sequence leg", file=sys.stderr, flush=True)
        seqres = sequence_leg(args, pm, torch)
    if seqres is not None:
        value = seqres["ms_per_solve"]["learned_fp32"]
        config = {"workload": f"poisson{n}_drift8_sequence_learned_fp32", "n": n, "levels": Lf,
                  "cycle": "FAS V(2,2) RB-GS, FW restriction, trilinear prolongation",
                  "rtol": args.rtol, "reference_norm": "||r(u=0)||",
                  "sequence": f"8 Gaussians drifting 0.002/solve; train on solves 0-{args.seq_train - 1}, "
                              f"timed solves {seqres['solves_timed'][0]}-{seqres['solves_timed'][1]}",
                  "warm_start": f"history projection over the last {args.seq_hist} solves (timed)",
                  "sweep_variant": variant, "zchunk": zchunk}
        dtype = "f32 inner / f64 outer; network bf16 (fp32 accumulate)"
    else:
        value = ms_mixed
        config = {"workload": f"poisson{n}_gauss8_from_zero", "n": n, "levels": Lf,
                  "cycle": "FAS V(2,2) RB-GS, FW restriction, trilinear prolongation",
                  "rtol": args.rtol, "sweep_variant": variant, "zchunk": zchunk}
        dtype = "f32 inner / f64 outer"
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "ms/solve",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(value, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic: seeded 8-Gaussian sources (splitmix64 seed 1234)",
        "config": config,
        "solves_per_s": round(1e3 / value, 2),
        "sequence": seqres,
        "single_solve": single,
        "amr_hierarchy": amrres,
        "excised": exres,
        "roofline": roof,
        "roofline_fp32_sweep": roof32,
        "roofline_resid_restrict": roof_rr,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    S.close()


if __name__ == "__main__":
    main()
