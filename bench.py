#!/usr/bin/env python3
"""bench.py -- time to a 1e-10 relative FP64 residual per 257^3 Poisson solve.

Workload (the 257^3 configuration of the north-star metric): -lap(u) = f on
the unit cube, vertex-centred 257^3 grid, u = 0 Dirichlet, f a seeded sum of
8 Gaussians, MG hierarchy 257 -> 5 (7 levels), FAS V(2,2) red-black
Gauss-Seidel, full-weight restriction, trilinear prolongation.

One step = one solve from u = 0 to ||r||/||r0|| <= 1e-10 (RMS norms), inputs
already resident in HBM.  Two solvers are timed:

* plain: FP64 V-cycles -- the reference method (amrd/mg.c) on the GPU;
* mixed: FP32 inner V-cycles inside FP64 defect correction (FP64 residual,
  update and test) -- the reported ``value``.

Extra fields: the plain time and the speedup, the roofline of the dominant
kernel (the finest-level fused red-black sweep, timed live with HIP events
on the solver's stream) and the CPU baseline (the oracle's C restatement of
the reference CPU path, single core, on a bounded sample).
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=257)
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-cycles", type=int, default=3)
    ap.add_argument("--variant", type=int, default=-1, help="sweep tile variant (-1: tuned)")
    ap.add_argument("--zchunk", type=int, default=-1)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus N with N > 1 is not supported "
              "(see DESIGN.md for how the solve would shard).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import pamr_amrd_amd as pm

    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    n = args.n
    print(f"bench: generating {n}^3 source", file=sys.stderr, flush=True)
    rho = -pm.gaussian_source(n, seed=1234)
    S = pm.Solver(n)
    S.set_params(pre=2, pst=2, max_citer=5000, crtol=1e-3, rop_u=3, rop_res=3, rop_src=3)
    variant, zchunk = args.variant, args.zchunk
    if variant >= 0 and zchunk >= 0:
        S.set_level_sweep_config(S.nlev, variant, zchunk)
    else:
        variant, zchunk = pm.tuned_sweep_config(n)
    S.upload(S.nlev, pm.PMG_F_RHO, rho)
    S.init()
    Lf = S.nlev

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

    # dominant kernel: finest-level fused red-black sweep (rhs == 0 there):
    # reads u and rho, writes u -> 24 B (FP64) / 12 B (FP32) per interior point
    npts = (n - 2) ** 3
    t64 = S.time_kernel(0, Lf, 0, 40)
    t32 = S.time_kernel(0, Lf, 1, 40)
    bytes64 = 24 * npts
    bytes32 = 12 * npts
    ach64 = bytes64 / (t64 * 1e-3) / 1e9
    ach32 = bytes32 / (t32 * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(ach64, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach64 / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": "k_sweep<double> finest level (fused red-black sweep, 255^3 interior)",
            "algorithmic_bytes": bytes64, "avg_ms": round(t64, 5)}
    roof32 = {"bound": "hbm", "achieved": round(ach32, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
              "frac": round(ach32 / HBM_PEAK_GBS, 4), "algorithmic_bytes": bytes32,
              "avg_ms": round(t32, 5)}
    cpu = None
    if not args.no_cpu:
        print("bench: CPU baseline (oracle, 1 core)", file=sys.stderr, flush=True)
        cpu = cpu_baseline(rho, n, len(hp), args.cpu_cycles)
    out = {
        "metric": METRIC,
        "value": round(ms_mixed, 4),
        "unit": "ms/solve",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_mixed, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 inner / f64 outer",
        "data": "synthetic: seeded 8-Gaussian source (splitmix64 seed 1234)",
        "config": {"workload": f"poisson{n}_gauss8_rtol1e-10", "n": n, "levels": Lf,
                   "cycle": "FAS V(2,2) RB-GS, FW restriction, trilinear prolongation",
                   "rtol": args.rtol, "sweep_variant": variant, "zchunk": zchunk},
        "plain_fp64_ms_per_solve": round(ms_plain, 4),
        "speedup_vs_plain_fp64": round(ms_plain / ms_mixed, 3),
        "plain_cycles": len(hp),
        "mixed_outer_iters": len(hm),
        "final_rel_residual": {"plain": hp[-1] / r0p, "mixed": hm[-1] / r0m},
        "solves_per_s": round(1e3 / ms_mixed, 2),
        "roofline": roof,
        "roofline_fp32_sweep": roof32,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    S.close()


if __name__ == "__main__":
    main()
