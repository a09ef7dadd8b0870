"""Tile-shape autotuner for the fused red-black sweep.

For a grid size it times the finest-level sweep (HIP events on the solver's
stream, pmg_time_kernel) for every compiled tile variant (TX x TY x threads)
and z-chunk length, in FP64 and FP32, and stores the fastest configuration
in tuning.json next to this file.  bench.py and the solvers read the table
through tuned_sweep_config(); sizes that were never tuned get the default.
"""
from __future__ import annotations

import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning.json")
DEFAULT = (1, 32)
VARIANTS = {0: (64, 8, 256), 1: (64, 16, 512), 2: (32, 16, 512), 3: (128, 8, 512), 4: (64, 32, 512),
            5: (128, 16, 1024), 6: (128, 16, 1024, "2 planes in flight"), 7: (64, 16, 512, "2 planes in flight"),
            8: (128, 8, 512, "2 planes in flight"), 9: (128, 16, 1024, "2 planes in flight, 5-plane ring"),
            10: (128, 16, 1024, "5-plane ring"), 11: (128, 8, 512, "5-plane ring")}


def load_table():
    if not os.path.exists(TABLE):
        return {}
    with open(TABLE) as f:
        return json.load(f)


def table_key(n, op=0, do_ex=0):
    """Table key: the grid size, with the operator and excision when they are
    not the plain Poisson operator (the variable-coefficient sweep moves more
    bytes per point and may prefer another tile)."""
    return str(n) if (op == 0 and not do_ex) else f"{n}/op{op}{'ex' if do_ex else ''}"


def tuned_level_config(n, prec="f64", op=0, do_ex=0):
    """(variant, zchunk) tuned for a level of n^3 points (and this operator
    and excision when tuned for them), None when untuned."""
    tab = load_table()
    t = tab.get(table_key(n, op, do_ex), {}).get(prec) or tab.get(str(n), {}).get(prec)
    if not t:
        return None
    return int(t["variant"]), int(t["zchunk"])


def tuned_sweep_config(n, prec="f64"):
    return tuned_level_config(n, prec) or DEFAULT


def autotune(n, zchunks=(2, 4, 8, 16, 32, 64), nrep=20, write=True, verbose=True, solver=None):
    """Tune the finest level of an n^3 Poisson solver, or of `solver` (set up
    by the caller: operator, coefficients, excision; stored under its own
    table key)."""
    import numpy as np

    import pamr_amrd_amd as pm
    S = solver
    if S is None:
        S = pm.Solver(n)
        rho = -pm.gaussian_source(n, seed=1234) if n <= 129 else np.zeros((n, n, n))
        S.upload(S.nlev, pm.PMG_F_RHO, rho)
        S.init()
    key = table_key(n, getattr(S, "op", 0), getattr(S, "do_ex", 0))
    results = {}
    for prec, name in ((0, "f64"), (1, "f32")):
        best = None
        rows = []
        for v in range(pm.lib().pmg_num_sweep_variants()):
            for z in zchunks:
                if z > n - 2:
                    continue
                try:
                    S.set_level_sweep_config(S.nlev, v, z)
                    S.time_kernel(0, S.nlev, prec, 4)  # warm
                    ms = S.time_kernel(0, S.nlev, prec, nrep)
                except RuntimeError:
                    continue
                rows.append((v, z, ms))
                if verbose:
                    print(f"autotune n={n} {name} variant={v} {VARIANTS[v]} zchunk={z}: {ms:.4f} ms",
                          flush=True)
                if best is None or ms < best[2]:
                    best = (v, z, ms)
        results[name] = {"variant": best[0], "zchunk": best[1], "ms": best[2],
                         "all": [list(r) for r in rows]}
    if solver is None:
        S.close()
    else:
        S.apply_tuning()
    if write:
        tab = load_table()
        tab[key] = results
        with open(TABLE, "w") as f:
            json.dump(tab, f, indent=1)
    return results


AMR_VARIANTS = {0: (128, 16, 1024), 1: (64, 16, 512), 2: (32, 16, 512), 3: (32, 8, 256), 4: (16, 16, 256)}


def autotune_amr(A, levels=None, zchunks=(4, 8, 16, 32), nrep=10, verbose=False):
    """Per-level tile selection for the batched sweeps of an AMR hierarchy
    (``pamr_amrd_amd.Amr``): every compiled tile variant and z-chunk is timed
    with pmg_amr_time_sweep on the level's own patches and the fastest is
    kept.  Timing relaxes u, so each level's u is saved and restored around
    it.  Returns {level: (variant, zchunk, ms, default_ms)}."""
    import pamr_amrd_amd as pm
    nvar = pm.lib().pmg_amr_num_sweep_variants()
    levels = levels or range(2, A.nlev + 1)
    out = {}
    for L in levels:
        saved = [A.download(L, g, pm.PMG_F_U) for g in range(A.ngrids(L))]
        v0, z0 = A.sweep_config(L)
        A.time_sweep(L, 2)
        t0 = A.time_sweep(L, nrep)
        best = (v0, z0, t0)
        nmax = max(max(A.grid(L, g)[0]) for g in range(A.ngrids(L)))
        for v in range(nvar):
            for z in zchunks:
                if z > nmax - 2 or (v, z) == (v0, z0):
                    continue
                try:
                    A.set_level_sweep_config(L, v, z)
                except RuntimeError:
                    continue
                A.time_sweep(L, 2)
                ms = A.time_sweep(L, nrep)
                if verbose:
                    print(f"amr autotune level {L} variant={v} {AMR_VARIANTS.get(v)} zchunk={z}: {ms:.4f} ms",
                          flush=True)
                if ms < best[2]:
                    best = (v, z, ms)
        A.set_level_sweep_config(L, best[0], best[1])
        for g, u in enumerate(saved):
            A.upload(L, g, pm.PMG_F_U, u)
        out[L] = (best[0], best[1], best[2], t0)
    return out


if __name__ == "__main__":
    import argparse
    import shutil
    ap = argparse.ArgumentParser()
    ap.add_argument("sizes", nargs="*", type=int, default=[257])
    ap.add_argument("--copy-to", default=None, help="also copy tuning.json to this path")
    a = ap.parse_args()
    for n in a.sizes:
        autotune(n)
    if a.copy_to:
        shutil.copy(TABLE, a.copy_to)
