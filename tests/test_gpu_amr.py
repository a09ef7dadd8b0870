"""Batched multigrid over an AMR hierarchy on the GPU vs the multi-patch
oracle (oracle/amr_oracle.c, pinned in test_oracle_amr.py).

Bar: cmask rules, every transfer (sync with AMR boundary width 0 and 1,
inject with straight / half-weight / full-weight restriction, second-order
interp) and whole V-cycles bit for bit on every grid of every level; the
per-cycle norms within 1e-12 relative (they differ only in summation order).
"""
import numpy as np
import pytest

import oracle as O
import pamr_amrd_amd as pm
from pamr_amrd_amd import hierarchy as H

pytestmark = pytest.mark.gpu

FIELD = {O.OA_U: pm.PMG_F_U, O.OA_RHO: pm.PMG_F_RHO, O.OA_RES: pm.PMG_F_RES, O.OA_A: pm.PMG_F_A,
         O.OA_C: pm.PMG_F_C}


def layout(seed, base_n=17, nref=2, nmin=3, nmax=6, smin=5, smax=13):
    return H.random_boxes(base_n=base_n, nref=nref, nmin=nmin, nmax=nmax, smin=smin, smax=smax, seed=seed)


def pair(seed, op=0, base_n=17, **kw):
    levels = layout(seed, base_n=base_n, **kw)
    print(f"case amr seed={seed} op={op} base={base_n} boxes={[len(l) for l in levels]}", flush=True)
    return O.AmrMG(base_n, levels, op=op), pm.Amr(base_n, levels, op=op)


def each_grid(m):
    for L in range(1, m.nlev + 1):
        for g in range(m.ngrids(L)):
            yield L, g


def test_hierarchy_and_cmask(torch_cuda):
    for seed in (1, 2, 3):
        m, a = pair(seed)
        assert a.nlev == m.nlev and a.base == m.base
        for L, g in each_grid(m):
            assert a.ngrids(L) == m.ngrids(L)
            sm, bm = m.grid(L, g)
            sa, ba = a.grid(L, g)
            assert sm == sa and np.array_equal(bm, ba)
        for L in range(1, m.nlev + 1):
            for which in (0, 1):
                m.set_cmask(L, which)
                for g in range(m.ngrids(L)):
                    assert np.array_equal(a.cmask(L, g, which), m.field(L, g, O.OA_CMASK)), (L, g, which)
        a.close()


def fill_random(m, a, levels, which, rng):
    for L in levels:
        for g in range(m.ngrids(L)):
            v = rng.standard_normal(m.grid(L, g)[0])
            m.field(L, g, which)[:] = v
            a.upload(L, g, FIELD[which], v)


def compare(m, a, levels, which):
    for L in levels:
        for g in range(m.ngrids(L)):
            got = a.download(L, g, FIELD[which])
            want = m.field(L, g, which)
            assert np.array_equal(got, want), (L, g, np.nanmax(np.abs(got - want)))


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_transfers_bit_exact(torch_cuda, seed):
    m, a = pair(seed)
    rng = np.random.default_rng(seed)
    Lf = m.nlev
    for bdy in (0, 1):
        fill_random(m, a, [Lf], O.OA_U, rng)
        m.sync_field(Lf, bdy, O.OA_U)
        a.sync(Lf, bdy, pm.PMG_F_U)
        compare(m, a, [Lf], O.OA_U)
    for op in (1, 2, 3):
        for lf in (Lf, Lf - 1, m.base):
            fill_random(m, a, [lf, lf - 1], O.OA_RES, rng)
            m.inject_field(lf, O.OA_RES, op)
            a.inject(lf, pm.PMG_F_RES, op)
            compare(m, a, [lf - 1], O.OA_RES)
    for ordr in (2, 4):
        for lc in (m.base - 1, m.base, m.base + 1):
            fill_random(m, a, [lc, lc + 1], O.OA_RES, rng)
            m.interp_field(lc, O.OA_RES, ord=ordr)
            a.interp(lc, pm.PMG_F_RES, ord=ordr)
            compare(m, a, [lc + 1], O.OA_RES)
    a.close()


@pytest.mark.parametrize("seed,op", [(6, 0), (7, 1)])
def test_vcycles_fourth_order_cgc_bit_exact(torch_cuda, seed, op):
    """the coarse-grid correction interpolated at fourth order (u_res's
    mg_interp = PAMR_FOURTH_ORDER): V-cycles bit for bit"""
    m, a = pair(seed, op=op)
    set_sources(m, a, op, 4321 + seed)
    for x in (m, a):
        x.set_params(interp_ord=4)
        x.init()
    for c in range(4):
        rm, nm, cm = m.vcycle()
        ra, na, ca = a.vcycle()
        print(f"cycle {c}: oracle {rm:.15e} gpu {ra:.15e}", flush=True)
        assert abs(ra - rm) <= 1e-12 * rm, (c, ra, rm)
        assert cm == ca
        assert np.allclose(na, nm, rtol=1e-12, atol=0), (c, na, nm)
    for L, g in each_grid(m):
        assert np.array_equal(a.download(L, g, pm.PMG_F_U), m.field(L, g, O.OA_U)), ("u", L, g)
    a.close()


def set_sources(m, a, op, seed):
    for L in range(m.base, m.nlev + 1):
        for g in range(m.ngrids(L)):
            shape, bb = m.grid(L, g)
            rho = -H.gaussian_on(shape, bb, seed=seed)
            m.field(L, g, O.OA_RHO)[:] = rho
            a.upload(L, g, pm.PMG_F_RHO, rho)
            if op == 1:
                av = H.lognormal_on(shape, bb)
                cv = H.uniform_on(shape, bb)
                m.field(L, g, O.OA_A)[:] = av
                m.field(L, g, O.OA_C)[:] = cv
                a.upload(L, g, pm.PMG_F_A, av)
                a.upload(L, g, pm.PMG_F_C, cv)


@pytest.mark.parametrize("seed,op,graph", [(1, 0, True), (2, 0, True), (3, 0, False), (4, 1, True), (5, 1, True)])
def test_vcycles_bit_exact(torch_cuda, seed, op, graph):
    m, a = pair(seed, op=op)
    a.set_graph(graph)
    set_sources(m, a, op, 1234 + seed)
    for x in (m, a):
        x.set_params()
        x.init()
    for L, g in each_grid(m):
        for w in (O.OA_U, O.OA_RHO):
            assert np.array_equal(a.download(L, g, FIELD[w]), m.field(L, g, w)), ("init", L, g, w)
    for c in range(5):
        rm, nm, cm = m.vcycle()
        ra, na, ca = a.vcycle()
        print(f"cycle {c}: oracle {rm:.15e} gpu {ra:.15e} citer {cm} {ca}", flush=True)
        assert abs(ra - rm) <= 1e-12 * rm, (c, ra, rm)
        assert cm == ca
        assert np.allclose(na, nm, rtol=1e-12, atol=0), (c, na, nm)
    for L, g in each_grid(m):
        assert np.array_equal(a.download(L, g, pm.PMG_F_U), m.field(L, g, O.OA_U)), ("u", L, g)
    a.close()


def test_larger_layout_cycles(torch_cuda):
    """base 33^3 with 6-10 boxes of 9-25 points per refined level"""
    m, a = pair(11, base_n=33, nmin=6, nmax=10, smin=9, smax=25)
    set_sources(m, a, 0, 99)
    for x in (m, a):
        x.set_params()
        x.init()
    for c in range(3):
        rm, _, _ = m.vcycle()
        ra, _, _ = a.vcycle()
        assert abs(ra - rm) <= 1e-12 * rm, (c, ra, rm)
    for L, g in each_grid(m):
        assert np.array_equal(a.download(L, g, pm.PMG_F_U), m.field(L, g, O.OA_U)), ("u", L, g)
    print("plan stats", [a.plan_stats(L) for L in range(1, a.nlev + 1)], flush=True)
    a.close()


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4])
def test_sweep_tilings_bit_exact(torch_cuda, variant):
    """Every compiled tile variant, with z-chunks 2/4/8 spread over the
    levels, gives the oracle's V-cycles bit for bit (the tiling changes
    only the summation order of the norms)."""
    m, a = pair(21 + variant, op=variant % 2, base_n=33, nmin=4, nmax=8, smin=9, smax=25)
    for L in range(1, a.nlev + 1):
        a.set_level_sweep_config(L, variant, (2, 4, 8)[L % 3])
        assert a.sweep_config(L) == (variant, (2, 4, 8)[L % 3])
    set_sources(m, a, variant % 2, 500 + variant)
    for x in (m, a):
        x.set_params()
        x.init()
    for c in range(3):
        rm, nm, cm = m.vcycle()
        ra, na, ca = a.vcycle()
        print(f"variant {variant} cycle {c}: oracle {rm:.15e} gpu {ra:.15e}", flush=True)
        assert abs(ra - rm) <= 1e-12 * rm, (c, ra, rm)
        assert cm == ca
        assert np.allclose(na, nm, rtol=1e-12, atol=0), (c, na, nm)
    for L, g in each_grid(m):
        assert np.array_equal(a.download(L, g, pm.PMG_F_U), m.field(L, g, O.OA_U)), ("u", L, g)
    a.close()


def test_autotune_keeps_state(torch_cuda):
    """The per-level autotuner restores u around its timing runs: cycles
    after tuning still equal the oracle's bit for bit."""
    from pamr_amrd_amd import autotune as AT
    m, a = pair(31, op=1, base_n=33, nmin=6, nmax=10, smin=9, smax=25)
    set_sources(m, a, 1, 77)
    for x in (m, a):
        x.set_params()
        x.init()
    rm, _, _ = m.vcycle()
    ra, _, _ = a.vcycle()
    assert abs(ra - rm) <= 1e-12 * rm
    chosen = AT.autotune_amr(a, zchunks=(2, 4, 8), nrep=2)
    print(f"autotuned {chosen}", flush=True)
    assert set(chosen) == set(range(2, a.nlev + 1))
    for L, (v, z, _, _) in chosen.items():
        assert a.sweep_config(L) == (v, z)
    for c in range(2):
        rm, _, _ = m.vcycle()
        ra, _, _ = a.vcycle()
        assert abs(ra - rm) <= 1e-12 * rm, (c, ra, rm)
    for L, g in each_grid(m):
        assert np.array_equal(a.download(L, g, pm.PMG_F_U), m.field(L, g, O.OA_U)), ("u", L, g)
    a.close()


def test_sweep_config_rejects(torch_cuda):
    m, a = pair(3)
    v0 = a.sweep_config(a.nlev)
    for bad in ((99, 4), (0, 0)):
        with pytest.raises(RuntimeError):
            a.set_level_sweep_config(a.nlev, *bad)
        assert a.sweep_config(a.nlev) == v0
    a.close()


@pytest.mark.parametrize("op", [0, 1])
def test_coarsest_cycle_skip_bit_exact(torch_cuda, op):
    """66 cycles on a 33^3-base hierarchy: from about cycle 60 the coarsest
    loop cannot reach crtol and the reference relaxes to max_citer; the
    periodic states are skipped -- the same iteration counts every cycle and
    u bit for bit on every grid against the oracle"""
    levels = layout(5, base_n=33, nmin=4, nmax=6, smin=9, smax=17)
    print(f"case amr coarsest skip op={op} boxes={[len(l) for l in levels]}", flush=True)
    m = O.AmrMG(33, levels, op=op)
    a = pm.Amr(33, levels, op=op)
    for L in range(m.base, m.nlev + 1):
        for g in range(m.ngrids(L)):
            shape, bb = m.grid(L, g)
            vals = {O.OA_RHO: -H.gaussian_on(shape, bb, seed=77)}
            if op == 1:
                vals[O.OA_A] = H.lognormal_on(shape, bb)
                vals[O.OA_C] = H.uniform_on(shape, bb)
            for w, v in vals.items():
                m.field(L, g, w)[:] = v
                a.upload(L, g, FIELD[w], v)
    for x in (m, a):
        x.set_params()
        x.init()
    cits = []
    for c in range(66):
        rm, nm, cm = m.vcycle()
        ra, na, ca = a.vcycle()
        assert cm == ca, (c, cm, ca)
        assert abs(ra - rm) <= 1e-12 * rm, (c, ra, rm)
        cits.append(ca)
    print(f"coarsest iterations {cits}", flush=True)
    assert cits[-1] == 5000
    for L, g in each_grid(m):
        assert np.array_equal(a.download(L, g, pm.PMG_F_U), m.field(L, g, O.OA_U)), ("u", L, g)
    a.close()


@pytest.mark.parametrize("op", [0, 1])
def test_mixed_precision_amr_reaches_1e10(torch_cuda, op):
    """FP32 V-cycles of the hierarchy in FP64 defect correction
    (pmg_amr_solve_rel mode 1) reach a 1e-10 relative FP64 residual.  The
    defect is taken after u is injected down the AMR levels, as the FAS cycle
    does, so in exact arithmetic the two iterations are the same: their
    residual histories agree while FP32 rounding is far below the residual.
    The composite solution -- every finest-level grid, the coarser AMR levels
    where no finer grid covers them -- is compared with the FP64 solve's to
    1e-13.  A fine grid's AMR-boundary values are set only by coarse-grid
    corrections, which no later defect measures, so the FP32 rounding of the
    corrections stays in them: FP32 from the start agrees to 1e-6 (FP32
    rounding of the accumulated corrections), FP32 after FP64 cycles down to
    1e-3 (pmg_amr_set_mixed_switch) to 1e-9.  Measured values are printed."""
    levels = layout(5, base_n=33, nmin=4, nmax=6, smin=9, smax=17)
    print(f"case amr mixed op={op} boxes={[len(l) for l in levels]}", flush=True)
    runs = []
    for mode, switch, rtol in ((0, 1.0, 1e-13), (1, 1.0, 1e-10), (1, 1e-3, 1e-10)):
        a = pm.Amr(33, levels, op=op)
        for L in range(a.base, a.nlev + 1):
            for g in range(a.ngrids(L)):
                shape, bb = a.grid(L, g)
                a.upload(L, g, pm.PMG_F_RHO, -H.gaussian_on(shape, bb, seed=77))
                if op == 1:
                    a.upload(L, g, pm.PMG_F_A, H.lognormal_on(shape, bb))
                    a.upload(L, g, pm.PMG_F_C, H.uniform_on(shape, bb))
        a.set_params()
        a.init()
        a.set_mixed_switch(switch)
        r0, h = a.solve_rel(mode, rtol, 600)
        print(f"mode {mode} switch {switch:g}: {len(h)} iterations, final {h[-1] / r0:.3e}", flush=True)
        assert len(h) and h[-1] <= rtol * r0, (mode, switch, len(h), h[-3:] / r0)
        sol = {}
        for L in range(a.base, a.nlev + 1):
            for g in range(a.ngrids(L)):
                own = a.cmask(L, g, 1) == 0.0 if L < a.nlev else np.ones(a.grid(L, g)[0], dtype=bool)
                sol[L, g] = (a.download(L, g, pm.PMG_F_U), own)
        runs.append((r0, np.asarray(h), sol))
        a.close()
    r64, h64, s64 = runs[0]
    scale = max(np.abs(u[own]).max() for u, own in s64.values() if own.any())
    for (r32, h32, s32), bound in zip(runs[1:], (1e-6, 1e-9)):
        assert abs(r32 - r64) <= 1e-12 * r64
        k = int(np.sum(h64[:len(h32)] > 1e-5 * r64))
        assert k >= 3
        rel = np.abs(h32[:k] - h64[:k]) / h64[:k]
        err = max(np.abs(um[own] - u64[own]).max() for (um, own), (u64, _) in zip(s32.values(), s64.values())
                  if own.any())
        print(f"residual histories: first {k} iterations agree to {rel.max():.2e}; composite solution vs FP64: "
              f"{err / scale:.3e} relative (bound {bound:g})", flush=True)
        assert rel.max() <= 1e-3
        assert err <= bound * scale
