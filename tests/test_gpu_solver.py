"""GPU parity of the uniform-hierarchy FAS solver against the oracle driver
(amrd/mg.c restated).  Per-cycle residual norms within 1e-12 relative; the
solution after the cycles bit-exact; masks bit-exact."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pm(torch_cuda):
    import pamr_amrd_amd as pm
    pm.lib()
    return pm


def _pair(pm, n, rho, op=0, a=None, c=None, pre=2, pst=2, variant=1, zchunk=32, graph=True):
    S = pm.Solver(n, op=op)
    S.set_params(pre=pre, pst=pst)
    S.set_sweep_config(variant, zchunk)
    S.set_graph(graph)
    m = O.UniformMG(n, op=op)
    m.set_params(pre=pre, pst=pst)
    L = S.nlev
    S.upload(L, pm.PMG_F_RHO, rho)
    m.field(L, O.F_RHO)[:] = rho
    if op == 1:
        S.upload(L, pm.PMG_F_A, a)
        S.upload(L, pm.PMG_F_C, c)
        m.field(L, O.F_A)[:] = a
        m.field(L, O.F_C)[:] = c
    S.init()
    m.init()
    return S, m


@pytest.mark.parametrize("variant,zchunk", [(v, 32) for v in range(12)] + [(6, 5), (7, 4), (8, 7), (9, 5), (11, 3)])
def test_c1_ten_cycles_65(pm, variant, zchunk):
    """65^3 Poisson, seeded Gaussian source, V(2,2), FW, 10 cycles, every
    sweep tile variant (z-chunks that leave every remainder of the
    three-step march of the two-planes-in-flight variants)."""
    print(f"case c1 65^3 variant={variant} zchunk={zchunk}", flush=True)
    n = 65
    rho = -pm.gaussian_source(n, seed=1234)
    S, m = _pair(pm, n, rho, variant=variant, zchunk=zchunk)
    assert S.nlev == m.nlev == 5
    for cyc in range(10):
        r_gpu, ln_gpu, ci_gpu = S.vcycle()
        r_ref, ln_ref, ci_ref = m.vcycle()
        assert ci_gpu == ci_ref, (cyc, ci_gpu, ci_ref)
        assert abs(r_gpu - r_ref) <= 1e-12 * abs(r_ref), (cyc, r_gpu, r_ref)
    u_gpu = S.download(S.nlev, pm.PMG_F_U)
    u_ref = m.field(m.nlev, O.F_U)
    assert np.array_equal(u_gpu, u_ref), np.abs(u_gpu - u_ref).max()
    S.close()


@pytest.mark.parametrize("n", [17, 33, 129])
def test_cycles_other_sizes(pm, n):
    print(f"case cycles n={n}", flush=True)
    rho = -pm.gaussian_source(n, seed=99 + n)
    S, m = _pair(pm, n, rho, zchunk=8)
    for cyc in range(4):
        r_gpu, _, _ = S.vcycle()
        r_ref, _, _ = m.vcycle()
        assert abs(r_gpu - r_ref) <= 1e-12 * abs(r_ref), (cyc, r_gpu, r_ref)
    assert np.array_equal(S.download(S.nlev, pm.PMG_F_U), m.field(m.nlev, O.F_U))
    S.close()


@pytest.mark.parametrize("n,variant,zchunk", [(33, 1, 32), (65, 5, 32), (65, 6, 5), (65, 9, 7), (65, 10, 32),
                                               (65, 11, 3), (65, 2, 8)])
def test_varcoef_cycles(pm, n, variant, zchunk):
    """Variable-coefficient cycles, the coefficient a in its LDS ring (4 or 5
    planes) where both rings fit, read from global memory otherwise (the
    FP64 128 x 16 tiles)."""
    print(f"case varcoef {n} variant={variant} zchunk={zchunk}", flush=True)
    rng = np.random.default_rng(5)
    rho = -pm.gaussian_source(n, seed=5)
    a = np.exp(0.5 * rng.standard_normal((n, n, n)))
    c = rng.random((n, n, n))
    S, m = _pair(pm, n, rho, op=1, a=a, c=c, variant=variant, zchunk=zchunk)
    for cyc in range(4):
        r_gpu, _, _ = S.vcycle()
        r_ref, _, _ = m.vcycle()
        assert abs(r_gpu - r_ref) <= 1e-12 * abs(r_ref), (cyc, r_gpu, r_ref)
    assert np.array_equal(S.download(S.nlev, pm.PMG_F_U), m.field(m.nlev, O.F_U))
    S.close()


def test_masks_bit_exact(pm):
    n = 33
    S = pm.Solver(n)
    m = O.UniformMG(n)
    m.init()
    m.vcycle()
    for L in range(1, S.nlev + 1):
        # after a cycle the oracle's cmask holds set_cmask_bdy of every level
        assert np.array_equal(S.cmask(L, 0), m.field(L, O.F_CMASK))
        if L < S.nlev:
            assert S.cmask(L, 1).sum() == np.prod(S.shape(L))
    S.close()


def test_graph_replays_match_eager(pm):
    n = 65
    rho = -pm.gaussian_source(n, seed=3)
    Sg, _ = _pair(pm, n, rho, graph=True)
    Se, _ = _pair(pm, n, rho, graph=False)
    for _ in range(5):
        rg = Sg.vcycle()[0]
        re = Se.vcycle()[0]
        assert rg == re
    assert np.array_equal(Sg.download(Sg.nlev, pm.PMG_F_U), Se.download(Se.nlev, pm.PMG_F_U))
    Sg.close()
    Se.close()


@pytest.mark.parametrize("n,pre,pst", [(65, 2, 2), (129, 2, 2), (129, 3, 1), (257, 2, 2)])
def test_sweep_blocking_matches_single(pm, n, pre, pst):
    """pairs of sweeps run as one temporally blocked pass give the same u,
    norms and coarsest iteration counts as single sweeps (also with odd
    sweep counts, where u ends a level in the ping-pong twin)"""
    rho = -pm.gaussian_source(n, seed=8)
    Sb, _ = _pair(pm, n, rho, pre=pre, pst=pst) if n < 257 else (_solver_only(pm, n, rho, pre, pst), None)
    Ss, _ = _pair(pm, n, rho, pre=pre, pst=pst) if n < 257 else (_solver_only(pm, n, rho, pre, pst), None)
    Sb.set_sweep_blocking(True)
    Ss.set_sweep_blocking(False)
    for _ in range(3):
        rb, nb, cb = Sb.vcycle()
        rs, ns, cs = Ss.vcycle()
        assert cb == cs
        assert abs(rb - rs) <= 1e-13 * rs
        assert np.allclose(nb, ns, rtol=1e-12, atol=0)
    for L in range(1, Sb.nlev + 1):
        assert np.array_equal(Sb.download(L, pm.PMG_F_U), Ss.download(L, pm.PMG_F_U)), L
    Sb.zero_solution()
    Ss.zero_solution()
    r0b, hb = Sb.solve_rel(1, 1e-10, 40)
    r0s, hs = Ss.solve_rel(1, 1e-10, 40)
    assert hb[-1] <= 1e-10 * r0b and hs[-1] <= 1e-10 * r0s
    assert abs(len(hb) - len(hs)) <= 1
    Sb.close()
    Ss.close()


def _solver_only(pm, n, rho, pre, pst):
    S = pm.Solver(n)
    S.set_params(pre=pre, pst=pst)
    S.upload(S.nlev, pm.PMG_F_RHO, rho)
    S.init()
    return S


@pytest.mark.parametrize("kind", [1, 2, 3])
@pytest.mark.parametrize("n,op", [(33, 0), (65, 0), (129, 0), (65, 1), (97, 1)])
def test_restrict_fusion_matches_unfused(pm, n, op, kind):
    """the finest residual computed inside the injection gives the same
    coarse u and res, the same cycles and the same u as residual + dmfwr3d
    (kind 1: residual planes in an LDS ring; 2: restriction streamed in z)"""
    print(f"case restrict fusion n={n} op={op} kind={kind}", flush=True)
    rho = -pm.gaussian_source(n, seed=6)
    a = c = None
    if op == 1:
        rng = np.random.default_rng(n)
        a = np.exp(0.5 * rng.standard_normal((n, n, n)))
        c = rng.random((n, n, n))
    Sf, _ = _pair(pm, n, rho, op=op, a=a, c=c)
    Su, _ = _pair(pm, n, rho, op=op, a=a, c=c)
    Sf.set_restrict_fusion(kind)
    Su.set_restrict_fusion(False)
    for _ in range(3):
        rf, nf, cf = Sf.vcycle()
        ru, nu, cu = Su.vcycle()
        assert cf == cu
        assert abs(rf - ru) <= 1e-13 * ru
        assert np.allclose(nf, nu, rtol=1e-12, atol=0)
    for L in range(1, Sf.nlev + 1):
        for w in (pm.PMG_F_U, pm.PMG_F_RES):
            if L == Sf.nlev and w == pm.PMG_F_RES:
                continue
            assert np.array_equal(Sf.download(L, w), Su.download(L, w)), (L, w)
    # the FP32 inner cycles fuse the same way: identical outer histories and u
    Sf.zero_solution()
    Su.zero_solution()
    r0f, hf = Sf.solve_rel(1, 1e-10, 40)
    r0u, hu = Su.solve_rel(1, 1e-10, 40)
    assert hf[-1] <= 1e-10 * r0f and hu[-1] <= 1e-10 * r0u
    assert r0f == r0u
    assert np.array_equal(hf, hu)
    assert np.array_equal(Sf.download(Sf.nlev, pm.PMG_F_U), Su.download(Su.nlev, pm.PMG_F_U))
    Sf.close()
    Su.close()


@pytest.mark.parametrize("kind", [1, 2, 3])
@pytest.mark.parametrize("cz,n", [(1, 65), (3, 65), (5, 65), (16, 65), (2, 257), (13, 257)])
def test_resid_restrict_zchunk_bit_exact(pm, cz, n, kind):
    """every z-chunk of the fused finest residual + injection (both kernels)
    gives the same cycles (FP64) and the same mixed-precision history (FP32)
    as the default"""
    print(f"case resid_restrict zchunk {cz} n={n} kind={kind}", flush=True)
    rho = -pm.gaussian_source(n, seed=11)
    Sd, _ = _pair(pm, n, rho)
    Sz, _ = _pair(pm, n, rho)
    Sz.set_restrict_fusion(kind)
    Sz.set_resid_restrict_zchunk(0, cz)
    Sz.set_resid_restrict_zchunk(1, cz)
    for _ in range(3):
        rd, nd, cd = Sd.vcycle()
        rz, nz_, cz_ = Sz.vcycle()
        assert cd == cz_
        assert abs(rd - rz) <= 1e-13 * rd
    assert np.array_equal(Sd.download(Sd.nlev, pm.PMG_F_U), Sz.download(Sz.nlev, pm.PMG_F_U))
    Sd.zero_solution()
    Sz.zero_solution()
    _, hd = Sd.solve_rel(1, 1e-10, 40)
    _, hz = Sz.solve_rel(1, 1e-10, 40)
    assert np.array_equal(hd, hz)
    assert np.array_equal(Sd.download(Sd.nlev, pm.PMG_F_U), Sz.download(Sz.nlev, pm.PMG_F_U))
    Sd.close()
    Sz.close()


@pytest.mark.parametrize("n", [65, 129])
def test_bottom_fusion_matches_unfused(pm, n):
    """the single-workgroup bottom V (levels <= 17^3) gives the same u, the
    same coarsest iteration counts and the same norms (up to summation order)
    as the per-level launches"""
    rho = -pm.gaussian_source(n, seed=5)
    Sf, _ = _pair(pm, n, rho)
    Su, _ = _pair(pm, n, rho)
    Su.set_bottom_fusion(False)
    for _ in range(4):
        rf, nf, cf = Sf.vcycle()
        ru, nu, cu = Su.vcycle()
        assert cf == cu
        assert abs(rf - ru) <= 1e-13 * ru
        assert np.allclose(nf, nu, rtol=1e-12, atol=0)
    for L in range(1, Sf.nlev + 1):
        assert np.array_equal(Sf.download(L, pm.PMG_F_U), Su.download(L, pm.PMG_F_U)), L
    # FP32 inner cycles take the fused path too (from u = 0 again)
    Sf.zero_solution()
    Su.zero_solution()
    r0f, hf = Sf.solve_rel(1, 1e-10, 40)
    r0u, hu = Su.solve_rel(1, 1e-10, 40)
    assert hf[-1] <= 1e-10 * r0f and hu[-1] <= 1e-10 * r0u
    assert abs(len(hf) - len(hu)) <= 1
    Sf.close()
    Su.close()


@pytest.mark.parametrize("op,fused", [(0, True), (0, False), (1, True), (1, False)])
def test_coarsest_cycle_skip_bit_exact(pm, op, fused):
    """once the solution is converged to rounding, the coarsest loop cannot
    reach crtol and the reference relaxes to max_citer: the periodic states
    are skipped (same iteration count, same u, same norms) -- 20 cycles at
    33^3 against the oracle, with the bottom fused and with per-level
    launches"""
    n = 33
    print(f"case coarsest cycle skip op={op} fused={fused}", flush=True)
    rho = -pm.gaussian_source(n, seed=5)
    a = c = None
    if op == 1:
        rng = np.random.default_rng(3)
        a = np.exp(0.5 * rng.standard_normal((n, n, n)))
        c = rng.random((n, n, n))
    S, m = _pair(pm, n, rho, op=op, a=a, c=c)
    S.set_bottom_fusion(fused)
    cits = []
    for k in range(20):
        rm, nm, cm = m.vcycle()
        rs, ns, cs = S.vcycle()
        assert cm == cs, (k, cm, cs)
        assert abs(rs - rm) <= 1e-12 * rm, (k, rs, rm)
        assert np.allclose(ns, nm, rtol=1e-12, atol=0), (k, ns, nm)
        cits.append(cs)
    print(f"coarsest iterations {cits}", flush=True)
    assert cits[-1] == 5000
    for L in range(1, S.nlev + 1):
        assert np.array_equal(S.download(L, pm.PMG_F_U), m.field(L, O.F_U)), L
    S.close()


@pytest.mark.parametrize("op", [0, 1])
def test_bottom_lds_bit_identical(pm, op):
    """the bottom V with its arrays in LDS gives the same bits as from global
    memory: FP64 cycles with the bottom up to 9^3 (its arrays fit), FP32
    inner cycles with the bottom up to 17^3 -- u of every level, the norms,
    the coarsest iteration counts and the FP64 residual history"""
    n = 65
    print(f"case bottom lds op={op}", flush=True)
    rho = -pm.gaussian_source(n, seed=8)
    a = c = None
    if op == 1:
        rng = np.random.default_rng(3)
        a = np.exp(0.5 * rng.standard_normal((n, n, n)))
        c = rng.random((n, n, n))
    Sl, _ = _pair(pm, n, rho, op=op, a=a, c=c)
    Sg, _ = _pair(pm, n, rho, op=op, a=a, c=c)
    Sg.set_bottom_lds(False)
    for S in (Sl, Sg):
        S.set_bottom_fusion(True, max_n=9)
    for _ in range(3):
        rl, nl, cl = Sl.vcycle()
        rg, ng, cg = Sg.vcycle()
        assert rl == rg and cl == cg and np.array_equal(nl, ng)
    for L in range(1, Sl.nlev + 1):
        assert np.array_equal(Sl.download(L, pm.PMG_F_U), Sg.download(L, pm.PMG_F_U)), L
    for S in (Sl, Sg):
        S.set_bottom_fusion(True)
        S.zero_solution()
    r0l, hl = Sl.solve_rel(1, 1e-10, 40)
    r0g, hg = Sg.solve_rel(1, 1e-10, 40)
    assert r0l == r0g and np.array_equal(hl, hg)
    assert np.array_equal(Sl.download(Sl.nlev, pm.PMG_F_U), Sg.download(Sg.nlev, pm.PMG_F_U))
    Sl.close()
    Sg.close()


def test_mixed_precision_reaches_1e10(pm):
    """FP32 inner V-cycles in FP64 defect correction reach ||r||/||r0|| <= 1e-10;
    the error against a 40-cycle FP64 solution is reported and bounded."""
    n = 65
    rho = -pm.gaussian_source(n, seed=1234)
    S, m = _pair(pm, n, rho)
    r0, hist = S.solve_rel(1, 1e-10, 40)
    assert hist[-1] <= 1e-10 * r0, (r0, hist)
    u_mixed = S.download(S.nlev, pm.PMG_F_U)
    for _ in range(40):
        m.vcycle()
    u_ref = m.field(m.nlev, O.F_U)
    err = np.abs(u_mixed - u_ref).max() / np.abs(u_ref).max()
    print(f"mixed: iters={len(hist)} final rel res={hist[-1] / r0:.3e} rel err vs FP64 oracle={err:.3e}")
    assert err < 1e-9
    S.close()


def test_plain_fp64_rel_solve(pm):
    n = 65
    rho = -pm.gaussian_source(n, seed=1234)
    S, m = _pair(pm, n, rho)
    r0, hist = S.solve_rel(0, 1e-10, 40)
    assert hist[-1] <= 1e-10 * r0
    assert len(hist) < 15
    S.close()


@pytest.mark.parametrize("n", [33, 129])
def test_fp32_exact_bottom(pm, n):
    """FP32 inner cycles with the 17^3 level solved exactly by the sine
    transform (the default) against the relaxed bottom V: both reach 1e-10,
    the exact one in no more iterations, and the two solutions agree with a
    40-cycle FP64 solve to 1e-9."""
    print(f"case fp32 exact bottom n={n}", flush=True)
    rho = -pm.gaussian_source(n, seed=77)
    S, m = _pair(pm, n, rho)
    S.init()
    S.solve(40, 40)
    u64 = S.download(S.nlev, pm.PMG_F_U)
    out = {}
    for on in (False, True):
        S.set_fp32_exact_bottom(on)
        S.zero_solution()
        S.init()
        r0, h = S.solve_rel(1, 1e-10, 60)
        assert h[-1] <= 1e-10 * r0, (on, h[-3:])
        u = S.download(S.nlev, pm.PMG_F_U)
        err = np.abs(u - u64).max() / np.abs(u64).max()
        out[on] = (len(h), err, (h[-1] / r0) ** (1.0 / len(h)))
        assert err < 1e-9, (on, err)
    print(f"relaxed bottom: {out[False]}, exact bottom: {out[True]}", flush=True)
    assert out[True][0] <= out[False][0]
    S.close()


def test_time_kernel_in_cycle(pm):
    """pmg_time_kernel which 3: the finest FP64 sweep timed inside ungraphed
    V-cycles -- positive and of the order of the back-to-back figure; FP32
    is refused."""
    n = 65
    rho = -pm.gaussian_source(n, seed=1234)
    S, m = _pair(pm, n, rho)
    t_in = S.time_kernel(3, S.nlev, 0, 4)
    t_b2b = S.time_kernel(0, S.nlev, 0, 4)
    print(f"in-cycle {t_in:.4f} ms, back to back {t_b2b:.4f} ms", flush=True)
    assert 0 < t_in < 10 * t_b2b
    with pytest.raises(RuntimeError):
        S.time_kernel(3, S.nlev, 1, 4)
    S.close()
