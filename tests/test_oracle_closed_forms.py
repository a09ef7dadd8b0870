"""Pins of the oracle driver by what the mathematics fixes.

* the discrete eigenfunction problem has the closed-form solution phi;
* for a linear operator the FAS V-cycle (amrd/mg.c) equals the correction
  scheme built independently here from the oracle's transfer and hook
  kernels (themselves pinned bit-exactly to the reference Fortran);
* zero data is a fixed point; V(2,2) contracts by the textbook factor;
* transfers reproduce affine / trilinear functions;
* mask and coarsening rules agree with brute-force/closed-form values.
A dropped term, a wrong sign, a swapped f_lop/f_res or a wrong coarsest-loop
rule breaks at least one of these.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O

L = O.lib()


def grid(n):
    x = np.linspace(0.0, 1.0, n)
    return x[None, None, :], x[None, :, None], x[:, None, None]


def test_level_counts_closed_form():
    # (n-1)/2^k + 1 >= 5 and the next halving < 5 marks the coarsest level
    for n, nl in [(65, 5), (129, 6), (257, 7), (513, 8), (17, 3), (9, 2)]:
        m = O.UniformMG(n)
        assert m.nlev == nl
        assert m.shape(1) == (5, 5, 5)
        assert m.shape(nl) == (n, n, n)


def test_island_no_overlap_rule_as_written():
    # src/gh.c:1455-1456 tests bbox_i.max > bbox_j.min: two disjoint boxes
    # with i below j are merged, overlapping ones with i above j are not
    dx = np.array([0.1, 0.1, 0.1])
    mc = (C.c_int * 3)(5, 5, 5)
    for boxes, expect_same in [
        ([0, 1, 0, 1, 0, 1, 2, 3, 2, 3, 2, 3], True),   # i entirely below j
        ([0.5, 1.5, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1], False),  # overlapping, i above
    ]:
        bb = np.array(boxes, dtype=np.float64)
        isl = np.zeros(2)
        L.or_calc_island_no(O.P(bb), 2, O.P(isl), O.P(dx), 2, 3, mc)
        assert (isl[0] == isl[1]) == expect_same


def test_cmask_rules_brute_force():
    shape = (C.c_int * 3)(9, 9, 9)
    base = np.array([0.0, 1.0, 0.0, 1.0, 0.0, 1.0])
    per = (C.c_int * 3)(0, 0, 0)
    dx = np.array([0.125, 0.125, 0.125])
    m = np.zeros(729)
    # base-covering grid: every point ON
    L.or_set_cmask_bdy(O.P(m), shape, O.P(base), O.P(base), O.P(dx), per)
    assert m.sum() == 729
    # interior grid [0.25,0.75]^3 with dx 1/16: AMR boundaries are OFF
    g = np.array([0.25, 0.75, 0.25, 0.75, 0.25, 0.75])
    dx2 = np.array([0.0625] * 3)
    L.or_set_cmask_bdy(O.P(m), shape, O.P(g), O.P(base), O.P(dx2), per)
    assert m.reshape(9, 9, 9)[1:-1, 1:-1, 1:-1].sum() == 343 and m.sum() == 343
    # child box [0.25,0.5]^3 inside the base grid (dx 1/8): indices 2..4,
    # shrunk by one at the non-physical child edges -> 3..3
    cb = np.array([0.25, 0.5, 0.25, 0.5, 0.25, 0.5])
    L.or_set_cmask_child(O.P(m), shape, O.P(base), O.P(base), O.P(dx), per, 1, O.P(cb))
    mm = m.reshape(9, 9, 9)
    assert mm.sum() == 1 and mm[3, 3, 3] == 1
    # child touching the physical boundary keeps the boundary index
    cb = np.array([0.0, 0.5, 0.0, 0.5, 0.0, 0.5])
    L.or_set_cmask_child(O.P(m), shape, O.P(base), O.P(base), O.P(dx), per, 1, O.P(cb))
    mm = m.reshape(9, 9, 9)
    assert mm.sum() == 4 ** 3 and mm[0:4, 0:4, 0:4].all()


def test_restriction_reproduces_affine():
    n = 17
    X, Y, Z = grid(n)
    f = 0.3 + 1.5 * X - 2.0 * Y + 0.75 * Z + 0 * X * Y * Z
    f = np.broadcast_to(f, (n, n, n)).copy()
    Xc, Yc, Zc = grid(9)
    fc = np.broadcast_to(0.3 + 1.5 * Xc - 2.0 * Yc + 0.75 * Zc, (9, 9, 9))
    for op in (2, 3):
        r = O.restrict(f, op=op)
        # exact at every coarse point (interior and reduced boundary stencils)
        assert np.allclose(r, fc, rtol=0, atol=1e-14)


def test_interpolation_exact_for_trilinear():
    X, Y, Z = grid(9)
    g = (1 + X) * (2 - Y) * (0.5 + Z) + 3 * X * Y * Z
    g = np.broadcast_to(g, (9, 9, 9)).copy()
    Xf, Yf, Zf = grid(17)
    gf = (1 + Xf) * (2 - Yf) * (0.5 + Zf) + 3 * Xf * Yf * Zf
    assert np.allclose(O.interp(g), gf, rtol=0, atol=1e-14)


def test_cubic_interpolation_exact_for_tricubic():
    """dvi4q1d is 4-point Lagrange interpolation: exact for polynomials of
    degree <= 3 in each variable (the premise of test/test1.c:287-288)"""
    X, Y, Z = grid(9)
    p = lambda x, y, z: (1 + x - 2 * x ** 3) * (0.5 - y ** 2 + y ** 3) * (2 + z + 3 * z ** 3) + x * y * z ** 2
    g = np.broadcast_to(p(X, Y, Z), (9, 9, 9)).copy()
    Xf, Yf, Zf = grid(17)
    assert np.allclose(O.interp(g, ord=4), p(Xf, Yf, Zf), rtol=0, atol=1e-13)
    # a cubic term breaks the trilinear interpolation, as it must
    assert not np.allclose(O.interp(g, ord=2), p(Xf, Yf, Zf), rtol=0, atol=1e-6)


def _setup(n, rho):
    m = O.UniformMG(n)
    m.set_params(pre=2, pst=2, max_citer=5000, crtol=1e-3, rop_u=3, rop_res=3, rop_src=3)
    m.field(m.nlev, O.F_RHO)[:] = rho
    m.init()
    return m


def test_zero_data_fixed_point():
    m = _setup(17, np.zeros((17, 17, 17)))
    for _ in range(3):
        r, _, _ = m.vcycle()
        assert r == 0.0
    assert not m.field(m.nlev, O.F_U).any()


def test_discrete_eigenfunction_solution():
    n = 33
    h = 1.0 / (n - 1)
    X, Y, Z = grid(n)
    phi = np.sin(np.pi * X) * np.sin(np.pi * Y) * np.sin(np.pi * Z)
    lam = 12.0 / h ** 2 * np.sin(np.pi * h / 2) ** 2
    rho = np.broadcast_to(-lam * phi, (n, n, n)).copy()
    rho[0], rho[-1], rho[:, 0], rho[:, -1], rho[:, :, 0], rho[:, :, -1] = 0, 0, 0, 0, 0, 0
    m = _setup(n, rho)
    for _ in range(14):
        m.vcycle()
    u = m.field(m.nlev, O.F_U)
    assert np.max(np.abs(u - phi)) < 1e-11


def test_contraction_factor():
    f = O.__dict__  # noqa: F841 (keep module import)
    import pamr_amrd_amd as pm
    n = 65
    rho = -pm.gaussian_source(n, seed=7)
    m = _setup(n, rho)
    norms = [m.vcycle()[0] for _ in range(8)]
    ratios = np.array(norms[1:]) / np.array(norms[:-1])
    # V(2,2) RB-GS, FW/trilinear, 3D Poisson: ~0.05-0.15 per cycle
    assert np.all(ratios < 0.2), ratios
    assert np.all(ratios > 0.005), ratios


def _cs_vcycle(levels, shapes, hdx, crtol=1e-3, pre=2, pst=2, max_citer=5000):
    """Correction-scheme V-cycle for lap(e) = b written from the oracle's
    pinned kernels: levels[-1] holds (u, rho) on the finest grid."""
    nl = len(shapes)
    e = [None] * nl
    rho = [None] * nl
    cm = [np.ones(s) for s in shapes]
    zero = [np.zeros(s) for s in shapes]
    u, rho_f = levels
    e[-1] = u
    rho[-1] = rho_f
    res_pre = [0.0] * nl

    def relax(l):
        nz, ny, nx = shapes[l]
        return L.or_relax(0, O.P(e[l]), O.P(zero[l]), O.P(rho[l]), None, None, O.P(cm[l]), None, 1.0, nx,
                          ny, nz, *hdx[l])

    def residual(l):
        nz, ny, nx = shapes[l]
        r = np.zeros(shapes[l])
        L.or_residual(0, O.P(r), O.P(zero[l]), O.P(e[l]), O.P(rho[l]), None, None, O.P(cm[l]), None, 1.0,
                      nx, ny, nz, *hdx[l])
        return r

    for l in range(nl - 1, -1, -1):
        if l < nl - 1:
            r = residual(l + 1)
            rho[l] = -O.restrict(r, op=3)
            e[l] = np.zeros(shapes[l])
        for _ in range(pre):
            res_pre[l] = relax(l)
    for l in range(nl):
        ctol = res_pre[l] * crtol
        if l > 0:
            e[l] += O.interp(e[l - 1])
        it = 0
        for it in range(pst):
            relax(l)
        it = pst
        if l == 0:
            rr = ctol + 1
            while it < max_citer and rr > ctol:
                rr = relax(l)
                it += 1
    nz, ny, nx = shapes[-1]
    r = np.zeros(shapes[-1])
    return L.or_residual(0, O.P(r), O.P(zero[-1]), O.P(e[-1]), O.P(rho[-1]), None, None, O.P(cm[-1]), None,
                         1.0, nx, ny, nz, *hdx[-1])


def test_fas_equals_correction_scheme():
    import pamr_amrd_amd as pm
    n = 33
    rho = -pm.gaussian_source(n, seed=11)
    m = _setup(n, rho)
    shapes = [m.shape(l) for l in range(1, m.nlev + 1)]
    hdx = [(1.0 / (s[2] - 1),) * 3 for s in shapes]
    u_cs = np.zeros((n, n, n))
    r_first = None
    for cyc in range(8):
        r_fas = m.vcycle()[0]
        r_cs = _cs_vcycle((u_cs, rho), shapes, hdx)
        r_first = r_first or r_cs
        # equal up to rounding relative to the first residual (the cycles
        # differ only in operation order: FAS carries R u, CS carries e)
        assert abs(r_fas - r_cs) <= 1e-13 * r_first, (cyc, r_fas, r_cs)
    assert np.allclose(m.field(m.nlev, O.F_U), u_cs, rtol=0, atol=1e-12 * np.abs(u_cs).max())


# ---------------------------------------------------------------------------
# MG extrapolation in time (amrd/evolve.c:346-368, 1661-1702)
# ---------------------------------------------------------------------------
def _dyadic(rng, shape):
    return rng.integers(-64, 64, size=shape).astype(np.float64) / 8.0


def test_extrap_etm1_exact_for_linear_in_time():
    """Method 0 is the line through tm1 (at tnp1_lm1 - dt_lm1) and tm2 (at
    tnp1_lm1 - 2 dt_lm1): data linear in time are reproduced exactly at
    tnp1 (dyadic values and times: no rounding)."""
    rng = np.random.default_rng(1)
    a, b = _dyadic(rng, (5, 6, 7)), _dyadic(rng, (5, 6, 7))
    tnp1_lm1, dt_lm1 = 1.5, 0.25
    for dt in (0.125, 0.0625, 0.25):
        tnp1 = tnp1_lm1 - dt_lm1 + dt
        tm1, tm2 = a + b * (tnp1_lm1 - dt_lm1), a + b * (tnp1_lm1 - 2 * dt_lm1)
        got = O.mg_extrap(None, tm1, tm2, 0, tnp1, dt, tnp1_lm1, dt_lm1)
        assert np.array_equal(got, a + b * tnp1)
        # a quadratic is not reproduced (the method is second order)
        q1, q2 = tm1 + (tnp1_lm1 - dt_lm1) ** 2, tm2 + (tnp1_lm1 - 2 * dt_lm1) ** 2
        assert not np.allclose(O.mg_extrap(None, q1, q2, 0, tnp1, dt, tnp1_lm1, dt_lm1), a + b * tnp1 + tnp1 ** 2)


def test_extrap_tn():
    """Method 1 adds dt/dt_lm1 times the difference of the extrapolation
    variables to tl 2: nothing when they agree, half the difference for
    dt = dt_lm1 / 2 (the reference's sign: tm2 - tm1)."""
    rng = np.random.default_rng(2)
    v2, tm1, tm2 = (_dyadic(rng, (4, 5, 6)) for _ in range(3))
    assert np.array_equal(O.mg_extrap(v2, tm1, tm1.copy(), 1, 2.0, 0.125, 1.9, 0.25), v2)
    assert np.array_equal(O.mg_extrap(v2, tm1, tm2, 1, 2.0, 0.125, 1.9, 0.25), v2 + 0.5 * (tm2 - tm1))


def test_update_extrap_rotation_and_eps_correction():
    """eps_c = 0: tm2 <- tm1, tm1 <- v(tl 2) on every level; eps_c = 1 with
    ratio 2: on level L > L1 the correction (tm1 - brs)(1 - eps_c / 2^(L - L1))
    of the reference's comment (amrd/evolve.c:1647-1659), level L1 itself
    skipped when L1 > 1."""
    rng = np.random.default_rng(3)
    rho_tm = [2] * 10
    for L1, eps in ((1, 0.0), (1, 1.0), (2, 1.0), (2, 0.0)):
        levels = {lev: {k: _dyadic(rng, (3, 4, 5)) for k in ("tm1", "tm2", "v2", "brs")} for lev in range(1, 5)}
        before = {lev: {k: v.copy() for k, v in d.items()} for lev, d in levels.items()}
        O.update_mg_extrap_vars(levels, L1, 4, eps, rho_tm)
        first = L1 + 1 if L1 > 1 else L1
        for lev in range(1, 5):
            d, b = levels[lev], before[lev]
            if lev < first:
                for k in d:
                    assert np.array_equal(d[k], b[k])
                continue
            assert np.array_equal(d["tm1"], b["v2"])
            want = b["tm1"]
            if eps and lev > 1:
                want = want + (b["v2"] - b["brs"]) * (1.0 - eps / 2.0 ** (lev - L1))
            assert np.array_equal(d["tm2"], want), (L1, eps, lev)
