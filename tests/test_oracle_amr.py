"""Pins of the multi-patch oracle (oracle/amr_oracle.c): the multigrid driver
over an AMR hierarchy and PAMR's sync / inject / interp bookkeeping.

* A refined level that covers the whole domain turns the AMR hierarchy into a
  uniform one: per-cycle norms and the solution must equal the uniform
  oracle's bit for bit (which is itself pinned in test_oracle_closed_forms).
* sync (src/pamr.c:601-617): brute force over the lattice -- every point takes
  the value of the first grid (level order) whose interior, reduced by the
  AMR boundary width at non-physical faces, contains it.
* interp (src/pamr.c:623-638): a trilinear polynomial with dyadic
  coefficients on the coarse grids is reproduced exactly at every fine point.
* inject (src/pamr.c:667-685): an affine field on the fine grids is
  reproduced exactly at every coarse point inside a fine grid, and no other
  coarse point is written.
"""
import numpy as np
import pytest

import oracle as O
from pamr_amrd_amd import hierarchy as H


def small_levels(seed, base_n=17, nref=2):
    return H.random_boxes(base_n=base_n, nref=nref, nmin=3, nmax=6, smin=5, smax=13, seed=seed)


def lattice_box(bbox, dx):
    """integer lattice box (x0, x1, y0, y1, z0, z1) of a physical bbox"""
    return [int(round(v / dx)) for v in bbox]


def test_full_domain_refinement_equals_uniform():
    n = 33
    m = O.AmrMG(17, [[(0.0, 1.0, 0.0, 1.0, 0.0, 1.0)]])
    assert m.nlev == 4 and m.base == 3
    u = O.UniformMG(n)
    assert u.nlev == 4
    rho = -H.gaussian_on((n, n, n), (0, 1, 0, 1, 0, 1))
    rho[0, :, :] = rho[-1, :, :] = rho[:, 0, :] = rho[:, -1, :] = rho[:, :, 0] = rho[:, :, -1] = 0
    u.field(4, O.F_RHO)[:] = rho
    m.field(4, 0, O.OA_RHO)[:] = rho
    m.field(3, 0, O.OA_RHO)[:] = O.restrict(rho, 3)  # AMR data of the base level
    for x in (u, m):
        x.set_params()
        x.init()
    for c in range(5):
        ru, nu, cu = u.vcycle()
        ra, na, ca = m.vcycle()
        assert ru == ra and cu == ca, (c, ru, ra, cu, ca)
        assert np.array_equal(nu, na)
    assert np.array_equal(u.field(4, O.F_U), m.field(4, 0, O.OA_U))
    assert np.array_equal(u.field(2, O.F_U), m.field(2, 0, O.OA_U))


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
@pytest.mark.parametrize("bdy", [0, 1])
def test_sync_ownership_brute_force(seed, bdy):
    levels = small_levels(seed)
    m = O.AmrMG(17, levels)
    L = m.nlev  # finest refined level
    dx = 1.0 / 16 / 4
    rng = np.random.default_rng(seed)
    grids = []
    for g in range(m.ngrids(L)):
        shape, bb = m.grid(L, g)
        f = m.field(L, g, O.OA_U)
        f[:] = rng.standard_normal(shape)
        grids.append((lattice_box(bb, dx), f.copy()))
    lim = int(round(1.0 / dx))

    def reduced(box):
        r = list(box)
        for d in range(3):
            if r[2 * d] > 0:
                r[2 * d] += bdy
            if r[2 * d + 1] < lim:
                r[2 * d + 1] -= bdy
        return r

    def inside(box, p):
        return all(box[2 * d] <= p[d] <= box[2 * d + 1] for d in range(3))

    m.sync_field(L, bdy, O.OA_U)
    red = [reduced(b) for b, _ in grids]
    for g, (box, before) in enumerate(grids):
        after = m.field(L, g, O.OA_U)
        for k in range(box[4], box[5] + 1):
            for j in range(box[2], box[3] + 1):
                for i in range(box[0], box[1] + 1):
                    p = (i, j, k)
                    owner = next((h for h in range(len(grids)) if inside(red[h], p)), None)
                    loc = (k - box[4], j - box[2], i - box[0])
                    if owner is None or owner == g:
                        want = before[loc]
                    else:
                        ob = grids[owner][0]
                        want = grids[owner][1][k - ob[4], j - ob[2], i - ob[0]]
                    assert after[loc] == want, (g, p, owner)


def poly_trilinear(X, Y, Z):
    return 1.0 + 2.0 * X - 3.0 * Y + 0.5 * Z + 4.0 * X * Y - 2.0 * Y * Z + X * Z + 8.0 * X * Y * Z


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_interp_exact_for_trilinear(seed):
    m = O.AmrMG(17, small_levels(seed))
    for Lc in (m.base, m.base + 1):
        for g in range(m.ngrids(Lc)):
            shape, bb = m.grid(Lc, g)
            m.field(Lc, g, O.OA_RES)[:] = poly_trilinear(*H.grid_points(shape, bb))
        for g in range(m.ngrids(Lc + 1)):
            m.field(Lc + 1, g, O.OA_RES)[:] = np.nan
        m.interp_field(Lc, O.OA_RES)
        for g in range(m.ngrids(Lc + 1)):
            shape, bb = m.grid(Lc + 1, g)
            got = m.field(Lc + 1, g, O.OA_RES)
            assert not np.isnan(got).any(), "fine point not reached by interp"
            assert np.array_equal(got, poly_trilinear(*H.grid_points(shape, bb)))


def poly_tricubic(X, Y, Z):
    return (1.0 + X - 2.0 * X ** 2 + X ** 3) * (2.0 - Y + 0.5 * Y ** 3) * (1.0 + Z ** 2 - Z ** 3)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_interp4_exact_for_tricubic(seed):
    """PAMR_interp at PAMR_FOURTH_ORDER (dvi4q1d on the buffered temporary
    of every transfer): a tensor-product cubic on the coarse grids is
    reproduced at every fine point to rounding (the coefficients hold 1/6)."""
    m = O.AmrMG(17, small_levels(seed))
    for Lc in (m.base, m.base + 1):
        for g in range(m.ngrids(Lc)):
            shape, bb = m.grid(Lc, g)
            m.field(Lc, g, O.OA_RES)[:] = poly_tricubic(*H.grid_points(shape, bb))
        for g in range(m.ngrids(Lc + 1)):
            m.field(Lc + 1, g, O.OA_RES)[:] = np.nan
        m.interp_field(Lc, O.OA_RES, ord=4)
        for g in range(m.ngrids(Lc + 1)):
            shape, bb = m.grid(Lc + 1, g)
            got = m.field(Lc + 1, g, O.OA_RES)
            want = poly_tricubic(*H.grid_points(shape, bb))
            assert not np.isnan(got).any(), "fine point not reached by interp"
            assert np.allclose(got, want, rtol=0, atol=1e-13), np.max(np.abs(got - want))


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("op", [1, 2, 3])
def test_inject_exact_for_affine(seed, op):
    m = O.AmrMG(17, small_levels(seed))
    aff = lambda X, Y, Z: 0.25 + 2.0 * X - 1.5 * Y + 3.0 * Z
    for Lf in (m.base + 1, m.base + 2):
        Lc = Lf - 1
        dxc = 1.0 / 16 / (1 << (Lc - m.base))
        fine_boxes = []
        for g in range(m.ngrids(Lf)):
            shape, bb = m.grid(Lf, g)
            m.field(Lf, g, O.OA_U)[:] = aff(*H.grid_points(shape, bb))
            fine_boxes.append(lattice_box(bb, dxc / 2))
        for g in range(m.ngrids(Lc)):
            m.field(Lc, g, O.OA_U)[:] = np.nan
        m.inject_field(Lf, O.OA_U, op)
        for g in range(m.ngrids(Lc)):
            shape, bb = m.grid(Lc, g)
            got = m.field(Lc, g, O.OA_U)
            want = aff(*H.grid_points(shape, bb))
            cb = lattice_box(bb, dxc)
            for k in range(shape[0]):
                for j in range(shape[1]):
                    for i in range(shape[2]):
                        p = (2 * (cb[0] + i), 2 * (cb[2] + j), 2 * (cb[4] + k))
                        covered = any(all(b[2 * d] <= p[d] <= b[2 * d + 1] for d in range(3))
                                      for b in fine_boxes)
                        v = got[k, j, i]
                        if covered:
                            assert v == want[k, j, i], (Lf, g, i, j, k, v, want[k, j, i])
                        else:
                            assert np.isnan(v), (Lf, g, i, j, k)


def test_owned_segments_partition():
    """owned pieces are disjoint and cover the union of the grids"""
    m = O.AmrMG(17, small_levels(5))
    L = m.nlev
    dx = 1.0 / 64
    segs = m.owned(L, 1, 0)
    cover = {}
    for g, bb in segs:
        b = lattice_box(bb, dx)
        for k in range(b[4], b[5] + 1):
            for j in range(b[2], b[3] + 1):
                for i in range(b[0], b[1] + 1):
                    assert (i, j, k) not in cover, "owned pieces overlap"
                    cover[(i, j, k)] = g
    union = set()
    for g in range(m.ngrids(L)):
        b = lattice_box(m.grid(L, g)[1], dx)
        union |= {(i, j, k) for k in range(b[4], b[5] + 1) for j in range(b[2], b[3] + 1)
                  for i in range(b[0], b[1] + 1)}
    assert set(cover) == union


def _lop0(u, rho, h):
    """interior of Delta_h u - rho (op 0), zero on the boundary"""
    out = np.zeros_like(u)
    c = u[1:-1, 1:-1, 1:-1]
    lap = ((u[2:, 1:-1, 1:-1] - 2 * c + u[:-2, 1:-1, 1:-1]) + (u[1:-1, 2:, 1:-1] - 2 * c + u[1:-1, :-2, 1:-1])
           + (u[1:-1, 1:-1, 2:] - 2 * c + u[1:-1, 1:-1, :-2])) / (h * h)
    out[1:-1, 1:-1, 1:-1] = lap - rho[1:-1, 1:-1, 1:-1]
    return out


def _prolong(c):
    """vertex-centred 2:1 trilinear prolongation"""
    n = 2 * (c.shape[0] - 1) + 1
    f = np.zeros((n, n, n))
    f[::2, ::2, ::2] = c
    f[1::2, :, :][:, ::2, ::2] = 0.5 * (f[:-1:2, ::2, ::2] + f[2::2, ::2, ::2])
    f[:, 1::2, :][:, :, ::2] = 0.5 * (f[:, :-1:2, ::2] + f[:, 2::2, ::2])
    f[:, :, 1::2] = 0.5 * (f[:, :, :-1:2] + f[:, :, 2::2])
    return f


@pytest.mark.parametrize("ntre", [1, 4])
def test_tre_closed_form(ntre):
    """compute_MG_rhs's truncation-error estimate (amrd/mg.c:288-295,306-313)
    on a uniform MG hierarchy 33 -> 17 -> 9 -> 5 with no sweeps (pre = pst =
    max_citer = 0), so every level keeps u_L = R u_{L+1} after the cycle:
    level 1 holds |L_1 u_1| sqrt(ntre) (Laplacian minus source, zero on the
    boundary); every finer level L holds the trilinear interpolation of level
    L-1's own estimate, which overwrote what compute_MG_rhs(L) wrote there."""
    m = O.AmrMG(33, [])
    assert m.nlev == 4
    rng = np.random.default_rng(3)
    n = 33
    x = np.linspace(0, 1, n)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    u0 = np.sin(np.pi * X) * np.sin(2 * np.pi * Y) * np.sin(np.pi * Z) + 0.01 * rng.standard_normal((n, n, n))
    u0[0], u0[-1], u0[:, 0], u0[:, -1], u0[:, :, 0], u0[:, :, -1] = 0, 0, 0, 0, 0, 0
    m.field(4, 0, O.OA_U)[:] = u0
    m.field(4, 0, O.OA_RHO)[:] = -H.gaussian_on((n, n, n), (0, 1, 0, 1, 0, 1), seed=9)
    m.set_params(pre=0, pst=0, max_citer=0)
    m.set_tre(True, ntre)
    m.init()
    m.vcycle()
    assert np.array_equal(m.field(4, 0, O.OA_U), u0)
    own = {}
    for L in range(1, 4):
        nL = m.grid(L, 0)[0][0]
        own[L] = np.abs(_lop0(m.field(L, 0, O.OA_U), m.field(L, 0, O.OA_RHO), 1.0 / (nL - 1))) * np.sqrt(ntre)
    t1 = m.field(1, 0, O.OA_TRE)
    assert np.allclose(t1, own[1], rtol=1e-13, atol=0), np.max(np.abs(t1 - own[1]))
    assert np.max(t1) > 0
    for L in range(2, 5):
        want = _prolong(own[L - 1])
        got = m.field(L, 0, O.OA_TRE)
        assert np.allclose(got, want, rtol=1e-13, atol=1e-300), (L, np.max(np.abs(got - want)))


@pytest.mark.parametrize("seed", [1, 2])
def test_tre_is_diagnostic(seed):
    """save_tre changes neither the solution nor the norms; the estimate is
    non-negative and scales with sqrt(ntre) (x2 for ntre = 4, exactly)."""
    runs = {}
    for key, (on, ntre) in {"off": (False, 1), "one": (True, 1), "four": (True, 4)}.items():
        levels = small_levels(seed)
        m = O.AmrMG(17, levels)
        for L in range(m.base, m.nlev + 1):
            for g in range(m.ngrids(L)):
                shape, bb = m.grid(L, g)
                m.field(L, g, O.OA_RHO)[:] = -H.gaussian_on(shape, bb, seed=seed)
        m.set_params()
        m.set_tre(on, ntre)
        m.init()
        norms = [m.vcycle()[1] for _ in range(3)]
        u = {(L, g): m.field(L, g, O.OA_U).copy() for L in range(1, m.nlev + 1) for g in range(m.ngrids(L))}
        t = {(L, g): m.field(L, g, O.OA_TRE).copy() for L in range(1, m.nlev + 1) for g in range(m.ngrids(L))}
        runs[key] = (norms, u, t)
    for a, b in zip(runs["off"][0], runs["one"][0]):
        assert np.array_equal(a, b)
    for k in runs["off"][1]:
        assert np.array_equal(runs["off"][1][k], runs["one"][1][k])
        assert np.all(runs["one"][2][k] >= 0)
        assert np.array_equal(runs["four"][2][k], 2.0 * runs["one"][2][k])
        assert not np.any(runs["off"][2][k])
    assert any(np.any(v > 0) for v in runs["one"][2].values())
