"""The oracle's m_util_r8 and hook restatements against golden vectors
written by the reference's own Fortran (tests/golden/make_golden.py).
All comparisons are bit-exact."""
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_kernels.npz")


@pytest.fixture(scope="module")
def g():
    return np.load(GOLD, allow_pickle=False)


def _restr(u, op, chr_=None):
    return O.restrict(u, op=op, chr_f=chr_, ex=1.0)


@pytest.mark.parametrize("tag", ["a", "b"])
def test_restriction_matches_reference(g, tag):
    u = g[f"restr_{tag}_u"]
    chr_ = g[f"restr_{tag}_chr"]
    assert np.array_equal(_restr(u, 3), g[f"restr_{tag}_fw0"])
    assert np.array_equal(_restr(u, 3, chr_), g[f"restr_{tag}_fw1"])
    assert np.array_equal(_restr(u, 2), g[f"restr_{tag}_hw0"])
    assert np.array_equal(_restr(u, 2, chr_), g[f"restr_{tag}_hw1"])
    assert np.array_equal(_restr(u, 1), g[f"restr_{tag}_copy"])


@pytest.mark.parametrize("tag", ["a", "b"])
def test_interpolation_matches_reference(g, tag):
    u = g[f"interp_{tag}_u"]
    chr_ = g[f"interp_{tag}_chr"]
    assert np.array_equal(O.interp(u), g[f"interp_{tag}_ex0"])
    assert np.array_equal(O.interp(u, chr_c=chr_, ex=1.0), g[f"interp_{tag}_ex1"])


def test_hooks_match_reference(g):
    L = O.lib()
    V, rho, rhs, cm = g["hook_V"], g["hook_rho"], g["hook_rhs"], g["hook_cmask"]
    x, y, z = g["hook_x"], g["hook_y"], g["hook_z"]
    nz, ny, nx = V.shape
    dx, dy, dz = x[1] - x[0], y[1] - y[0], z[1] - z[0]
    Vr = V.copy()
    n = L.or_relax(0, O.P(Vr), O.P(rhs), O.P(rho), None, None, O.P(cm), None, 1.0, nx, ny, nz, dx, dy, dz)
    assert np.array_equal(Vr, g["hook_relax_V"])
    assert n == g["hook_relax_norm"][0]
    res = np.zeros_like(V)
    n = L.or_residual(0, O.P(res), O.P(rhs), O.P(V), O.P(rho), None, None, O.P(cm), None, 1.0, nx, ny, nz,
                      dx, dy, dz)
    assert np.array_equal(res, g["hook_residual_res"])
    assert n == g["hook_residual_norm"][0]
    lv = np.zeros_like(V)
    L.or_lop(0, O.P(lv), O.P(V), O.P(rho), None, None, O.P(cm), None, 1.0, nx, ny, nz, dx, dy, dz)
    assert np.array_equal(lv, g["hook_lop"])
    # nbs relax: source phi_r^2 + phi_i^2 with phi_i = 0
    phr = g["hook_nbs_phr"]
    Vb = V.copy()
    src = phr * phr
    n = L.or_relax(0, O.P(Vb), O.P(rhs), O.P(src), None, None, O.P(cm), None, 1.0, nx, ny, nz, dx, dy, dz)
    assert np.array_equal(Vb, g["hook_nbs_relax_V"])
    assert n == g["hook_nbs_relax_norm"][0]


@pytest.mark.parametrize("tag", ["a", "b"])
def test_cubic_interpolation_matches_reference(g, tag):
    """ord = 4 (dvi4q1d) into a pre-filled target: interior segment (tag a)
    and whole grid (tag b)"""
    u1 = g[f"interp4_{tag}_u1"].copy()
    u2 = g[f"interp4_{tag}_u2"].copy()
    O.dminterp3d(u1, u2, g[f"interp4_{tag}_args"])
    assert np.array_equal(u2, g[f"interp4_{tag}_out"])


@pytest.mark.parametrize("tag", ["a", "b"])
def test_cubic_interpolation_with_excision_matches_reference(g, tag):
    """ord = 4 with excision (dvi4q1d_ex): the fine result and the source
    array, which the reference also modifies (dminterp3d's NOTE)"""
    u1 = g[f"interp4ex_{tag}_u1"].copy()
    u2 = g[f"interp4ex_{tag}_u2"].copy()
    O.dminterp3d(u1, u2, g[f"interp4ex_{tag}_args"], chr1=g[f"interp4ex_{tag}_chr"], ex=1.0, do_ex=1)
    assert np.array_equal(u2, g[f"interp4ex_{tag}_out"])
    assert np.array_equal(u1, g[f"interp4ex_{tag}_u1_after"])
