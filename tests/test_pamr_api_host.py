"""Host logic of the PAMR / AMRD calls (include/pamr.h, include/amrd_mg.h)
that needs no GPU: contexts, the variable table and its gfn numbering, the
attributes AMRD's MG wiring gives its variables, level spacings, and the
error convention (1 success, 0 failure).

Pins (each a rule stated in the reference, checked on the values, not by
restating the code):
  * a variable takes num_tl AMR gfns and then one MG gfn, from 1 in
    definition order (src/pamr.c:1126-1170, 1735-1760);
  * AMRD relies on cmask's MG gfn being its AMR gfn + 1 (amrd/io.c:997) and
    on AMR time level k of an elliptic variable being its MG gfn - num_tl +
    k - 1 (amrd/io.c:848-849);
  * <v>_res restricts half-weight unless listed in mg_fw_restr, and
    interpolates second order when v is in mg_interp2 (amrd/io.c:853-861);
  * dx of level l is dx_1 / 2^(l-1), dt_1 = lambda * min dx_1
    (src/gh.c:55-79).
"""
import ctypes as C

import numpy as np
import pytest

import pamr_amrd_amd as pm

AMRH, MGH = 1, 2


@pytest.fixture
def L():
    lib = pm.lib()
    shape = (C.c_int * 3)(33, 17, 9)
    bb = np.array([0.0, 1.0, 0.0, 0.5, 0.0, 0.25])
    assert lib.PAMR_init_context(None, 0, 3, shape, pm.P(bb), None, 0) == 1
    yield lib
    lib.PAMR_free_context(1)


def defvar(lib, name, in_amrh, in_mgh, num_tl, mg_inject=0, mg_sync=0, noinj=0):
    pbt = (C.c_int * 6)()
    return lib.PAMR_def_var_full(name.encode(), in_amrh, in_mgh, num_tl, 0, 0, 0, 0, mg_inject, 0, mg_sync, noinj,
                                 0, pbt)


def attribs(lib, name):
    v = [C.c_int() for _ in range(12)]
    pbt = (C.c_int * 6)()
    sg = lib.PAMR_get_var_attribs(name.encode(), *[C.byref(x) for x in v], pbt)
    keys = ("in_amrh in_mgh num_tl amr_inject amr_interp amr_bdy_interp amr_sync mg_inject mg_interp mg_sync "
            "mg_noinj_to_amr regrid_transfer").split()
    return sg, dict(zip(keys, [x.value for x in v]))


def test_gfn_numbering(L):
    assert defvar(L, "V", 1, 1, 3) == 1           # AMR gfns 1..3, MG gfn 4
    assert defvar(L, "rho", 1, 1, 1) == 5         # 5, MG 6
    assert defvar(L, "w", 0, 1, 0) == 7           # MG only
    assert defvar(L, "e", 1, 0, 2) == 8           # AMR only: 8, 9
    assert defvar(L, "z", 1, 1, 1) == 10
    assert [L.PAMR_get_gfn(b"V", AMRH, k) for k in (1, 2, 3)] == [1, 2, 3]
    assert L.PAMR_get_gfn(b"V", MGH, 0) == 4
    mg = L.PAMR_get_gfn(b"V", MGH, 0)
    assert [mg - 3 + k - 1 for k in (1, 2, 3)] == [1, 2, 3]  # amrd/io.c:848-849
    assert L.PAMR_get_gfn(b"rho", MGH, 0) == 6
    assert L.PAMR_get_gfn(b"w", MGH, 0) == 7
    assert [L.PAMR_get_gfn(b"e", AMRH, k) for k in (1, 2)] == [8, 9]
    assert L.PAMR_get_gfn(b"z", MGH, 0) == 11


def test_errors_return_zero(L):
    assert defvar(L, "V", 1, 1, 1) == 1
    assert defvar(L, "V", 1, 1, 1) == 0            # exists
    assert defvar(L, "m", 0, 1, 2) == 0            # MG variable with time levels
    assert defvar(L, "n", 1, 1, -1) == 0           # AMR variable, negative num_tl
    assert L.PAMR_get_gfn(b"nope", AMRH, 1) == 0
    assert defvar(L, "e", 1, 0, 1) > 0
    assert L.PAMR_get_gfn(b"e", MGH, 0) == 0       # not in the MG hierarchy
    assert L.PAMR_get_gfn(b"V", AMRH, 2) == 0      # invalid time level
    assert L.PAMR_get_gfn(b"V", 7, 1) == 0         # unknown hierarchy
    assert L.PAMR_init_s_iter(1, MGH, 0) == 0      # no MG hierarchy
    assert L.PAMR_get_max_lev(AMRH) == 0           # no AMR hierarchy
    assert L.PAMR_next_g() == 0
    assert L.PAMR_pop_iter() == 0                  # empty stack
    assert L.PAMR_build_mgh(1, 1, 1) == 0          # no AMR hierarchy
    shape = (C.c_int * 3)(9, 9, 9)
    bb = np.array([0.0, 1.0] * 3)
    assert L.PAMR_init_context(None, 0, 3, shape, pm.P(bb), None, 0) == 0  # second context
    assert L.PAMR_init_context(None, 0, 2, shape, pm.P(bb), None, 0) == 0  # dim 2


def test_amrd_mg_wiring(L):
    assert defvar(L, "phi", 1, 1, 1, mg_inject=3, mg_sync=1) > 0
    assert defvar(L, "psi", 1, 1, 2, mg_inject=3, mg_sync=1) > 0
    assert L.amrd_mg_def_vars(b"phi, psi", b"phi_res", b"psi", 0) == 1
    _, a = attribs(L, "phi_res")
    assert (a["in_amrh"], a["in_mgh"], a["num_tl"]) == (0, 1, 0)
    assert (a["mg_inject"], a["mg_interp"], a["mg_sync"]) == (3, 0, 0)   # FW listed, no interp2
    _, a = attribs(L, "psi_res")
    assert (a["mg_inject"], a["mg_interp"]) == (2, 2)                     # HW default, interp2 listed
    for n in ("phi_lop", "phi_rhs", "psi_lop", "psi_rhs"):
        _, a = attribs(L, n)
        assert (a["in_amrh"], a["in_mgh"], a["mg_inject"], a["mg_interp"], a["mg_sync"]) == (0, 1, 0, 0, 0), n
    _, a = attribs(L, "cmask")
    assert (a["in_amrh"], a["in_mgh"], a["num_tl"], a["mg_noinj_to_amr"], a["mg_inject"]) == (1, 1, 1, 1, 0)
    assert L.PAMR_get_gfn(b"cmask", MGH, 0) == L.PAMR_get_gfn(b"cmask", AMRH, 1) + 1   # amrd/io.c:997
    assert L.PAMR_get_gfn(b"chr", AMRH, 1) == 0                           # only with excision
    assert L.amrd_mg_def_vars(b"undefined", b"", b"", 0) == 0


def test_tre_vars(L):
    """amrd/io.c:810-838: <v>_tre in the AMR hierarchy with one time level and
    a second-order regrid transfer; only the first also in the MG hierarchy
    (where the MG estimate goes)."""
    assert L.amrd_def_tre_vars(b"phi, psi") == 1
    _, a = attribs(L, "phi_tre")
    assert (a["in_amrh"], a["in_mgh"], a["num_tl"], a["mg_noinj_to_amr"], a["regrid_transfer"]) == (1, 1, 1, 1, 2)
    assert (a["amr_inject"], a["amr_interp"], a["amr_sync"], a["mg_inject"], a["mg_interp"], a["mg_sync"]) == \
        (0, 0, 0, 0, 0, 0)
    _, a = attribs(L, "psi_tre")
    assert (a["in_amrh"], a["in_mgh"], a["num_tl"], a["regrid_transfer"]) == (1, 0, 1, 2)
    assert L.PAMR_get_gfn(b"phi_tre", MGH, 0) == L.PAMR_get_gfn(b"phi_tre", AMRH, 1) + 1
    assert L.PAMR_get_gfn(b"psi_tre", MGH, 0) == 0
    assert L.amrd_def_tre_vars(b"phi") == 0       # phi_tre exists


def test_level_spacing(L):
    lam = 0.25
    L.PAMR_set_lambda(lam)
    dx, dt = np.zeros(3), C.c_double()
    for lev in (1, 2, 5):
        L.PAMR_get_dxdt(lev, pm.P(dx), C.byref(dt))
        f = 2.0 ** (lev - 1)
        assert np.array_equal(dx, np.array([1.0 / 32, 0.5 / 16, 0.25 / 8]) / f)
        assert dt.value == lam * (1.0 / 32) / f
    w = (C.c_int * 3)()
    L.PAMR_get_MG_coarse_width(w)
    assert list(w) == [5, 5, 5]
    bb = np.zeros(6)
    L.PAMR_get_global_bbox(pm.P(bb))
    assert np.array_equal(bb, [0.0, 1.0, 0.0, 0.5, 0.0, 0.25])
