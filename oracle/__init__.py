"""CPU oracle for the PAMR/AMRD multigrid hot path -- TEST INFRASTRUCTURE ONLY.

This package is the checker: only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` import it.  It wraps

* ``liboracle.so`` -- the project's plain-C restatement (``mg_oracle.c``) of
  ``amrd/mg.c``, ``src/m_util_r8.f``, the nps/nbs MG hooks and the transfer
  bookkeeping, compiled ``-O2 -ffp-contract=off``;
* ``_ref/*.so`` -- the reference's own Fortran kernels and hooks compiled from
  their sources by ``oracle/Makefile`` (present only where the reference tree
  was available at build time).

Arrays are numpy float64 in Fortran order (x fastest), i.e. numpy shape
``(nz, ny, nx)`` C-contiguous.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = "/root/reference"

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def build(ref: bool | None = None) -> None:
    """Compile liboracle.so (and oracle/_ref when the reference tree exists)."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    if ref is None:
        ref = os.path.isdir(REF_DIR)
    if ref:
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def _load(name):
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        return None
    return C.CDLL(path)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(os.path.join(HERE, "liboracle.so")):
            build(ref=False)
        _lib = C.CDLL(os.path.join(HERE, "liboracle.so"))
        _declare(_lib)
    return _lib


def ref_lib(which: str):
    """which in {'mutil', 'nps', 'nbs'}; None when not built."""
    return _load(os.path.join("_ref", f"libref_{which}.so"))


def P(a):
    """numpy float64 array -> double*  (None -> NULL)."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "need contiguous float64"
    return a.ctypes.data_as(_dp)


def _declare(L):
    i, d = C.c_int, C.c_double
    L.or_dmcopy3d.argtypes = [_dp, _dp] + [i] * 21
    L.or_dvi2q1d.argtypes = [_dp, _dp, i, i, i, i, i, i]
    L.or_dminterp3d.argtypes = [_dp, _dp] + [i] * 21 + [i, _dp, d, i]
    L.or_dmhwr3d.argtypes = [_dp, _dp] + [i] * 21 + [_dp, d, i]
    L.or_dmfwr3d.argtypes = [_dp, _dp] + [i] * 21 + [_dp, d, i]
    L.or_lop.argtypes = [i, _dp, _dp, _dp, _dp, _dp, _dp, _dp, d, i, i, i, d, d, d]
    L.or_residual.argtypes = [i, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, d, i, i, i, d, d, d]
    L.or_residual.restype = d
    L.or_relax.argtypes = [i, _dp, _dp, _dp, _dp, _dp, _dp, _dp, d, i, i, i, d, d, d]
    L.or_relax.restype = d
    L.or_set_cmask_bdy.argtypes = [_dp, _ip, _dp, _dp, _dp, _ip]
    L.or_set_cmask_child.argtypes = [_dp, _ip, _dp, _dp, _dp, _ip, i, _dp]
    L.or_calc_island_no.argtypes = [_dp, i, _dp, _dp, i, i, _ip]
    L.or_calc_gs_ibbox.argtypes = [_dp, _dp, _ip, i, _ip, _dp]
    L.or_mg_create.argtypes = [i, i, i, _dp, i, i, i, d]
    L.or_mg_create.restype = C.c_void_p
    L.or_mg_destroy.argtypes = [C.c_void_p]
    L.or_mg_nlev.argtypes = [C.c_void_p]
    L.or_mg_shape.argtypes = [C.c_void_p, i, _ip]
    L.or_mg_field.argtypes = [C.c_void_p, i, i]
    L.or_mg_field.restype = _dp
    L.or_mg_set_params.argtypes = [C.c_void_p, i, i, i, d, d, d, i, i, i, i]
    L.or_mg_init.argtypes = [C.c_void_p]
    L.or_mg_vcycle.argtypes = [C.c_void_p, _dp, _ip]
    L.or_mg_vcycle.restype = d
    L.or_mg_solve.argtypes = [C.c_void_p, i, i, d, _dp]
    L.or_mg_solve.restype = i
    L.or_mg_residual_fine.argtypes = [C.c_void_p]
    L.or_mg_residual_fine.restype = d
    vp = C.c_void_p
    L.or_amr_create.argtypes = [_ip, _dp, i, _ip, _dp, i, i]
    L.or_amr_create.restype = vp
    L.or_amr_destroy.argtypes = [vp]
    L.or_amr_create_mgh.argtypes = [vp, i]
    L.or_amr_create_mgh.restype = vp
    L.or_amr_grid_coarsest.argtypes = [vp, i, i]
    for f in ("or_amr_nlev", "or_amr_base_level"):
        getattr(L, f).argtypes = [vp]
    L.or_amr_ngrids.argtypes = [vp, i]
    L.or_amr_grid_info.argtypes = [vp, i, i, _ip, _dp]
    L.or_amr_field.argtypes = [vp, i, i, i]
    L.or_amr_field.restype = _dp
    L.or_amr_set_params.argtypes = [vp, i, i, i, d, d, d, i, i, i, i]
    L.or_amr_init.argtypes = [vp]
    L.or_amr_set_tre.argtypes = [vp, i, i]
    L.or_amr_set_excision.argtypes = [vp, i, d]
    L.or_amr_interp_field_ord.argtypes = [vp, i, i, i]
    L.or_amr_vcycle.argtypes = [vp, _dp, _ip]
    L.or_amr_vcycle.restype = d
    L.or_amr_residual_fine.argtypes = [vp]
    L.or_amr_residual_fine.restype = d
    L.or_amr_owned.argtypes = [vp, i, i, i, _ip, _dp, i]
    L.or_amr_transfers.argtypes = [vp, i, i, i, _ip, i]
    L.or_amr_set_cmask.argtypes = [vp, i, i]
    L.or_amr_sync_field.argtypes = [vp, i, i, i]
    L.or_amr_inject_field.argtypes = [vp, i, i, i]
    L.or_amr_interp_field.argtypes = [vp, i, i]
    L.or_net_forward.argtypes = [i] + [_dp] * 8 + [i, i, i]
    L.or_net_loss_grad.argtypes = [i] + [_dp] * 8 + [i, i, i] + [_dp] * 6
    L.or_net_loss_grad.restype = d


# --------------------------------------------------------------------------
# thin helpers
# --------------------------------------------------------------------------
F_U, F_RHO, F_RHS, F_RES, F_LOP, F_CMASK, F_A, F_C, F_CHR = range(9)


def restrict(u_f, op=3, chr_f=None, ex=1.0):
    """Whole-grid restriction fine->coarse (factor 2) via or_dmfwr3d/dmhwr3d/dmcopy3d."""
    nz, ny, nx = u_f.shape
    cx, cy, cz = (nx - 1) // 2 + 1, (ny - 1) // 2 + 1, (nz - 1) // 2 + 1
    out = np.zeros((cz, cy, cx))
    L = lib()
    do_ex = 0 if chr_f is None else 1
    ch = u_f if chr_f is None else chr_f
    args = [P(u_f), P(out), nx, ny, nz, cx, cy, cz, 1, 1, 1, 1, 1, 1, cx, cy, cz, 2, 2, 2, 1, 1, 1]
    if op == 1:
        L.or_dmcopy3d(*args)
    elif op == 2:
        L.or_dmhwr3d(*args, P(ch), ex, do_ex)
    else:
        L.or_dmfwr3d(*args, P(ch), ex, do_ex)
    return out


def interp(u_c, chr_c=None, ex=1.0, ord=2):
    """Whole-grid prolongation coarse->fine via or_dminterp3d: trilinear
    (ord 2) or cubic (ord 4, no excision)."""
    cz, cy, cx = u_c.shape
    nx, ny, nz = 2 * cx - 1, 2 * cy - 1, 2 * cz - 1
    out = np.zeros((nz, ny, nx))
    src = u_c.copy()
    do_ex = 0 if chr_c is None else 1
    ch = src if chr_c is None else chr_c
    lib().or_dminterp3d(P(src), P(out), cx, cy, cz, nx, ny, nz, 1, 1, 1, 1, 1, 1, cx, cy, cz,
                        1, 1, 1, 2, 2, 2, ord, P(ch), ex, do_ex)
    return out


def dminterp3d(u1, u2, args, chr1=None, ex=1.0, do_ex=0):
    """or_dminterp3d with the reference's argument list (args: the 22 ints
    nx1..nz2, i1..k2, ni..nk, si1..sk1, irho..krho, ord); u2 is updated in
    place and returned."""
    ch = u1 if chr1 is None else chr1
    lib().or_dminterp3d(P(u1), P(u2), *[int(a) for a in args], P(ch), ex, do_ex)
    return u2


class UniformMG:
    """Owning wrapper of the oracle's uniform-hierarchy FAS driver."""

    def __init__(self, n, bbox=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0), mcw=5, op=0, do_ex=0, ex=1.0,
                 shape=None):
        self.L = lib()
        nx, ny, nz = shape if shape is not None else (n, n, n)
        bb = np.array(bbox, dtype=np.float64)
        self.h = self.L.or_mg_create(nx, ny, nz, P(bb), mcw, op, do_ex, ex)
        self.nlev = self.L.or_mg_nlev(self.h)

    def __del__(self):
        try:
            if self.h:
                self.L.or_mg_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def shape(self, lev):
        s = (C.c_int * 3)()
        self.L.or_mg_shape(self.h, lev, s)
        return (s[2], s[1], s[0])

    def field(self, lev, which):
        p = self.L.or_mg_field(self.h, lev, which)
        if not p:
            return None
        n = int(np.prod(self.shape(lev)))
        return np.ctypeslib.as_array(p, shape=(n,)).reshape(self.shape(lev))

    def set_params(self, pre=2, pst=2, max_citer=5000, crtol=1e-3, w0_r=1.0, w0_i=1.0,
                   rop_u=3, rop_res=3, rop_src=3, interp_ord=2):
        self.L.or_mg_set_params(self.h, pre, pst, max_citer, crtol, w0_r, w0_i, rop_u, rop_res,
                                rop_src, interp_ord)

    def init(self):
        self.L.or_mg_init(self.h)

    def vcycle(self):
        norms = np.zeros(self.nlev)
        cit = C.c_int(0)
        r = self.L.or_mg_vcycle(self.h, P(norms), C.byref(cit))
        return r, norms, cit.value

    def solve(self, min_iter, max_iter, tol=0.0):
        out = np.zeros(max_iter)
        it = self.L.or_mg_solve(self.h, min_iter, max_iter, tol, P(out))
        return out[:it]

    def residual_fine(self):
        return self.L.or_mg_residual_fine(self.h)


OA_U, OA_RHO, OA_RHS, OA_RES, OA_LOP, OA_CMASK, OA_A, OA_C, OA_TRE, OA_CHR = range(10)


class AmrMG:
    """Owning wrapper of the oracle's multigrid driver over an AMR hierarchy.

    ``levels`` is a list (refined AMR levels 2, 3, ...) of lists of physical
    boxes (x0, x1, y0, y1, z0, z1); the base grid covers ``bbox``."""

    def __init__(self, base_n, levels, bbox=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0), mcw=5, op=0):
        self.L = lib()
        shape = (C.c_int * 3)(base_n, base_n, base_n)
        bb = np.array(bbox, dtype=np.float64)
        nbox = (C.c_int * max(1, len(levels)))(*[len(l) for l in levels])
        boxes = np.array([b for l in levels for b in l] or [[0.0] * 6], dtype=np.float64).reshape(-1)
        self.h = self.L.or_amr_create(shape, P(bb), len(levels), nbox, P(boxes), mcw, op)
        self.nlev = self.L.or_amr_nlev(self.h)
        self.base = self.L.or_amr_base_level(self.h)

    def __del__(self):
        try:
            if self.h:
                self.L.or_amr_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def sub_mgh(self, amr_lev):
        """the MG hierarchy PAMR_build_mgh(amr_lev, finest, tl) builds: island
        levels below AMR level amr_lev, which and the finer AMR levels share
        this hierarchy's grid functions (keep this object alive meanwhile)"""
        sub = object.__new__(AmrMG)
        sub.L = self.L
        sub.h = self.L.or_amr_create_mgh(self.h, amr_lev)
        if not sub.h:
            raise ValueError(f"no MG hierarchy from AMR level {amr_lev}")
        sub.nlev = self.L.or_amr_nlev(sub.h)
        sub.base = self.L.or_amr_base_level(sub.h)
        sub.parent = self
        return sub

    def coarsest(self, lev, g):
        return bool(self.L.or_amr_grid_coarsest(self.h, lev, g))

    def ngrids(self, lev):
        return self.L.or_amr_ngrids(self.h, lev)

    def grid(self, lev, g):
        """(shape (nz, ny, nx), bbox) of grid g of MG level lev."""
        s = (C.c_int * 3)()
        b = np.zeros(6)
        self.L.or_amr_grid_info(self.h, lev, g, s, P(b))
        return (s[2], s[1], s[0]), b

    def field(self, lev, g, which):
        p = self.L.or_amr_field(self.h, lev, g, which)
        if not p:
            return None
        shape = self.grid(lev, g)[0]
        return np.ctypeslib.as_array(p, shape=(int(np.prod(shape)),)).reshape(shape)

    def set_params(self, pre=2, pst=2, max_citer=5000, crtol=1e-3, w0_r=1.0, w0_i=1.0,
                   rop_u=3, rop_res=3, rop_src=3, interp_ord=2):
        self.L.or_amr_set_params(self.h, pre, pst, max_citer, crtol, w0_r, w0_i, rop_u, rop_res,
                                 rop_src, interp_ord)

    def init(self):
        self.L.or_amr_init(self.h)

    def set_excision(self, ex, do_ex=True):
        """PAMR's excision: OA_CHR of every grid is the mask (value ex =
        excised); fill it on every level before init()"""
        self.L.or_amr_set_excision(self.h, 1 if do_ex else 0, ex)

    def set_tre(self, save_tre, ntre=1):
        self.L.or_amr_set_tre(self.h, 1 if save_tre else 0, ntre)

    def vcycle(self):
        norms = np.zeros(self.nlev)
        cit = C.c_int(0)
        r = self.L.or_amr_vcycle(self.h, P(norms), C.byref(cit))
        return r, norms, cit.value

    def residual_fine(self):
        return self.L.or_amr_residual_fine(self.h)

    def owned(self, lev, interior, extend):
        cap = 4096
        gi = (C.c_int * cap)()
        bb = np.zeros(6 * cap)
        n = self.L.or_amr_owned(self.h, lev, interior, extend, gi, P(bb), cap)
        assert n <= cap
        return [(gi[k], bb[6 * k:6 * k + 6].copy()) for k in range(n)]

    def transfers(self, kind, lev, arg=0):
        """rows of 16 ints: src g, dst g, delta, irho, src ibbox[6], dst ibbox[6]"""
        cap = 1 << 16
        out = (C.c_int * (16 * cap))()
        n = self.L.or_amr_transfers(self.h, kind, lev, arg, out, cap)
        assert n <= cap
        return np.array(out[:16 * n], dtype=np.int64).reshape(n, 16)

    def set_cmask(self, lev, which):
        self.L.or_amr_set_cmask(self.h, lev, which)

    def sync_field(self, lev, bdy, which):
        self.L.or_amr_sync_field(self.h, lev, bdy, which)

    def inject_field(self, lev_fine, which, op=3):
        self.L.or_amr_inject_field(self.h, lev_fine, which, op)

    def interp_field(self, lev_coarse, which, ord=2):
        self.L.or_amr_interp_field_ord(self.h, lev_coarse, which, ord)


# --------------------------------------------------------------------------
# MG extrapolation in time (numpy restatement, elementwise FP64)
# --------------------------------------------------------------------------
def mg_extrap(v2, tm1, tm2, method, tnp1, dt, tnp1_lm1, dt_lm1):
    """v(tl 1) of an elliptic variable on a level > 1 (amrd/evolve.c:346-368):
    method 0 (AMRD_MG_EXTRAP_2ND_FROM_ETM1) the line through tm1 at
    tnp1_lm1 - dt_lm1 and tm2 at tnp1_lm1 - 2 dt_lm1, evaluated at tnp1;
    method 1 (2ND_FROM_TN) v2 + dt/dt_lm1 (tm2 - tm1)."""
    if method == 0:
        f1 = (tnp1 - (tnp1_lm1 - 2 * dt_lm1)) / dt_lm1
        f2 = -(tnp1 - (tnp1_lm1 - dt_lm1)) / dt_lm1
        return f1 * tm1 + f2 * tm2
    return v2 + (dt / dt_lm1) * (tm2 - tm1)


def update_mg_extrap_vars(levels, L1, Lf, eps_c, rho_tm):
    """amrd/evolve.c:1661-1702 over ``levels[lev] = {"tm1", "tm2", "v2",
    "brs"}`` (numpy arrays, updated in place); rho_tm[l-1] is the time
    refinement ratio below level l."""
    net_rho = 1
    if L1 > 1:
        net_rho = rho_tm[L1 - 1]
        L1 += 1
    for lev in range(L1, Lf + 1):
        d = levels.get(lev)
        if d is not None:
            d["tm2"][...] = d["tm1"]
            d["tm1"][...] = d["v2"]
            if lev > 1 and eps_c != 0:
                f = 1.0 - eps_c / float(net_rho)
                d["tm2"][...] = d["tm2"] + (d["tm1"] - d["brs"]) * f
        net_rho *= rho_tm[L1 - 1]


def net_forward(w1, b1, w2, b2, w3, b3, x):
    H = b1.shape[0]
    n2, n1, n0 = x.shape
    y = np.zeros_like(x)
    lib().or_net_forward(H, P(w1), P(b1), P(w2), P(b2), P(w3), P(b3), P(x), P(y), n0, n1, n2)
    return y


def net_loss_grad(w1, b1, w2, b2, w3, b3, x, t):
    H = b1.shape[0]
    n2, n1, n0 = x.shape
    g = [np.zeros_like(a) for a in (w1, b1, w2, b2, w3, b3)]
    loss = lib().or_net_loss_grad(H, P(w1), P(b1), P(w2), P(b2), P(w3), P(b3), P(x), P(t),
                                  n0, n1, n2, *[P(a) for a in g])
    return loss, g
