"""Driving the PAMR / AMRD calls (include/pamr.h, include/amrd_mg.h) the way a
C application does, with the nps example hooks (apps/nps_hooks.c).

``NpsContext`` is the sequence of an application's main program: init the
context, define the variables (nps_define_vars: V, rho [, a, c] and the MG
variables AMRD adds), compose the hierarchy from per-level boxes, register
the hooks (nps_init).  Grid functions are reached through the PAMR iterator
and getters -- device pointers -- and moved with the app's copy helpers.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import HERE, P, lib

APP_PATH = os.path.join(HERE, "apps", "libnps_hooks.so")
PAMR_AMRH, PAMR_MGH = 1, 2
MAX_GFNS = 512

_app = None
# void (*app_fill_ex_mask_fnc)(real *mask, int dim, int *shape, real *bbox, real excised)
FILL_EX_MASK = C.CFUNCTYPE(None, C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double),
                           C.c_double)


def nps_lib():
    """The example application's hooks (built by the package build)."""
    global _app
    if _app is None:
        lib()  # the application links libpamr_mg.so: load it first
        if not os.path.exists(APP_PATH):
            raise RuntimeError(f"{APP_PATH} is not built")
        A = C.CDLL(APP_PATH)
        i, vp, dp = C.c_int, C.c_void_p, C.POINTER(C.c_double)
        A.nps_define_vars.argtypes = [i, i, i, i]
        A.nps_define_vars_tl.argtypes = [i, i, i, i, i]
        A.nps_define_vars_ex.argtypes = [i, i, i, i, C.c_double]
        A.nps_init.argtypes = []
        A.nps_set_level_hooks.argtypes = [i]
        A.nps_set_level_hooks.restype = None
        A.nps_release.argtypes = []
        A.nps_release.restype = None
        A.nps_upload.argtypes = [vp, dp, C.c_long]
        A.nps_download.argtypes = [dp, vp, C.c_long]
        _app = A
    return _app


class NpsContext:
    """A PAMR context with the nps variables over base_n^3 on bbox plus the
    refined levels' boxes (``levels[l]``: boxes of AMR level l + 2)."""

    def __init__(self, base_n, levels, op=0, rop_u=3, rop_res=3, rop_src=3, mcw=5,
                 bbox=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0), tre_vars=None, num_tl=1, extrap=False, extra_defs=None,
                 excision=None):
        """excision: (ex, fill) -- AMRD's chr variable is defined, the hooks
        skip chr == ex, and PAMR_excision_on("chr", ...) is called before the
        hierarchy is composed with a fill function that writes
        fill((nz, ny, nx), bbox) (numpy, ex = excised) into each grid's mask."""
        self.L, self.app = lib(), nps_lib()
        L = self.L
        shape = (C.c_int * 3)(base_n, base_n, base_n)
        self.bbox = np.array(bbox, dtype=np.float64)
        if not L.PAMR_init_context(None, 0, 3, shape, P(self.bbox), None, 0):
            raise RuntimeError("PAMR_init_context failed")
        self.open = True
        w = (C.c_int * 3)(mcw, mcw, mcw)
        L.PAMR_set_MG_coarse_width(w)
        if excision is not None:
            ex, fill = excision
            if num_tl != 1 or not self.app.nps_define_vars_ex(op, rop_u, rop_res, rop_src, ex):
                raise RuntimeError("nps_define_vars_ex failed")

            def _fill(mask, dim, shape, bb, excised):
                shp = (shape[2], shape[1], shape[0])
                box = np.array([bb[k] for k in range(6)])
                vals = np.ascontiguousarray(fill(shp, box), dtype=np.float64)
                C.memmove(mask, vals.ctypes.data, vals.nbytes)

            self._fill_cb = FILL_EX_MASK(_fill)  # kept alive while the context is
            if not L.PAMR_excision_on(b"chr", C.cast(self._fill_cb, C.c_void_p), ex, 0):
                raise RuntimeError("PAMR_excision_on failed")
        elif not self.app.nps_define_vars_tl(op, rop_u, rop_res, rop_src, num_tl):
            raise RuntimeError("nps_define_vars failed")
        if extrap and not L.amrd_mg_def_extrap_vars(b"", b""):
            raise RuntimeError("amrd_mg_def_extrap_vars failed")
        if extra_defs is not None:
            extra_defs(L)  # more variables, before the hierarchy is composed
        if tre_vars and not L.amrd_def_tre_vars(tre_vars.encode()):
            raise RuntimeError("amrd_def_tre_vars failed")
        boxes = [list(self.bbox)] + [list(b) for lv in levels for b in lv]
        levs = [1] + [l + 2 for l, lv in enumerate(levels) for _ in lv]
        bb = np.array(boxes, dtype=np.float64).reshape(-1)
        lv = (C.c_int * len(levs))(*levs)
        self.max_lev = len(levels) + 1
        if not L.PAMR_compose_hierarchy(1, self.max_lev, len(levs), lv, P(bb), 0.0):
            raise RuntimeError("PAMR_compose_hierarchy failed")
        if not self.app.nps_init():
            raise RuntimeError("nps_init failed")

    def close(self):
        if getattr(self, "open", False):
            self.L.PAMR_destroy_mgh()
            self.L.PAMR_free_context(1)
            self.app.nps_release()
            self.open = False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------
    def gfn(self, name, hier, tl=1):
        return self.L.PAMR_get_gfn(name.encode(), hier, 0 if hier == PAMR_MGH else tl)

    def grids(self, hier, lev):
        """[(shape (z, y, x), bbox, gfs pointers)] of the grids of lev, in
        iterator order."""
        L = self.L
        out = []
        valid = L.PAMR_init_s_iter(lev, hier, 0)
        while valid:
            rank, dim, ngfs = C.c_int(), C.c_int(), C.c_int()
            shp = (C.c_int * 3)()
            gw = (C.c_int * 6)()
            bb = np.zeros(6)
            t = C.c_double()
            x = (C.POINTER(C.c_double) * 3)()
            gfs = (C.c_void_p * MAX_GFNS)()
            if not L.PAMR_get_g_attribs(C.byref(rank), C.byref(dim), shp, P(bb), gw, C.byref(t), C.byref(ngfs), x,
                                        gfs):
                raise RuntimeError("PAMR_get_g_attribs failed")
            out.append(((shp[2], shp[1], shp[0]), bb, [gfs[k] for k in range(ngfs.value)]))
            valid = L.PAMR_next_g()
        return out

    def upload(self, hier, lev, g, name, arr, tl=1):
        shape, _, gfs = self.grids(hier, lev)[g]
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.shape == shape
        ptr = gfs[self.gfn(name, hier, tl) - 1]
        assert ptr, (name, hier, lev, g)
        if not self.app.nps_upload(ptr, P(a), a.size):
            raise RuntimeError("upload failed")

    def download(self, hier, lev, g, name, tl=1):
        shape, _, gfs = self.grids(hier, lev)[g]
        ptr = gfs[self.gfn(name, hier, tl) - 1]
        assert ptr, (name, hier, lev, g)
        out = np.empty(shape)
        if not self.app.nps_download(P(out), ptr, out.size):
            raise RuntimeError("download failed")
        return out

    def set_params(self, pre=2, pst=2, min_iter=1, max_iter=50, tol=0.0, crtol=1e-3, max_citer=5000, w0_r=1.0,
                   w0_i=1.0):
        self.L.amrd_mg_set_params(pre, pst, min_iter, max_iter, tol, crtol, max_citer, w0_r, w0_i)

    def build_mgh(self, tl=1):
        nl = self.L.PAMR_build_mgh(1, self.max_lev, tl)
        if nl <= 0:
            raise RuntimeError("PAMR_build_mgh failed")
        return nl

    def vcycle(self, giter=0, save_tre=False, reference_name=False):
        """One V-cycle through amrd_vcycle_tre, or through the reference's
        own entry point vcycle(giter, save_tre) when reference_name."""
        fn = self.L.vcycle if reference_name else self.L.amrd_vcycle_tre
        r = fn(giter, 1 if save_tre else 0)
        if r < 0:
            raise RuntimeError("amrd_vcycle failed")
        nl = self.L.PAMR_get_max_lev(PAMR_MGH)
        norms = np.zeros(nl)
        cit = C.c_int()
        self.L.amrd_mg_last_norms(P(norms), C.byref(cit))
        return r, norms, cit.value
