"""Write golden vectors of the reference's own Fortran kernels and hooks.

Runs in the build container only: it loads oracle/_ref/*.so, which
oracle/Makefile compiles from the reference's sources (src/m_util_r8.f,
w/examples/nps/num_w.f, examples/nbs/num.f).  Inputs are seeded random
arrays; outputs are what the reference code returns for them.  The result,
tests/golden/ref_kernels.npz, is plain data (np.savez, no pickles).

Shapes are kept small, and the excised full-weight restriction uses fine
grids with nx <= 17 so that the reference's w3d(i,1,.) index quirk
(src/m_util_r8.f:931) stays inside its 27-entry table (defined behaviour).
"""
import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_kernels.npz")


def ci(v):
    return C.byref(C.c_int(v))


def cd(v):
    return C.byref(C.c_double(v))


def main():
    R = oracle.ref_lib("mutil")
    NPS = oracle.ref_lib("nps")
    NBS = oracle.ref_lib("nbs")
    if R is None or NPS is None or NBS is None:
        raise SystemExit("oracle/_ref not built (make -C oracle ref)")
    rng = np.random.default_rng(20261202)
    out = {}
    P = oracle.P
    # restrictions
    for tag, (nx, ny, nz) in {"a": (9, 7, 5), "b": (17, 9, 9)}.items():
        u = rng.standard_normal((nz, ny, nx))
        chr_ = np.where(rng.random((nz, ny, nx)) < 0.2, 1.0, 0.0)
        cx, cy, cz = (nx - 1) // 2 + 1, (ny - 1) // 2 + 1, (nz - 1) // 2 + 1
        args = [nx, ny, nz, cx, cy, cz, 1, 1, 1, 1, 1, 1, cx, cy, cz, 2, 2, 2, 1, 1, 1]
        out[f"restr_{tag}_u"] = u
        out[f"restr_{tag}_chr"] = chr_
        for name, f in (("fw", R.dmfwr3d_), ("hw", R.dmhwr3d_)):
            for ex in (0, 1):
                o = np.zeros((cz, cy, cx))
                f(P(u), P(o), *[ci(v) for v in args], P(chr_ if ex else u), cd(1.0), ci(ex))
                out[f"restr_{tag}_{name}{ex}"] = o
        o = np.zeros((cz, cy, cx))
        R.dmcopy3d_(P(u), P(o), *[ci(v) for v in args])
        out[f"restr_{tag}_copy"] = o
    # interpolation
    for tag, (cx, cy, cz) in {"a": (5, 4, 3), "b": (9, 5, 5)}.items():
        u = rng.standard_normal((cz, cy, cx))
        chr_ = np.where(rng.random((cz, cy, cx)) < 0.15, 1.0, 0.0)
        nx, ny, nz = 2 * cx - 1, 2 * cy - 1, 2 * cz - 1
        args = [cx, cy, cz, nx, ny, nz, 1, 1, 1, 1, 1, 1, cx, cy, cz, 1, 1, 1, 2, 2, 2, 2]
        out[f"interp_{tag}_u"] = u
        out[f"interp_{tag}_chr"] = chr_
        for ex in (0, 1):
            o = np.zeros((nz, ny, nx))
            uu = u.copy()
            R.dminterp3d_(P(uu), P(o), *[ci(v) for v in args], P(chr_ if ex else uu), cd(1.0), ci(ex))
            out[f"interp_{tag}_ex{ex}"] = o
    # hooks (nps: vertex point counts nx = Nx+1)
    nx, ny, nz = 11, 9, 8
    V = rng.standard_normal((nz, ny, nx))
    rho = rng.standard_normal((nz, ny, nx))
    rhs = rng.standard_normal((nz, ny, nx))
    cm = np.where(rng.random((nz, ny, nx)) < 0.8, 1.0, 0.0)
    x = 0.3 + np.arange(nx) * 0.125
    y = -0.2 + np.arange(ny) * 0.1
    z = np.arange(nz) * 0.0625
    for k, v in dict(V=V, rho=rho, rhs=rhs, cmask=cm, x=x, y=y, z=z).items():
        out[f"hook_{k}"] = v
    Vr = V.copy()
    nrm = C.c_double(0)
    pb = (C.c_int * 6)()
    NPS.relax_(P(Vr), P(rhs), P(rho), P(cm), pb, P(x), P(y), P(z), C.byref(nrm), ci(nx - 1), ci(ny - 1),
               ci(nz - 1), pb)
    out["hook_relax_V"] = Vr
    out["hook_relax_norm"] = np.array([nrm.value])
    res = np.zeros_like(V)
    NPS.residual_(P(res), P(rhs), P(V), P(rho), P(cm), P(x), P(y), P(z), C.byref(nrm), ci(nx - 1), ci(ny - 1),
                  ci(nz - 1))
    out["hook_residual_res"] = res
    out["hook_residual_norm"] = np.array([nrm.value])
    lv = np.zeros_like(V)
    NPS.lop_(P(lv), P(V), P(rho), P(cm), P(x), P(y), P(z), ci(nx - 1), ci(ny - 1), ci(nz - 1))
    out["hook_lop"] = lv
    # nbs relax with phi_i = 0 (source phi_r^2)
    phr = rng.standard_normal((nz, ny, nx))
    phi = np.zeros_like(phr)
    Vb = V.copy()
    NBS.relax_(P(Vb), P(rhs), P(phr), P(phi), P(cm), pb, P(x), P(y), P(z), C.byref(nrm), ci(nx), ci(ny), ci(nz))
    out["hook_nbs_phr"] = phr
    out["hook_nbs_relax_V"] = Vb
    out["hook_nbs_relax_norm"] = np.array([nrm.value])
    # fourth-order interpolation (dvi4q1d, no excision) into a pre-filled
    # target: an interior segment (anchors outside it are read) and the
    # whole grid (both boundary stencils)
    rng4 = np.random.default_rng(20261203)
    for tag, (ni, nj, nk, i2, j2, k2, rho, extra) in {"a": (6, 4, 5, 3, 3, 5, 2, 3),
                                                       "b": (5, 5, 5, 1, 1, 1, 2, 0)}.items():
        cx, cy, cz = ni + 2, nj + 1, nk + 3
        u1 = rng4.standard_normal((cz, cy, cx))
        nx = i2 + (ni - 1) * rho + extra
        ny = j2 + (nj - 1) * rho + extra
        nz = k2 + (nk - 1) * rho + extra
        u2 = rng4.standard_normal((nz, ny, nx))
        args = [cx, cy, cz, nx, ny, nz, 2, 1, 3, i2, j2, k2, ni, nj, nk, 1, 1, 1, rho, rho, rho, 4]
        o = u2.copy()
        R.dminterp3d_(P(u1), P(o), *[ci(v) for v in args], P(u1), cd(1.0), ci(0))
        out[f"interp4_{tag}_u1"] = u1
        out[f"interp4_{tag}_u2"] = u2
        out[f"interp4_{tag}_args"] = np.array(args, dtype=np.int64)
        out[f"interp4_{tag}_out"] = o
    # ... and with excision (dvi4q1d_ex): whole grid, random coarse masks
    for tag, (cx, cy, cz, dens) in {"a": (9, 7, 6, 0.3), "b": (9, 9, 9, 0.1)}.items():
        u1 = rng4.standard_normal((cz, cy, cx))
        chr_ = np.where(rng4.random((cz, cy, cx)) < dens, 1.0, 0.0)
        nx, ny, nz = 2 * cx - 1, 2 * cy - 1, 2 * cz - 1
        u2 = rng4.standard_normal((nz, ny, nx))
        args = [cx, cy, cz, nx, ny, nz, 1, 1, 1, 1, 1, 1, cx, cy, cz, 1, 1, 1, 2, 2, 2, 4]
        o, uu = u2.copy(), u1.copy()
        R.dminterp3d_(P(uu), P(o), *[ci(v) for v in args], P(chr_), cd(1.0), ci(1))
        out[f"interp4ex_{tag}_u1"] = u1
        out[f"interp4ex_{tag}_chr"] = chr_
        out[f"interp4ex_{tag}_u2"] = u2
        out[f"interp4ex_{tag}_args"] = np.array(args, dtype=np.int64)
        out[f"interp4ex_{tag}_out"] = o
        out[f"interp4ex_{tag}_u1_after"] = uu
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
