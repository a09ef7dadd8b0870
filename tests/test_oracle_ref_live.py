"""The oracle against the reference Fortran compiled from source into
oracle/_ref (present in the build container only), on more sizes and
random masks than the committed golden vectors."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle as O

REF_TREE = "/root/reference"


def _ref(which):
    # loaded only where the reference source tree is (the build container):
    # the reference, compiled or not, is never loaded anywhere else
    if not os.path.isdir(REF_TREE):
        pytest.skip("reference source tree absent")
    lib = O.ref_lib(which)
    if lib is None:
        pytest.skip("oracle/_ref not built")
    return lib


@pytest.fixture(scope="module")
def R():
    return _ref("mutil")


@pytest.fixture(scope="module")
def NPS():
    return _ref("nps")


def ci(v):
    return C.byref(C.c_int(v))


def cd(v):
    return C.byref(C.c_double(v))


@pytest.mark.parametrize("shape", [(5, 5, 5), (9, 9, 9), (17, 9, 5), (17, 13, 11), (33, 17, 9)])
@pytest.mark.parametrize("op", [2, 3])
@pytest.mark.parametrize("ex", [0, 1])
def test_restrict_live(R, shape, op, ex):
    nx, ny, nz = shape
    rng = np.random.default_rng(nx * 100 + ny + op + 7 * ex)
    u = rng.standard_normal((nz, ny, nx))
    chr_ = np.where(rng.random((nz, ny, nx)) < 0.25, 1.0, 0.0) if ex else u
    if ex and op == 3 and nx > 17:
        # keep the last x face zero: there the reference reads past w3d
        # (src/m_util_r8.f:931); with zero data every weight gives 0
        u[:, :, -1] = 0.0
    cx, cy, cz = (nx - 1) // 2 + 1, (ny - 1) // 2 + 1, (nz - 1) // 2 + 1
    a = np.zeros((cz, cy, cx))
    b = np.zeros((cz, cy, cx))
    args = [nx, ny, nz, cx, cy, cz, 1, 1, 1, 1, 1, 1, cx, cy, cz, 2, 2, 2, 1, 1, 1]
    (O.lib().or_dmfwr3d if op == 3 else O.lib().or_dmhwr3d)(O.P(u), O.P(a), *args, O.P(chr_), 1.0, ex)
    (R.dmfwr3d_ if op == 3 else R.dmhwr3d_)(O.P(u), O.P(b), *[ci(v) for v in args], O.P(chr_), cd(1.0), ci(ex))
    assert np.array_equal(a, b)


@pytest.mark.parametrize("cshape", [(3, 3, 3), (5, 5, 5), (9, 5, 3), (17, 9, 9)])
@pytest.mark.parametrize("ex", [0, 1])
def test_interp_live(R, cshape, ex):
    cx, cy, cz = cshape
    rng = np.random.default_rng(cx * 31 + cy + ex)
    u = rng.standard_normal((cz, cy, cx))
    chr_ = np.where(rng.random((cz, cy, cx)) < 0.2, 1.0, 0.0) if ex else u
    nx, ny, nz = 2 * cx - 1, 2 * cy - 1, 2 * cz - 1
    a = np.zeros((nz, ny, nx))
    b = np.zeros((nz, ny, nx))
    ua, ub = u.copy(), u.copy()
    args = [cx, cy, cz, nx, ny, nz, 1, 1, 1, 1, 1, 1, cx, cy, cz, 1, 1, 1, 2, 2, 2, 2]
    O.lib().or_dminterp3d(O.P(ua), O.P(a), *args, O.P(chr_ if ex else ua), 1.0, ex)
    R.dminterp3d_(O.P(ub), O.P(b), *[ci(v) for v in args], O.P(chr_ if ex else ub), cd(1.0), ci(ex))
    assert np.array_equal(a, b)
    assert np.array_equal(ua, ub)


@pytest.mark.parametrize("seg", [
    # (ni, nj, nk), (i2, j2, k2) offsets in u2, rho, extra fine points past the segment
    ((5, 5, 5), (1, 1, 1), 2, 0),       # whole grid: both boundary stencils everywhere
    ((6, 4, 5), (3, 3, 5), 2, 3),       # interior segment: anchors outside it are read
    ((4, 7, 3), (2, 1, 3), 2, 1),       # nk = 3 < 4: that axis falls back to dvi2q1d
    ((5, 4, 4), (5, 9, 1), 4, 8),       # rho 4
    ((9, 5, 6), (3, 1, 3), 2, 2)])
def test_interp4_live(R, seg):
    """ord = 4 (dvi4q1d), no excision, with u2 pre-filled: the interior
    cubic stencils read u2's anchors outside the interpolated segment"""
    (ni, nj, nk), (i2, j2, k2), rho, extra = seg
    rng = np.random.default_rng(ni * 97 + nj * 13 + nk + rho)
    cx, cy, cz = ni + 2, nj + 1, nk + 3
    u1 = rng.standard_normal((cz, cy, cx))
    nx = i2 + (ni - 1) * rho + extra
    ny = j2 + (nj - 1) * rho + extra
    nz = k2 + (nk - 1) * rho + extra
    u2 = rng.standard_normal((nz, ny, nx))
    a, b = u2.copy(), u2.copy()
    args = [cx, cy, cz, nx, ny, nz, 2, 1, 3, i2, j2, k2, ni, nj, nk, 1, 1, 1, rho, rho, rho, 4]
    O.lib().or_dminterp3d(O.P(u1), O.P(a), *args, O.P(u1), 1.0, 0)
    R.dminterp3d_(O.P(u1), O.P(b), *[ci(v) for v in args], O.P(u1), cd(1.0), ci(0))
    assert np.array_equal(a, b)
    assert not np.array_equal(a, u2)


@pytest.mark.parametrize("case", [
    # coarse shape, coarse offset (i1,j1,k1), segment (ni,nj,nk), fine offset, rho, mask density
    ((5, 5, 5), (1, 1, 1), (5, 5, 5), (1, 1, 1), 2, 0.15),
    ((9, 7, 6), (1, 1, 1), (9, 7, 6), (1, 1, 1), 2, 0.3),
    ((9, 9, 9), (1, 1, 1), (9, 9, 9), (1, 1, 1), 2, 0.05),
    ((14, 13, 15), (5, 5, 5), (6, 5, 7), (9, 9, 9), 2, 0.2),
    ((13, 13, 13), (5, 5, 5), (5, 5, 5), (17, 17, 17), 4, 0.25)])
def test_interp4_ex_live(R, case):
    """ord = 4 with excision (dvi4q1d_ex): extrapolation into excised
    anchors, run starts / ends, the dvi2q1d_ex fallback for short runs"""
    (cx, cy, cz), (i1, j1, k1), (ni, nj, nk), (i2, j2, k2), rho, dens = case
    rng = np.random.default_rng(cx * 7 + cy * 3 + cz + rho + int(dens * 100))
    u1 = rng.standard_normal((cz, cy, cx))
    chr_ = np.where(rng.random((cz, cy, cx)) < dens, 1.0, 0.0)
    if (i1, j1, k1) == (1, 1, 1):
        nx, ny, nz = (ni - 1) * rho + 1, (nj - 1) * rho + 1, (nk - 1) * rho + 1
    else:
        nx, ny, nz = [o + (m + 3) * rho + 1 for o, m in ((i2, ni), (j2, nj), (k2, nk))]
    u2 = rng.standard_normal((nz, ny, nx))
    a, b = u2.copy(), u2.copy()
    ua, ub = u1.copy(), u1.copy()
    args = [cx, cy, cz, nx, ny, nz, i1, j1, k1, i2, j2, k2, ni, nj, nk, 1, 1, 1, rho, rho, rho, 4]
    O.lib().or_dminterp3d(O.P(ua), O.P(a), *args, O.P(chr_), 1.0, 1)
    R.dminterp3d_(O.P(ub), O.P(b), *[ci(v) for v in args], O.P(chr_), cd(1.0), ci(1))
    assert np.array_equal(a, b)
    assert np.array_equal(ua, ub)


@pytest.mark.parametrize("shape", [(5, 5, 5), (12, 9, 7), (17, 17, 17)])
def test_hooks_live(NPS, shape):
    nx, ny, nz = shape
    rng = np.random.default_rng(nx + 1000 * nz)
    V = rng.standard_normal((nz, ny, nx))
    rho = rng.standard_normal((nz, ny, nx))
    rhs = rng.standard_normal((nz, ny, nx))
    cm = np.where(rng.random((nz, ny, nx)) < 0.7, 1.0, 0.0)
    x = np.cumsum(np.full(nx, 0.1)) - 0.1
    y = np.cumsum(np.full(ny, 0.07)) - 0.07
    z = np.cumsum(np.full(nz, 0.2)) - 0.2
    dx, dy, dz = x[1] - x[0], y[1] - y[0], z[1] - z[0]
    L = O.lib()
    for _ in range(2):
        Va, Vb = V.copy(), V.copy()
        na = L.or_relax(0, O.P(Va), O.P(rhs), O.P(rho), None, None, O.P(cm), None, 1.0, nx, ny, nz, dx, dy, dz)
        nb = C.c_double(0)
        pb = (C.c_int * 6)()
        NPS.relax_(O.P(Vb), O.P(rhs), O.P(rho), O.P(cm), pb, O.P(x), O.P(y), O.P(z), C.byref(nb), ci(nx - 1),
                   ci(ny - 1), ci(nz - 1), pb)
        assert np.array_equal(Va, Vb)
        assert na == nb.value
        V = Va
