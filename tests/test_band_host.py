"""The band plan of the learned excision correction, host side (no GPU):
pmg_band_mask against scipy's exact Euclidean distance transform and a brute
force over all point pairs, on cubic and non-cubic grids; pmg_band_cores
against a brute-force list of the cores that hold band points."""
import ctypes as C

import numpy as np
import pytest

from pamr_amrd_amd import hierarchy as H

BOX = (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)


@pytest.fixture(scope="module")
def L():
    import pamr_amrd_amd as pm
    return pm.lib()


def _mask(L, ch, depth, ex=1.0):
    nz, ny, nx = ch.shape
    band = np.zeros(ch.shape, dtype=np.uint8)
    c = np.ascontiguousarray(ch, dtype=np.float64)
    cnt = L.pmg_band_mask(c.ctypes.data_as(C.POINTER(C.c_double)), nx, ny, nz, ex, depth,
                          band.ctypes.data_as(C.POINTER(C.c_ubyte)))
    return cnt, band


def _expected(ch, depth, ex=1.0):
    from scipy import ndimage
    exc = ch == ex
    dist = ndimage.distance_transform_edt(~exc)
    inner = np.zeros(ch.shape, dtype=bool)
    inner[1:-1, 1:-1, 1:-1] = True
    return inner & ~exc & (dist <= depth)


@pytest.mark.parametrize("shape,seed,depth", [((33, 33, 33), 21, 3), ((65, 65, 65), 7, 3),
                                              ((20, 17, 23), 5, 2), ((33, 33, 33), 9, 5)])
def test_band_mask_is_the_distance_band(L, shape, seed, depth):
    spheres = H.excision_spheres(seed=seed, rmin=0.08, rmax=0.2)
    ch = H.chr_on(shape, BOX, spheres)
    cnt, band = _mask(L, ch, depth)
    want = _expected(ch, depth)
    assert want.sum() > 0
    assert cnt == want.sum()
    assert np.array_equal(band.astype(bool), want)


def test_band_mask_brute_force_random_points(L):
    rng = np.random.default_rng(3)
    shape = (9, 8, 10)
    ch = np.where(rng.random(shape) < 0.06, 2.0, 0.0)  # excised value 2 (AMRD_ex is a parameter)
    depth = 2
    cnt, band = _mask(L, ch, depth, ex=2.0)
    exc = np.argwhere(ch == 2.0)
    want = np.zeros(shape, dtype=bool)
    for k in range(1, shape[0] - 1):
        for j in range(1, shape[1] - 1):
            for i in range(1, shape[2] - 1):
                if ch[k, j, i] == 2.0:
                    continue
                d2 = ((exc - np.array([k, j, i])) ** 2).sum(axis=1)
                want[k, j, i] = d2.min() <= depth * depth
    assert cnt == want.sum()
    assert np.array_equal(band.astype(bool), want)


def test_band_mask_rejects_bad_arguments(L):
    ch = np.zeros((5, 5, 5))
    cnt, _ = _mask(L, ch, 0)
    assert cnt == -1
    cnt, band = _mask(L, ch, 3)  # nothing excised: empty band
    assert cnt == 0 and band.sum() == 0


@pytest.mark.parametrize("shape,core", [((65, 65, 65), 26), ((20, 17, 23), 8)])
def test_band_cores_cover_the_band(L, shape, core):
    spheres = H.excision_spheres(seed=21, rmin=0.08, rmax=0.2)
    ch = H.chr_on(shape, BOX, spheres)
    _, band = _mask(L, ch, 3)
    nz, ny, nx = shape
    margin = 3
    cap = (-(-nx // core)) * (-(-ny // core)) * (-(-nz // core))
    org = np.zeros(3 * cap, dtype=np.int32)
    m = L.pmg_band_cores(band.ctypes.data_as(C.POINTER(C.c_ubyte)), nx, ny, nz, core, margin,
                         org.ctypes.data_as(C.POINTER(C.c_int)), cap)
    want = set()
    for k, j, i in np.argwhere(band):
        want.add((i // core, j // core, k // core))
    got = {((org[3 * c] + margin) // core, (org[3 * c + 1] + margin) // core, (org[3 * c + 2] + margin) // core)
           for c in range(m)}
    assert m == len(want) == len(got)
    assert got == want
    for c in range(m):  # origins are core corners minus the margin
        assert all((org[3 * c + d] + margin) % core == 0 for d in range(3))
    # too small a capacity is refused
    if m > 1:
        assert L.pmg_band_cores(band.ctypes.data_as(C.POINTER(C.c_ubyte)), nx, ny, nz, core, margin,
                                org.ctypes.data_as(C.POINTER(C.c_int)), m - 1) == -1
