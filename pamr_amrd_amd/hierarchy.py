"""Seeded AMR hierarchies and grid functions for the batched multi-patch solve.

Workload shape (batched AMR hierarchy): a base grid of ``base_n`` points per
side over the unit cube, refinement ratio 2, and ``nref`` refined levels.
Each refined level holds between ``nmin`` and ``nmax`` boxes.  A box picks a
parent box of the level above at random, starts on a grid line of that
parent, has odd side lengths (in points) uniform in ``[smin, smax]``, and is
clipped to the parent -- so every box lies inside the coarser level
(the Berger-Oliger nesting that build_mgh assumes, src/gh.c:1216-1219).

Grid functions are evaluated from closed forms at each grid's own points, so
overlapping patches and coarser levels see the same data:

* ``f``: the 8-Gaussian source of :func:`pamr_amrd_amd.gaussian_params`;
* ``a = exp(0.5 g)``: ``g`` a unit-variance Gaussian random field with
  squared-exponential correlation of length ``ell`` (random Fourier
  features, 64 modes);
* ``c``: uniform in [0, 1], a hash of the point's index on the 2^12 lattice.
"""
from __future__ import annotations

import numpy as np

from . import gaussian_params, splitmix64

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _u01(gen):
    return float(int(next(gen)) >> 11) * (1.0 / 9007199254740992.0)


def random_boxes(base_n=129, nref=2, nmin=24, nmax=48, smin=17, smax=97, seed=7, min_side=5):
    """Per refined level, a list of physical boxes (x0, x1, y0, y1, z0, z1)."""
    gen = splitmix64(seed)
    dx0 = 1.0 / (base_n - 1)
    parents = [[(0, base_n - 1)] * 3]  # index boxes on the parent level's lattice
    out = []
    for lev in range(nref):
        nb = nmin + int(_u01(gen) * (nmax - nmin + 1))
        boxes = []
        tries = 0
        while len(boxes) < nb:
            tries += 1
            if tries > 100000:
                raise RuntimeError("random_boxes: cannot place boxes (parents too small)")
            p = parents[int(_u01(gen) * len(parents))]
            box = []
            for d in range(3):
                side = smin + 2 * int(_u01(gen) * ((smax - smin) // 2 + 1))
                lo, hi = 2 * p[d][0], 2 * p[d][1]  # parent extent in this level's indices
                start = lo + 2 * int(_u01(gen) * ((hi - lo) // 2 + 1))
                end = min(start + side - 1, hi)
                box.append((start, end))
            if any(e - s + 1 < min_side for s, e in box):
                continue
            boxes.append(box)
        dx = dx0 / float(2 << lev)
        out.append([tuple(v for s, e in box for v in (s * dx, e * dx)) for box in boxes])
        parents = boxes
    return out


def grid_points(shape, bbox):
    """Coordinates (x, y, z) of a grid's points, numpy-broadcastable (z, y, x),
    by repeated addition from the lower corner (src/gh.c:905-906)."""
    nz, ny, nx = shape
    axes = []
    for n, lo, hi in ((nx, bbox[0], bbox[1]), (ny, bbox[2], bbox[3]), (nz, bbox[4], bbox[5])):
        dx = (hi - lo) / (n - 1)
        x = np.empty(n)
        v = lo
        for m in range(n):
            x[m] = v
            v += dx
        axes.append(x)
    return axes[0][None, None, :], axes[1][None, :, None], axes[2][:, None, None]


def gaussian_on(shape, bbox, seed=1234, params=None):
    c, w, A = params if params is not None else gaussian_params(seed)
    X, Y, Z = grid_points(shape, bbox)
    f = np.zeros(shape)
    for k in range(len(A)):
        r2 = (X - c[k, 0]) ** 2 + (Y - c[k, 1]) ** 2 + (Z - c[k, 2]) ** 2
        f += A[k] * np.exp(-r2 / (w[k] * w[k]))
    return f


def _rff_params(seed, ell, modes=64):
    gen = splitmix64(seed)
    u = np.array([_u01(gen) for _ in range(5 * modes)]).reshape(modes, 5)
    # Box-Muller normals for the wave vectors (variance 1/ell^2 per axis)
    r = np.sqrt(-2.0 * np.log(np.maximum(u[:, 0], 1e-300)))
    k = np.stack([r * np.cos(2 * np.pi * u[:, 1]), r * np.sin(2 * np.pi * u[:, 1]),
                  np.sqrt(-2.0 * np.log(np.maximum(u[:, 2], 1e-300))) * np.cos(2 * np.pi * u[:, 3])],
                 axis=1) / ell
    phase = 2 * np.pi * u[:, 4]
    return k, phase


def lognormal_on(shape, bbox, seed=99, sigma=0.5, ell=0.1, modes=64):
    """a = exp(sigma * g), g ~ unit-variance field with correlation length ell."""
    k, phase = _rff_params(seed, ell, modes)
    X, Y, Z = grid_points(shape, bbox)
    g = np.zeros(shape)
    for m in range(modes):
        g += np.cos(k[m, 0] * X + k[m, 1] * Y + k[m, 2] * Z + phase[m])
    g *= np.sqrt(2.0 / modes)
    return np.exp(sigma * g)


def lognormal_on_device(shape, bbox, torch, seed=99, sigma=0.5, ell=0.1, modes=64):
    """lognormal_on evaluated with torch on the GPU (large grids); returns a
    float64 device tensor of the same values up to the last bits."""
    k, phase = _rff_params(seed, ell, modes)
    X, Y, Z = (torch.from_numpy(np.ascontiguousarray(v)).cuda() for v in grid_points(shape, bbox))
    g = torch.zeros(shape, dtype=torch.float64, device="cuda")
    for m in range(modes):
        g += torch.cos(float(k[m, 0]) * X + float(k[m, 1]) * Y + float(k[m, 2]) * Z + float(phase[m]))
    g *= float(np.sqrt(2.0 / modes))
    return torch.exp(sigma * g)


def excision_spheres(seed=21, count=3, rmin=0.05, rmax=0.12):
    """Seeded spheres (centre in [0.25, 0.75]^3, radius uniform in [rmin, rmax])."""
    gen = splitmix64(seed)
    out = []
    for _ in range(count):
        c = np.array([0.25 + 0.5 * _u01(gen) for _ in range(3)])
        out.append((c, rmin + (rmax - rmin) * _u01(gen)))
    return out


def chr_on(shape, bbox, spheres, ex=1.0):
    """chr grid function: ex inside any sphere (excised), 0 elsewhere."""
    X, Y, Z = grid_points(shape, bbox)
    inside = np.zeros(shape, dtype=bool)
    for c, r in spheres:
        inside |= (X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2 < r * r
    return np.where(inside, ex, 0.0)


def excision_data(shape, bbox):
    """Dirichlet data held on the excised points: g = 0.1 (1 + x + y + z)."""
    X, Y, Z = grid_points(shape, bbox)
    return np.broadcast_to(0.1 * (1.0 + X + Y + Z), shape).copy()


def _mix(z):
    with np.errstate(over="ignore"):
        z = (z + np.uint64(0x9E3779B97F4A7C15)) & _M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
        return z ^ (z >> np.uint64(31))


def uniform_on(shape, bbox, seed=5, lattice=4096):
    """c in [0, 1): a hash of each point's index on a fixed fine lattice."""
    X, Y, Z = grid_points(shape, bbox)
    ix = np.rint(X * lattice).astype(np.uint64)
    iy = np.rint(Y * lattice).astype(np.uint64)
    iz = np.rint(Z * lattice).astype(np.uint64)
    key = (ix + np.uint64(lattice + 1) * (iy + np.uint64(lattice + 1) * iz)) ^ np.uint64(seed)
    key = np.broadcast_to(key, shape)
    return (_mix(key.astype(np.uint64)) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
