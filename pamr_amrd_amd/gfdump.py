"""Reader of the built-in grid-function record files written by PAMR_save_gfn
and mg_dump when no application writer is registered (include/pamr.h).

A file "<output name>.pgf" holds one record per saved grid, in the host's
byte order: b"PGF1", int32 rank, float64 time, int32 shape[rank],
float64 bbox[2 rank], float64 data[prod(shape)] (x fastest).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Record:
    time: float
    shape: tuple          # (nx, ny, nz) as written
    bbox: np.ndarray      # (2 rank,)
    data: np.ndarray      # (nz, ny, nx): C order over the Fortran-ordered values


def read_pgf(path):
    """All records of a .pgf file, in the order they were written."""
    raw = open(path, "rb").read()
    out, pos = [], 0
    while pos < len(raw):
        if raw[pos:pos + 4] != b"PGF1":
            raise ValueError(f"{path}: bad record magic at byte {pos}")
        pos += 4
        rank = int(np.frombuffer(raw, np.int32, 1, pos)[0])
        pos += 4
        t = float(np.frombuffer(raw, np.float64, 1, pos)[0])
        pos += 8
        shape = tuple(int(v) for v in np.frombuffer(raw, np.int32, rank, pos))
        pos += 4 * rank
        bbox = np.frombuffer(raw, np.float64, 2 * rank, pos).copy()
        pos += 16 * rank
        n = int(np.prod(shape))
        data = np.frombuffer(raw, np.float64, n, pos).reshape(shape[::-1]).copy()
        pos += 8 * n
        out.append(Record(t, shape, bbox, data))
    return out
