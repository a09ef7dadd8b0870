"""The built-in grid-function record format (include/pamr.h, PAMR_save_gfn
without an application writer) read back by pamr_amrd_amd.gfdump: records
assembled here byte by byte from the format's definition."""
import struct

import numpy as np
import pytest

from pamr_amrd_amd.gfdump import read_pgf


def _record(t, shape, bbox, data):
    # data given as (nz, ny, nx); the file holds it x fastest
    out = b"PGF1" + struct.pack("<i", len(shape)) + struct.pack("<d", t)
    out += struct.pack(f"<{len(shape)}i", *shape) + struct.pack(f"<{2 * len(shape)}d", *bbox)
    return out + np.ascontiguousarray(data, dtype="<f8").tobytes()


def test_records_round_trip(tmp_path):
    rng = np.random.default_rng(5)
    a = rng.standard_normal((4, 3, 5))          # nz, ny, nx
    b = rng.standard_normal((2, 2, 2))
    p = tmp_path / "V_MG_0.pgf"
    p.write_bytes(_record(0.25, (5, 3, 4), (0, 1, 0, .5, 0, .75), a) + _record(-3.0, (2, 2, 2), (1, 2, 3, 4, 5, 6), b))
    r = read_pgf(p)
    assert len(r) == 2
    assert r[0].time == 0.25 and r[0].shape == (5, 3, 4)
    assert np.array_equal(r[0].bbox, [0, 1, 0, .5, 0, .75])
    assert np.array_equal(r[0].data, a)
    assert r[0].data[3, 2, 4] == a[3, 2, 4]      # (k, j, i) of the x-fastest layout
    assert r[1].time == -3.0 and np.array_equal(r[1].data, b)


def test_bad_magic_is_refused(tmp_path):
    p = tmp_path / "x.pgf"
    p.write_bytes(b"SDF?" + bytes(40))
    with pytest.raises(ValueError):
        read_pgf(p)
