"""GPU parity of the device m_util_r8 routines and MG hooks against the
oracle (bit-exact fields; norms within 1e-13 relative: only the order of the
norm's sum differs)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


@pytest.fixture(scope="module")
def pm(torch_cuda):
    import pamr_amrd_amd as pm
    pm.lib()
    return pm


@pytest.mark.parametrize("shape", [(5, 5, 5), (9, 7, 5), (17, 9, 9), (33, 17, 13), (65, 65, 65)])
@pytest.mark.parametrize("kind", ["fw", "hw", "copy"])
@pytest.mark.parametrize("ex", [0, 1])
def test_restriction(pm, torch_cuda, shape, kind, ex):
    torch = torch_cuda
    if kind == "copy" and ex:
        pytest.skip("dmcopy3d has no excision variant")
    print(f"case restriction {kind} ex={ex} shape={shape}", flush=True)
    nx, ny, nz = shape
    rng = np.random.default_rng(nx * 7 + ny + nz + ex)
    u = rng.standard_normal((nz, ny, nx))
    chr_ = np.where(rng.random((nz, ny, nx)) < 0.2, 1.0, 0.0) if ex else u
    cx, cy, cz = (nx - 1) // 2 + 1, (ny - 1) // 2 + 1, (nz - 1) // 2 + 1
    args = [nx, ny, nz, cx, cy, cz, 1, 1, 1, 1, 1, 1, cx, cy, cz, 2, 2, 2, 1, 1, 1]
    ref = np.zeros((cz, cy, cx))
    du, dchr = _dev(torch, u), _dev(torch, chr_)
    dout = torch.zeros((cz, cy, cx), dtype=torch.float64, device="cuda")
    L = pm.lib()
    OL = O.lib()
    if kind == "fw":
        OL.or_dmfwr3d(O.P(u), O.P(ref), *args, O.P(chr_), 1.0, ex)
        rc = L.pmg_dmfwr3d(pm.tptr(du), pm.tptr(dout), *args, pm.tptr(dchr), 1.0, ex, None)
    elif kind == "hw":
        OL.or_dmhwr3d(O.P(u), O.P(ref), *args, O.P(chr_), 1.0, ex)
        rc = L.pmg_dmhwr3d(pm.tptr(du), pm.tptr(dout), *args, pm.tptr(dchr), 1.0, ex, None)
    else:
        OL.or_dmcopy3d(O.P(u), O.P(ref), *args)
        rc = L.pmg_dmcopy3d(pm.tptr(du), pm.tptr(dout), *args, None)
    assert rc == 0
    torch.cuda.synchronize()
    got = dout.cpu().numpy()
    assert np.array_equal(got, ref), np.abs(got - ref).max()


@pytest.mark.parametrize("cshape", [(3, 3, 3), (5, 4, 3), (9, 9, 9), (17, 9, 5), (33, 33, 33)])
@pytest.mark.parametrize("ex", [0, 1])
def test_interpolation(pm, torch_cuda, cshape, ex):
    torch = torch_cuda
    print(f"case interp ex={ex} cshape={cshape}", flush=True)
    cx, cy, cz = cshape
    rng = np.random.default_rng(cx * 13 + cy + cz + ex)
    u = rng.standard_normal((cz, cy, cx))
    chr_ = np.where(rng.random((cz, cy, cx)) < 0.2, 1.0, 0.0) if ex else u
    nx, ny, nz = 2 * cx - 1, 2 * cy - 1, 2 * cz - 1
    args = [cx, cy, cz, nx, ny, nz, 1, 1, 1, 1, 1, 1, cx, cy, cz, 1, 1, 1, 2, 2, 2, 2]
    ref = np.zeros((nz, ny, nx))
    O.lib().or_dminterp3d(O.P(u.copy()), O.P(ref), *args, O.P(chr_), 1.0, ex)
    du, dchr = _dev(torch, u), _dev(torch, chr_)
    dout = torch.zeros((nz, ny, nx), dtype=torch.float64, device="cuda")
    rc = pm.lib().pmg_dminterp3d(pm.tptr(du), pm.tptr(dout), *args, pm.tptr(dchr), 1.0, ex, None)
    assert rc == 0
    torch.cuda.synchronize()
    assert np.array_equal(dout.cpu().numpy(), ref)


@pytest.mark.parametrize("seg", [
    # (ni, nj, nk), (i2, j2, k2), rho, extra fine points past the segment
    ((5, 5, 5), (1, 1, 1), 2, 0),
    ((6, 4, 5), (3, 3, 5), 2, 3),
    ((4, 7, 3), (2, 1, 3), 2, 1),
    ((5, 4, 4), (5, 9, 1), 4, 8),
    ((33, 33, 33), (1, 1, 1), 2, 0),
    ((30, 21, 17), (4, 3, 5), 2, 4)])
def test_cubic_interpolation(pm, torch_cuda, seg):
    """dminterp3d ord = 4 (dvi4q1d, no excision) into a pre-filled target,
    bit for bit against the oracle (pinned to the reference Fortran)"""
    torch = torch_cuda
    (ni, nj, nk), (i2, j2, k2), rho, extra = seg
    print(f"case interp4 seg={seg}", flush=True)
    rng = np.random.default_rng(ni * 97 + nj * 13 + nk + rho)
    cx, cy, cz = ni + 2, nj + 1, nk + 3
    u1 = rng.standard_normal((cz, cy, cx))
    nx, ny, nz = i2 + (ni - 1) * rho + extra, j2 + (nj - 1) * rho + extra, k2 + (nk - 1) * rho + extra
    u2 = rng.standard_normal((nz, ny, nx))
    args = [cx, cy, cz, nx, ny, nz, 2, 1, 3, i2, j2, k2, ni, nj, nk, 1, 1, 1, rho, rho, rho, 4]
    ref = O.dminterp3d(u1.copy(), u2.copy(), args)
    du, dout = _dev(torch, u1), _dev(torch, u2)
    rc = pm.lib().pmg_dminterp3d(pm.tptr(du), pm.tptr(dout), *args, None, 1.0, 0, None)
    assert rc == 0
    torch.cuda.synchronize()
    assert np.array_equal(dout.cpu().numpy(), ref)


@pytest.mark.parametrize("case", [
    # coarse shape, coarse offset, segment, fine offset, rho, mask density
    ((5, 5, 5), (1, 1, 1), (5, 5, 5), (1, 1, 1), 2, 0.15),
    ((9, 7, 6), (1, 1, 1), (9, 7, 6), (1, 1, 1), 2, 0.3),
    ((33, 33, 33), (1, 1, 1), (33, 33, 33), (1, 1, 1), 2, 0.1),
    ((14, 13, 15), (5, 5, 5), (6, 5, 7), (9, 9, 9), 2, 0.2),
    ((13, 13, 13), (5, 5, 5), (5, 5, 5), (17, 17, 17), 4, 0.25)])
def test_cubic_interpolation_excised(pm, torch_cuda, case):
    """dminterp3d ord = 4 with excision (dvi4q1d_ex) bit for bit against
    the oracle; segments keep 4 coarse points of margin, the reach of the
    reference's anchor extension"""
    torch = torch_cuda
    (cx, cy, cz), (i1, j1, k1), (ni, nj, nk), (i2, j2, k2), rho, dens = case
    print(f"case interp4 ex {case}", flush=True)
    rng = np.random.default_rng(cx * 7 + cy * 3 + cz + rho + int(dens * 100))
    u1 = rng.standard_normal((cz, cy, cx))
    chr_ = np.where(rng.random((cz, cy, cx)) < dens, 1.0, 0.0)
    if (i1, j1, k1) == (1, 1, 1):
        nx, ny, nz = (ni - 1) * rho + 1, (nj - 1) * rho + 1, (nk - 1) * rho + 1
    else:
        nx, ny, nz = [o + (m + 3) * rho + 1 for o, m in ((i2, ni), (j2, nj), (k2, nk))]
    u2 = rng.standard_normal((nz, ny, nx))
    args = [cx, cy, cz, nx, ny, nz, i1, j1, k1, i2, j2, k2, ni, nj, nk, 1, 1, 1, rho, rho, rho, 4]
    ref = O.dminterp3d(u1.copy(), u2.copy(), args, chr1=chr_, ex=1.0, do_ex=1)
    du, dout, dchr = _dev(torch, u1), _dev(torch, u2), _dev(torch, chr_)
    rc = pm.lib().pmg_dminterp3d(pm.tptr(du), pm.tptr(dout), *args, pm.tptr(dchr), 1.0, 1, None)
    assert rc == 0
    torch.cuda.synchronize()
    assert np.array_equal(dout.cpu().numpy(), ref)


@pytest.mark.parametrize("shape", [(5, 5, 5), (12, 9, 7), (33, 33, 33), (65, 40, 17)])
@pytest.mark.parametrize("op", [0, 1])
@pytest.mark.parametrize("ex", [0, 1])
def test_hooks(pm, torch_cuda, shape, op, ex):
    torch = torch_cuda
    print(f"case hooks op={op} ex={ex} shape={shape}", flush=True)
    nx, ny, nz = shape
    rng = np.random.default_rng(nx + 10 * ny + 100 * nz + op + 2 * ex)
    V = rng.standard_normal((nz, ny, nx))
    rho = rng.standard_normal((nz, ny, nx))
    rhs = rng.standard_normal((nz, ny, nx))
    a = np.exp(0.5 * rng.standard_normal((nz, ny, nx))) if op else None
    c = rng.random((nz, ny, nx)) if op else None
    cm = np.where(rng.random((nz, ny, nx)) < 0.85, 1.0, 0.0)
    chr_ = np.where(rng.random((nz, ny, nx)) < 0.1, 2.0, 0.0) if ex else None
    dx, dy, dz = 1.0 / (nx - 1), 0.7 / (ny - 1), 1.3 / (nz - 1)
    OL = O.lib()
    d = {k: (_dev(torch, v) if v is not None else None) for k, v in
         dict(V=V, rho=rho, rhs=rhs, a=a, c=c, cm=cm, chr=chr_).items()}
    nrm = torch.zeros(1, dtype=torch.float64, device="cuda")
    L = pm.lib()
    P = pm.tptr
    # relax twice in a row
    Vr = V.copy()
    for _ in range(2):
        n_ref = OL.or_relax(op, O.P(Vr), O.P(rhs), O.P(rho), O.P(a) if op else None,
                            O.P(c) if op else None, O.P(cm), O.P(chr_) if ex else None, 2.0,
                            nx, ny, nz, dx, dy, dz)
        assert L.pmg_hook_relax(op, P(d["V"]), P(d["rhs"]), P(d["rho"]), P(d["a"]), P(d["c"]),
                                P(d["cm"]), P(d["chr"]), 2.0, nx, ny, nz, dx, dy, dz, P(nrm), None) == 0
        torch.cuda.synchronize()
        assert np.array_equal(d["V"].cpu().numpy(), Vr)
        assert abs(nrm.item() - n_ref) <= 1e-13 * abs(n_ref)
    res_ref = np.zeros_like(V)
    n_ref = OL.or_residual(op, O.P(res_ref), O.P(rhs), O.P(Vr), O.P(rho), O.P(a) if op else None,
                           O.P(c) if op else None, O.P(cm), O.P(chr_) if ex else None, 2.0, nx, ny, nz,
                           dx, dy, dz)
    dres = torch.zeros_like(d["V"])
    assert L.pmg_hook_residual(op, P(dres), P(d["rhs"]), P(d["V"]), P(d["rho"]), P(d["a"]), P(d["c"]),
                               P(d["cm"]), P(d["chr"]), 2.0, nx, ny, nz, dx, dy, dz, P(nrm), None) == 0
    torch.cuda.synchronize()
    assert np.array_equal(dres.cpu().numpy(), res_ref)
    assert abs(nrm.item() - n_ref) <= 1e-13 * abs(n_ref)
    lop_ref = np.zeros_like(V)
    OL.or_lop(op, O.P(lop_ref), O.P(Vr), O.P(rho), O.P(a) if op else None, O.P(c) if op else None,
              O.P(cm), O.P(chr_) if ex else None, 2.0, nx, ny, nz, dx, dy, dz)
    dlop = torch.zeros_like(d["V"])
    assert L.pmg_hook_lop(op, P(dlop), P(d["V"]), P(d["rho"]), P(d["a"]), P(d["c"]), P(d["cm"]),
                          P(d["chr"]), 2.0, nx, ny, nz, dx, dy, dz, None) == 0
    torch.cuda.synchronize()
    assert np.array_equal(dlop.cpu().numpy(), lop_ref)
