"""Pins of the two oracle functions no reference routine covers: the
variable-coefficient operator (op 1, L[u] = div(a grad u) + c u - rho with
arithmetic-mean face coefficients) and the network oracle (or_net_forward /
or_net_loss_grad).

op 1 is fixed by properties a wrong sign, index or dropped term breaks:
reduction to the Poisson hook (pinned bit for bit against the reference
Fortran) for a = 1, c = 0; constants map to c*const; linear u with linear a
is differentiated exactly; the operator is symmetric; its diagonal is the
Jacobian the relaxation divides by (after a sweep the last colour's
residuals vanish); and the discrete solution of a manufactured problem
converges at second order.  The network oracle is fixed against torch's
conv3d / GELU (an independent implementation) and by central finite
differences of the loss."""
import numpy as np
import pytest

import oracle as O

EX = 1.0


def _lop(op, V, rho, a, c, cm, h):
    nz, ny, nx = V.shape
    out = np.zeros_like(V)
    O.lib().or_lop(op, O.P(out), O.P(V), O.P(rho), O.P(a), O.P(c), O.P(cm), None, EX, nx, ny, nz, *h)
    return out


def _rand(shape, seed):
    rng = np.random.default_rng(seed)
    V = rng.standard_normal(shape)
    a = np.exp(0.5 * rng.standard_normal(shape))
    c = rng.random(shape)
    rho = rng.standard_normal(shape)
    return V, a, c, rho


SHAPES = [(7, 9, 11), (13, 6, 5)]
H = (0.125, 0.1, 0.0625)


@pytest.mark.parametrize("shape", SHAPES)
def test_op1_reduces_to_poisson(shape):
    """a = 1, c = 0: op 1 equals op 0 (the reference's nps hook) to rounding,
    for lop, residual and relax, and the Jacobians agree exactly."""
    V, _, _, rho = _rand(shape, 1)
    one, zero, cm = np.ones(shape), np.zeros(shape), np.ones(shape)
    l0 = _lop(0, V, rho, one, zero, cm, H)
    l1 = _lop(1, V, rho, one, zero, cm, H)
    scale = np.abs(V).max() / min(H) ** 2
    assert np.abs(l0 - l1).max() <= 1e-14 * scale
    nz, ny, nx = shape
    rhs = np.random.default_rng(2).standard_normal(shape)
    V0, V1 = V.copy(), V.copy()
    L = O.lib()
    for _ in range(3):
        n0 = L.or_relax(0, O.P(V0), O.P(rhs), O.P(rho), O.P(one), O.P(zero), O.P(cm), None, EX, nx, ny, nz, *H)
        n1 = L.or_relax(1, O.P(V1), O.P(rhs), O.P(rho), O.P(one), O.P(zero), O.P(cm), None, EX, nx, ny, nz, *H)
        assert abs(n0 - n1) <= 1e-12 * n0
    assert np.abs(V0 - V1).max() <= 1e-12 * np.abs(V0).max()


@pytest.mark.parametrize("shape", SHAPES)
def test_op1_constants_and_linear(shape):
    """L[const] = c * const for any a; for u = x + 2y - z and a affine in
    (x, y, z), div(a grad u) = grad a . grad u exactly."""
    nz, ny, nx = shape
    _, a, c, _ = _rand(shape, 3)
    zero, cm = np.zeros(shape), np.ones(shape)
    l = _lop(1, np.full(shape, 3.0), zero, a, c, cm, H)
    inner = (slice(1, -1),) * 3
    assert np.allclose(l[inner], 3.0 * c[inner], rtol=0, atol=1e-12)
    z, y, x = np.meshgrid(np.arange(nz) * H[2], np.arange(ny) * H[1], np.arange(nx) * H[0], indexing="ij")
    u = x + 2 * y - z
    aa = 2.0 + 0.5 * x - 0.25 * y + 0.75 * z
    l = _lop(1, u, zero, aa, zero, cm, H)
    exact = 0.5 * 1 + (-0.25) * 2 + 0.75 * (-1)
    assert np.allclose(l[inner], exact, rtol=0, atol=1e-11)


@pytest.mark.parametrize("shape", SHAPES)
def test_op1_symmetric_with_jacobian_diagonal(shape):
    """With zero boundary data the operator is symmetric, <W, L V> = <V, L W>,
    and its diagonal equals the Jacobian the relaxation uses: after one
    sweep every point of the second colour has zero residual."""
    nz, ny, nx = shape
    V, a, c, _ = _rand(shape, 5)
    W = np.random.default_rng(6).standard_normal(shape)
    for f in (V, W):
        f[0], f[-1], f[:, 0], f[:, -1], f[:, :, 0], f[:, :, -1] = 0, 0, 0, 0, 0, 0
    zero, cm = np.zeros(shape), np.ones(shape)
    LV, LW = _lop(1, V, zero, a, c, cm, H), _lop(1, W, zero, a, c, cm, H)
    s1, s2 = np.sum(W * LV), np.sum(V * LW)
    assert abs(s1 - s2) <= 1e-12 * (np.abs(W * LV).sum())
    # relaxation zeroes the residual of the colour updated last
    rhs = np.random.default_rng(7).standard_normal(shape)
    rho = np.random.default_rng(8).standard_normal(shape)
    L = O.lib()
    L.or_relax(1, O.P(V), O.P(rhs), O.P(rho), O.P(a), O.P(c), O.P(cm), None, EX, nx, ny, nz, *H)
    res = np.zeros(shape)
    L.or_residual(1, O.P(res), O.P(rhs), O.P(V), O.P(rho), O.P(a), O.P(c), O.P(cm), None, EX, nx, ny, nz, *H)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    # pass 1 updates mod(i+j+k+1, 2) == 0 in the reference's 1-based indices
    last = ((i + j + k + 3 + 1) % 2 == 0)[1:-1, 1:-1, 1:-1]
    r = res[1:-1, 1:-1, 1:-1]
    assert np.abs(r[last]).max() <= 1e-10 * np.abs(r[~last]).max()


def _manufactured(n):
    h = 1.0 / (n - 1)
    z, y, x = np.meshgrid(*(np.arange(n) * h,) * 3, indexing="ij")
    pi = np.pi
    u = np.sin(pi * x) * np.sin(pi * y) * np.sin(pi * z)
    ux = pi * np.cos(pi * x) * np.sin(pi * y) * np.sin(pi * z)
    uy = pi * np.sin(pi * x) * np.cos(pi * y) * np.sin(pi * z)
    a = np.exp(x + y)
    c = x.copy()
    rho = a * (-3 * pi * pi * u + ux + uy) + c * u  # div(a grad u) + c u
    return u, a, c, rho


def test_op1_manufactured_second_order():
    """u = sin(pi x) sin(pi y) sin(pi z), a = exp(x + y), c = x: the V-cycle
    solution's max error falls by about 4x per refinement (17 -> 33 -> 65)."""
    errs = []
    for n in (17, 33, 65):
        u, a, c, rho = _manufactured(n)
        m = O.UniformMG(n, op=1)
        m.set_params(pre=2, pst=2)
        m.field(m.nlev, O.F_A)[:] = a
        m.field(m.nlev, O.F_C)[:] = c
        m.field(m.nlev, O.F_RHO)[:] = rho
        m.init()
        norms = m.solve(40, 40)
        assert norms[-1] < 1e-9 * norms[0]
        errs.append(np.abs(m.field(m.nlev, O.F_U) - u).max())
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    print("manufactured errors", errs, "ratios", r1, r2)
    assert 3.6 < r1 < 4.4 and 3.6 < r2 < 4.4


def test_op1_vcycle_reduces_to_poisson():
    """A whole op-1 V-cycle (coefficient restriction included) with a = 1,
    c = 0 gives op 0's per-cycle norms to rounding."""
    n = 33
    rho = -__import__("pamr_amrd_amd").gaussian_source(n, seed=11)
    ms = []
    for op in (0, 1):
        m = O.UniformMG(n, op=op)
        m.set_params(pre=2, pst=2)
        m.field(m.nlev, O.F_RHO)[:] = rho
        if op == 1:
            m.field(m.nlev, O.F_A)[:] = 1.0
            m.field(m.nlev, O.F_C)[:] = 0.0
        m.init()
        ms.append(m)
    for _ in range(4):
        r0 = ms[0].vcycle()[0]
        r1 = ms[1].vcycle()[0]
        assert abs(r0 - r1) <= 1e-10 * r0
    assert np.abs(ms[0].field(ms[0].nlev, O.F_U) - ms[1].field(ms[1].nlev, O.F_U)).max() < 1e-12


# --------------------------------------------------------------------------
# network oracle
# --------------------------------------------------------------------------
def _weights(Hc, seed, scale=0.3):
    rng = np.random.default_rng(seed)
    w1 = scale * rng.standard_normal((Hc, 1, 27))
    b1 = 0.1 * rng.standard_normal(Hc)
    w2 = scale / 2 * rng.standard_normal((Hc, Hc, 27))
    b2 = 0.1 * rng.standard_normal(Hc)
    w3 = scale * rng.standard_normal((1, Hc, 27))
    b3 = 0.1 * rng.standard_normal(1)
    return [w1, b1, w2, b2, w3, b3]


def _torch_net(ws, x, t=None):
    torch = pytest.importorskip("torch")
    F = torch.nn.functional
    Hc = ws[1].shape[0]
    tw = [torch.tensor(w, dtype=torch.float64, requires_grad=True) for w in ws]
    w1 = tw[0].reshape(Hc, 1, 3, 3, 3)
    w2 = tw[2].reshape(Hc, Hc, 3, 3, 3)
    w3 = tw[4].reshape(1, Hc, 3, 3, 3)
    xt = torch.tensor(x, dtype=torch.float64)[None, None]
    h = F.gelu(F.conv3d(xt, w1, tw[1], padding=1), approximate="tanh")
    h = F.gelu(F.conv3d(h, w2, tw[3], padding=1), approximate="tanh")
    y = F.conv3d(h, w3, tw[5], padding=1)[0, 0]
    if t is None:
        return y.detach().numpy()
    loss = 0.5 * ((y - torch.tensor(t, dtype=torch.float64)) ** 2).mean()
    loss.backward()
    return float(loss.detach()), [w.grad.numpy() for w in tw]


@pytest.mark.parametrize("Hc,shape", [(16, (6, 5, 7)), (4, (5, 5, 5))])
def test_net_forward_against_torch(Hc, shape):
    ws = _weights(Hc, 21)
    x = np.random.default_rng(22).standard_normal(shape)
    y = O.net_forward(*ws, x)
    yt = _torch_net(ws, x)
    assert np.abs(y - yt).max() <= 1e-12 * max(1.0, np.abs(yt).max())


@pytest.mark.parametrize("Hc", [4, 16])
def test_net_grad_against_torch(Hc):
    ws = _weights(Hc, 31)
    rng = np.random.default_rng(32)
    x = rng.standard_normal((5, 5, 5))
    t = rng.standard_normal((5, 5, 5))
    loss, g = O.net_loss_grad(*ws, x, t)
    lt, gt = _torch_net(ws, x, t)
    assert abs(loss - lt) <= 1e-12 * abs(lt)
    for a, b in zip(g, gt):
        assert np.abs(a - b.reshape(a.shape)).max() <= 1e-10 * max(1e-3, np.abs(b).max())


def test_net_grad_finite_differences():
    """Central differences of the loss for every parameter of a 4-channel
    network (657 parameters) and every bias plus 40 random weights per tensor
    of the 16-channel one."""
    rng = np.random.default_rng(41)
    x = rng.standard_normal((5, 5, 5))
    t = rng.standard_normal((5, 5, 5))
    for Hc, nsample in ((4, None), (16, 40)):
        ws = _weights(Hc, 40 + Hc)
        _, g = O.net_loss_grad(*ws, x, t)
        for ti, (w, gw) in enumerate(zip(ws, g)):
            flat = w.reshape(-1)
            gflat = gw.reshape(-1)
            idx = np.arange(flat.size)
            if nsample is not None and flat.size > nsample:
                idx = np.random.default_rng(ti).choice(flat.size, nsample, replace=False)
            floor = 1e-3 * np.abs(gflat).max()

            def central(q, eps):
                old = flat[q]
                flat[q] = old + eps
                lp = O.net_loss_grad(*ws, x, t)[0]
                flat[q] = old - eps
                lm = O.net_loss_grad(*ws, x, t)[0]
                flat[q] = old
                return (lp - lm) / (2 * eps)

            for q in idx:
                # Richardson extrapolation of two central differences: O(eps^4)
                eps = 1e-4 * max(1.0, abs(flat[q]))
                fd = (4 * central(q, eps / 2) - central(q, eps)) / 3
                assert abs(fd - gflat[q]) <= 1e-6 * max(abs(fd), floor), (Hc, ti, q, fd, gflat[q])
