"""The learned-smoother network on MFMA: the BF16 MFMA lane layout, the
forward pass and the gradients of the training kernels against the FP64
CPU oracle (or_net_forward / or_net_loss_grad), and a training sanity run.
Tolerances reflect BF16 activations with FP32 accumulation."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pm(torch_cuda):
    import pamr_amrd_amd as pm
    pm.lib()
    return pm


def test_mfma_lane_layout(pm):
    rng = np.random.default_rng(0)
    # values exactly representable in bf16 and products/sums exact in fp32
    A = rng.integers(-8, 9, (16, 16)).astype(np.float32) / 8
    B = rng.integers(-8, 9, (16, 16)).astype(np.float32) / 4
    D = pm.mfma_selftest(A, B)
    assert np.array_equal(D, A @ B)
    # the K = 32 instruction of the inference kernels
    A2 = rng.integers(-8, 9, (16, 32)).astype(np.float32) / 8
    B2 = rng.integers(-8, 9, (32, 16)).astype(np.float32) / 4
    assert np.array_equal(pm.mfma_selftest(A2, B2), A2 @ B2)


def _rand_weights(rng, scale=1.0):
    w = {"w1": rng.standard_normal((16, 1, 27)) * 0.3 * scale, "b1": rng.standard_normal(16) * 0.1,
         "w2": rng.standard_normal((16, 16, 27)) * 0.08 * scale, "b2": rng.standard_normal(16) * 0.1,
         "w3": rng.standard_normal((1, 16, 27)) * 0.1 * scale, "b3": rng.standard_normal(1) * 0.1}
    return w


def _bf16(a):
    a32 = np.ascontiguousarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def test_forward_and_gradients_match_oracle(pm, torch_cuda):
    torch = torch_cuda
    rng = np.random.default_rng(42)
    shape = (7, 9, 11)   # (nz, ny, nx), non-multiples of 16 to exercise the edges
    nb = 2
    ws = _rand_weights(rng)
    # weights rounded to bf16 (what the MFMA sees); the oracle uses the same
    wsb = {k: (_bf16(v) if k.startswith("w") else v) for k, v in ws.items()}
    x = _bf16(rng.standard_normal((nb,) + shape))
    t = rng.standard_normal((nb,) + shape)
    net = pm.Net(seed=3)
    net.set_weights(ws)
    xd = torch.from_numpy(x.astype(np.float32)).cuda()
    td = torch.from_numpy(t.astype(np.float32)).cuda()
    yd = torch.zeros_like(xd)
    net.forward(xd, yd, shape, nb)
    y = yd.cpu().numpy().astype(np.float64)
    loss, g = net.loss_grad(xd, td, shape, nb)
    yref = np.stack([O.net_forward(*[wsb[k] for k in ("w1", "b1", "w2", "b2", "w3", "b3")],
                                   np.ascontiguousarray(x[b])) for b in range(nb)])
    err_y = np.linalg.norm(y - yref) / np.linalg.norm(yref)
    print(f"forward rel err {err_y:.3e}")
    assert err_y < 2e-2
    lref, gref = 0.0, None
    for b in range(nb):
        lb, gb = O.net_loss_grad(*[wsb[k] for k in ("w1", "b1", "w2", "b2", "w3", "b3")],
                                 np.ascontiguousarray(x[b]), np.ascontiguousarray(t[b]))
        lref += lb / nb
        gref = [gi / nb for gi in gb] if gref is None else [a + gi / nb for a, gi in zip(gref, gb)]
    assert abs(loss - lref) <= 2e-2 * abs(lref)
    for name, gr in zip(("w1", "b1", "w2", "b2", "w3", "b3"), gref):
        e = np.linalg.norm(g[name] - gr) / np.linalg.norm(gr)
        print(f"grad {name} rel err {e:.3e}")
        assert e < 5e-2, name


@pytest.mark.parametrize("shape,nb", [((7, 9, 11), 2), ((20, 17, 35), 1)])
def test_inference_path_matches_oracle_and_forward(pm, torch_cuda, shape, nb):
    """The inference path of pmg_net_correct (layer 1 with the 27 taps as
    the MFMA's K, no stored pre-activations, channel 0 only at the end)
    against the FP64 oracle network and against the training forward
    (the same BF16 inputs and weights; FP32 summation order differs, and a
    BF16 activation may round the other way)."""
    torch = torch_cuda
    rng = np.random.default_rng(7)
    ws = _rand_weights(rng)
    wsb = {k: (_bf16(v) if k.startswith("w") else v) for k, v in ws.items()}
    scale = 0.5
    x = rng.standard_normal((nb,) + shape)
    xs = _bf16(x * scale)
    net = pm.Net(seed=3)
    net.set_weights(ws)
    xd = torch.from_numpy(x.astype(np.float32)).cuda()
    yi = torch.zeros_like(xd)
    net.infer(xd, yi, shape, nb, scale=scale)
    yf = torch.zeros_like(xd)
    net.forward(torch.from_numpy((x * scale).astype(np.float32)).cuda(), yf, shape, nb)
    yi, yf = yi.cpu().numpy().astype(np.float64), yf.cpu().numpy().astype(np.float64)
    yref = np.stack([O.net_forward(*[wsb[k] for k in ("w1", "b1", "w2", "b2", "w3", "b3")],
                                   np.ascontiguousarray(xs[b])) for b in range(nb)])
    e_ref = np.linalg.norm(yi - yref) / np.linalg.norm(yref)
    e_fwd = np.linalg.norm(yi - yf) / np.linalg.norm(yf)
    print(f"inference vs oracle {e_ref:.3e}, vs training forward {e_fwd:.3e}", flush=True)
    assert e_ref < 2e-2
    assert e_fwd < 5e-3


def test_training_reduces_loss(pm, torch_cuda):
    torch = torch_cuda
    rng = np.random.default_rng(7)
    shape = (16, 16, 16)
    nb = 8
    x = rng.standard_normal((nb,) + shape)
    # target: a fixed local linear map of x (7-point average) -- learnable
    t = x.copy()
    for ax in (1, 2, 3):
        t = t + 0.5 * (np.roll(x, 1, axis=ax) + np.roll(x, -1, axis=ax))
    xd = torch.from_numpy(x.astype(np.float32)).cuda()
    td = torch.from_numpy(t.astype(np.float32)).cuda()
    net = pm.Net(seed=11)
    losses = [net.train_step(xd, td, shape, nb, 3e-3, margin=2) for _ in range(300)]
    print(f"loss {losses[0]:.4e} -> {losses[-1]:.4e}")
    assert losses[-1] < 0.1 * losses[0]
