"""What the trained network gains: the learned correction of a drifting
sequence, trained on the GPU from plain FP64 V-cycles of the first solves,
evaluated on later (held-out) solves.

Measured on held-out states (after k plain V-cycles of a warm-started solve):
the error ||u* - u|| (u* a tightly converged FP64 solution) and the FP64
residual after the learned correction u += ||r|| h^2 N(r/||r||), relative to
their values before it, for the trained network and for the untrained
(initial) one, printed next to the same ratios of one plain V-cycle.  The
trained network must reduce the error and beat the untrained one; whether a
correction pays for its cost inside a solve is decided there, every cycle, by
the FP64 residual (pmg_solve_rel mode 2).
"""
import numpy as np
import pytest

import pamr_amrd_amd as pm
from pamr_amrd_amd import sequence as SQ

pytestmark = pytest.mark.gpu


def apply_and_measure(seq, net, r, ustar):
    """Apply u += ||r|| h^2 N(r/||r||); return (error ratio, residual ratio)
    against ||u* - u|| and ||r|| before it; the correction is then undone."""
    L = pm.lib()
    S, Lf, n, torch = seq.S, seq.Lf, seq.n, seq.torch
    h2 = seq.h * seq.h
    u = torch.empty((n, n, n), dtype=torch.float64, device="cuda")
    res = torch.empty_like(u)
    SQ._copy_grid(pm.tptr(u), S.field_ptr(Lf, pm.PMG_F_U), n)
    e0 = float(torch.linalg.vector_norm(ustar - u))
    SQ._copy_grid(pm.tptr(res), S.field_ptr(Lf, pm.PMG_F_RES), n)
    x = res.to(torch.float32)
    torch.cuda.synchronize()
    assert L.pmg_net_correct(net.h, pm.tptr(x), 1.0 / r, S.field_ptr(Lf, pm.PMG_F_U), r * h2, n, n, n, None) == 0
    torch.cuda.synchronize()
    rn = S.residual()
    SQ._copy_grid(pm.tptr(u), S.field_ptr(Lf, pm.PMG_F_U), n)
    e1 = float(torch.linalg.vector_norm(ustar - u))
    assert L.pmg_net_recorrect(net.h, S.field_ptr(Lf, pm.PMG_F_U), -r * h2, n, n, n, None) == 0
    torch.cuda.synchronize()
    S.residual()
    return e1 / e0, rn / r


def test_trained_network_gain(torch_cuda):
    torch = torch_cuda
    n = 65
    seq = SQ.Sequence(n=n, seed=1234, speed=0.004)
    print("case learned gain: collect", flush=True)
    X, Y = seq.collect(nsolves=8, pairs=2048, side=16, cycles=4)
    assert X.shape[0] >= 1024 and torch.isfinite(X).all() and torch.isfinite(Y).all()
    trained = pm.Net(seed=3)
    untrained = pm.Net(seed=3)
    print("case learned gain: train", flush=True)
    losses = SQ.train(trained, X, Y, side=16, steps=1000, batch=32, lr=2e-3, margin=3)
    first, last = float(np.mean(losses[:50])), float(np.mean(losses[-100:]))
    print(f"training loss {first:.4g} -> {last:.4g}", flush=True)
    assert last < 0.8 * first
    ratios = {"trained": [], "untrained": [], "vcycle": []}
    u0 = torch.empty((n, n, n), dtype=torch.float64, device="cuda")
    ustar = torch.empty_like(u0)
    L = pm.lib()
    for s in range(8, 12):
        rref = seq.prepare(s)
        SQ._copy_grid(pm.tptr(u0), seq.S.field_ptr(seq.Lf, pm.PMG_F_U), n)
        seq.solve(0, 1e-13, rref, max_iter=60)
        SQ._copy_grid(pm.tptr(ustar), seq.S.field_ptr(seq.Lf, pm.PMG_F_U), n)
        L.pmg_set_solution_lincomb(seq.S.h, pm.tptr(u0), None, 1.0, 0.0)
        seq.S.init()
        torch.cuda.synchronize()
        for k in range(3):
            r = seq.S.residual()
            ratios["trained"].append(apply_and_measure(seq, trained, r, ustar))
            ratios["untrained"].append(apply_and_measure(seq, untrained, r, ustar))
            u = torch.empty_like(u0)
            SQ._copy_grid(pm.tptr(u), seq.S.field_ptr(seq.Lf, pm.PMG_F_U), n)
            e0 = float(torch.linalg.vector_norm(ustar - u))
            rv, _, _ = seq.S.vcycle()
            SQ._copy_grid(pm.tptr(u), seq.S.field_ptr(seq.Lf, pm.PMG_F_U), n)
            ratios["vcycle"].append((float(torch.linalg.vector_norm(ustar - u)) / e0, rv / r))
        L.pmg_set_solution_lincomb(seq.S.h, pm.tptr(ustar), None, 1.0, 0.0)
        torch.cuda.synchronize()
        seq.push_solution()
    gm = {k: tuple(float(np.exp(np.mean(np.log([p[i] for p in v])))) for i in (0, 1)) for k, v in ratios.items()}
    print("learned gain, geometric means per application (error ratio, residual ratio): "
          + ", ".join(f"{k} {v[0]:.4f} {v[1]:.4f}" for k, v in gm.items()), flush=True)
    # the trained correction removes part of the error the untrained one leaves
    assert gm["trained"][0] < 1.0
    assert gm["trained"][0] < gm["untrained"][0]
    trained.close()
    untrained.close()
