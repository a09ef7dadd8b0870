"""History warm start (pmg_history_warm_start): the least-squares fit of the
new source by earlier ones, and the solution combination it sets.

Pins: a source inside the span of the history is reproduced (coefficients
and misfit exact to rounding); on random data the coefficients and misfit
equal numpy's least squares; on the drifting sequence the fit is never worse
than second-order extrapolation (which is one candidate combination), and
solves started from it converge to the same tolerance in no more cycles.
"""
import numpy as np
import pytest

import pamr_amrd_amd as pm
from pamr_amrd_amd import sequence as SQ

pytestmark = pytest.mark.gpu


def solver_with_rho(torch, n, rho):
    S = pm.Solver(n)
    S.set_params(pre=2, pst=2)
    SQ._copy_grid(S.field_ptr(S.nlev, pm.PMG_F_RHO), pm.tptr(rho), n)
    return S


def read_u(torch, S, n):
    u = torch.empty((n, n, n), dtype=torch.float64, device="cuda")
    SQ._copy_grid(pm.tptr(u), S.field_ptr(S.nlev, pm.PMG_F_U), n)
    return u


def test_span_is_reproduced(torch_cuda):
    torch = torch_cuda
    n = 33
    g = torch.Generator(device="cuda").manual_seed(5)
    F = [torch.randn((n, n, n), dtype=torch.float64, device="cuda", generator=g) for _ in range(3)]
    U = [torch.randn((n, n, n), dtype=torch.float64, device="cuda", generator=g) for _ in range(3)]
    rho = 2.0 * F[0] - F[1] + 0.5 * F[2]
    S = solver_with_rho(torch, n, rho)
    print("case warm start: span, n=33, K=3", flush=True)
    c, rel = S.history_warm_start([pm.tptr(f) for f in F], [pm.tptr(u) for u in U])
    assert np.allclose(c, [2.0, -1.0, 0.5], rtol=0, atol=1e-10), c
    assert rel < 1e-12, rel
    want = c[0] * U[0] + c[1] * U[1] + c[2] * U[2]
    got = read_u(torch, S, n)
    assert float(torch.max(torch.abs(got - want))) <= 1e-14 * float(torch.max(torch.abs(want)))
    S.close()


@pytest.mark.parametrize("K", [1, 4, 8])
def test_matches_numpy_least_squares(torch_cuda, K):
    torch = torch_cuda
    n = 17
    rng = np.random.default_rng(K)
    Fh = rng.standard_normal((K, n, n, n))
    rho_h = rng.standard_normal((n, n, n))
    F = [torch.from_numpy(f).cuda() for f in Fh]
    U = [torch.from_numpy(rng.standard_normal((n, n, n))).cuda() for _ in range(K)]
    S = solver_with_rho(torch, n, torch.from_numpy(rho_h).cuda())
    print(f"case warm start: lstsq, n=17, K={K}", flush=True)
    c, rel = S.history_warm_start([pm.tptr(f) for f in F], [pm.tptr(u) for u in U])
    A = Fh.reshape(K, -1).T
    cref = np.linalg.lstsq(A, rho_h.reshape(-1), rcond=None)[0]
    rel_ref = np.linalg.norm(rho_h.reshape(-1) - A @ cref) / np.linalg.norm(rho_h)
    # the normal equations lose digits with the conditioning of the basis;
    # the misfit is stationary at the optimum, so it agrees much more tightly
    err_c = float(np.max(np.abs(c - cref)) / np.max(np.abs(cref)))
    print(f"K={K}: max coefficient deviation {err_c:.2e}, misfit {rel:.15f} vs {rel_ref:.15f}", flush=True)
    assert err_c <= 1e-7, (c, cref)
    assert abs(rel - rel_ref) <= 1e-12, (rel, rel_ref)
    S.close()


def test_drifting_sequence(torch_cuda):
    torch = torch_cuda
    n = 65
    seq = SQ.Sequence(n=n, seed=1234, speed=0.002, hist_k=6)
    its = {"proj": [], "extrap2": []}
    for s in range(10):
        rref = seq.load_source(s)
        if seq.nhist >= 2:
            for method in ("extrap2", "proj"):
                seq.warm_start(method)
                _, h = seq.solve(0, 1e-10, rref, max_iter=40)
                assert h[-1] <= 1e-10 * rref
                its[method].append(len(h))
            # misfits: the projection's versus extrapolation's, measured directly
            e = seq.rho_cur - (2.0 * seq.hist_f[0] - seq.hist_f[1])
            rel_ex = float(torch.linalg.vector_norm(e) / torch.linalg.vector_norm(seq.rho_cur))
            seq.warm_start("proj")
            print(f"solve {s}: misfit proj {seq.last_misfit:.3e} extrap2 {rel_ex:.3e}", flush=True)
            assert seq.last_misfit <= rel_ex * (1 + 1e-9)
        else:
            seq.warm_start("zero")
        _, h = seq.solve(0, 1e-10, rref, max_iter=40)
        seq.push_solution()
    print(f"cycles: {its}", flush=True)
    assert sum(its["proj"]) <= sum(its["extrap2"])
    assert seq.last_misfit < 1e-4
