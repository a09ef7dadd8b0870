"""Flexible GMRES with the FP32 V-cycle as right preconditioner
(pmg_solve_fgmres).

Checks: the solve reaches the relative tolerance by the true FP64 residual;
its solution agrees with the plain FP64 V-cycle solution to the accuracy
both tolerances allow; it needs no more preconditioner applications than
mixed-precision defect correction needs outer iterations (GMRES minimises
the residual over the same preconditioned directions that defect correction
takes one at a time); with excision the Dirichlet data on the excised points
are untouched; small restart lengths still converge.
"""
import numpy as np
import pytest

import pamr_amrd_amd as pm
from pamr_amrd_amd import hierarchy as H

pytestmark = pytest.mark.gpu

BOX = (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)


def setup(n, op, ex, seed=21):
    S = pm.Solver(n, op=op, do_ex=1 if ex else 0, ex=1.0)
    shape = (n, n, n)
    rho = -H.gaussian_on(shape, BOX)
    rho[0], rho[-1] = 0, 0
    rho[:, 0], rho[:, -1], rho[:, :, 0], rho[:, :, -1] = 0, 0, 0, 0
    u0 = np.zeros(shape)
    ch = None
    if ex:
        spheres = H.excision_spheres(seed=seed)
        for lev in range(1, S.nlev + 1):
            S.set_chr(lev, H.chr_on(S.shape(lev), BOX, spheres))
        ch = H.chr_on(shape, BOX, spheres)
        u0 = np.where(ch == 1.0, H.excision_data(shape, BOX), 0.0)
    S.upload(S.nlev, pm.PMG_F_RHO, rho)
    if op == 1:
        S.upload(S.nlev, pm.PMG_F_A, H.lognormal_on(shape, BOX))
        S.upload(S.nlev, pm.PMG_F_C, H.uniform_on(shape, BOX))
    S.set_params(pre=2, pst=2)
    return S, u0, ch


def start(S, u0):
    S.upload(S.nlev, pm.PMG_F_U, u0)
    S.init()


@pytest.mark.parametrize("n,op,ex", [(65, 0, False), (65, 1, True), (129, 1, True)])
def test_fgmres_converges(torch_cuda, n, op, ex):
    print(f"case fgmres n={n} op={op} ex={ex}", flush=True)
    S, u0, ch = setup(n, op, ex)
    rtol = 1e-10
    start(S, u0)
    r0m, hm = S.solve_rel(1, rtol, 300)
    u_mixed = S.download(S.nlev, pm.PMG_F_U)
    start(S, u0)
    r0g, hg = S.solve_fgmres(8, rtol, 300)
    u_g = S.download(S.nlev, pm.PMG_F_U)
    print(f"mixed: {len(hm)} iterations; fgmres(8): {len(hg)} preconditioner applications, "
          f"final {hg[-1] / r0g:.3e}", flush=True)
    assert abs(r0m - r0g) <= 1e-12 * r0m
    assert hg[-1] <= rtol * r0g
    assert S.residual() <= rtol * r0g
    assert len(hg) <= len(hm)
    scale = np.max(np.abs(u_mixed))
    assert np.max(np.abs(u_g - u_mixed)) <= 1e-7 * scale
    if ch is not None:
        assert np.array_equal(u_g[ch == 1.0], u0[ch == 1.0])
    S.close()


def test_short_restarts(torch_cuda):
    S, u0, _ = setup(65, 1, True)
    start(S, u0)
    r0, h = S.solve_fgmres(2, 1e-10, 400)
    print(f"fgmres(2): {len(h)} applications", flush=True)
    assert h[-1] <= 1e-10 * r0
    S.close()
