"""Uniform-hierarchy V-cycles with excision (chr masks on every level, the
_ex transfers, Dirichlet data held on the excised points) on the GPU vs the
oracle: per-cycle norms within 1e-12 relative, u bit for bit."""
import numpy as np
import pytest

import oracle as O
import pamr_amrd_amd as pm
from pamr_amrd_amd import hierarchy as H

pytestmark = pytest.mark.gpu

BOX = (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)


@pytest.mark.parametrize("n,op", [(33, 0), (33, 1), (65, 0), (65, 1)])
def test_excised_cycles_bit_exact(torch_cuda, n, op):
    print(f"case excision n={n} op={op}", flush=True)
    spheres = H.excision_spheres(seed=21 + n)
    m = O.UniformMG(n, op=op, do_ex=1, ex=1.0)
    S = pm.Solver(n, op=op, do_ex=1, ex=1.0)
    for lev in range(1, m.nlev + 1):
        shape = m.shape(lev)
        ch = H.chr_on(shape, BOX, spheres)
        m.field(lev, O.F_CHR)[:] = ch
        S.set_chr(lev, ch)
    shape = (n, n, n)
    rho = -H.gaussian_on(shape, BOX)
    rho[0], rho[-1] = 0, 0
    rho[:, 0], rho[:, -1], rho[:, :, 0], rho[:, :, -1] = 0, 0, 0, 0
    ch = H.chr_on(shape, BOX, spheres)
    u0 = np.where(ch == 1.0, H.excision_data(shape, BOX), 0.0)
    for w, v in ((O.F_RHO, rho), (O.F_U, u0)):
        m.field(m.nlev, w)[:] = v
    S.upload(S.nlev, pm.PMG_F_RHO, rho)
    S.upload(S.nlev, pm.PMG_F_U, u0)
    if op == 1:
        a = H.lognormal_on(shape, BOX)
        c = H.uniform_on(shape, BOX)
        m.field(m.nlev, O.F_A)[:] = a
        m.field(m.nlev, O.F_C)[:] = c
        S.upload(S.nlev, pm.PMG_F_A, a)
        S.upload(S.nlev, pm.PMG_F_C, c)
    for x in (m, S):
        x.set_params()
        x.init()
    for cyc in range(6):
        rm, nm, cm = m.vcycle()
        rs, ns, cs = S.vcycle()
        print(f"cycle {cyc}: oracle {rm:.15e} gpu {rs:.15e}", flush=True)
        assert abs(rs - rm) <= 1e-12 * rm, (cyc, rs, rm)
        assert cs == cm
    got = S.download(S.nlev, pm.PMG_F_U)
    assert np.array_equal(got, m.field(m.nlev, O.F_U))
    # the excised points still hold the Dirichlet data
    assert np.array_equal(got[ch == 1.0], u0[ch == 1.0])
    S.close()
