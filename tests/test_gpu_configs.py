"""The workloads of the headline configurations, at their configured sizes,
each checked against what fixes it (the oracle where it finishes in seconds,
the FP64 residual and the untouched excised data elsewhere).

* 129^3 base + two levels of 24-48 boxes of sides 17-97, variable
  coefficient: two batched V-cycles bit for bit against the oracle on every
  grid (amrd/mg.c:84-201 over the hierarchy).
* 257^3 drifting sequence: the network trained on 4 solves, 3 later solves
  in plain FP64, FP32 defect correction and with the learned step, each to a
  1e-10 relative FP64 residual (the solve loop, amrd/mg.c:56-72); while the
  learned step is rejected, its solution equals the FP32 one bit for bit.
* 513^3 with three excised spheres: two right-hand sides by flexible GMRES
  and by FP32 defect correction to 1e-10, the excised points still holding
  their Dirichlet data.
* many right-hand sides of the excised problem solved concurrently (one
  solver per stream, one host thread each): every solution the same bits as
  the same right-hand side solved alone.
"""
import numpy as np
import pytest

import oracle as O
import pamr_amrd_amd as pm
from pamr_amrd_amd import hierarchy as H
from pamr_amrd_amd import sequence as SQ

pytestmark = pytest.mark.gpu

BOX = (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)


def test_amr129_two_cycles_bit_exact(torch_cuda):
    levels = H.random_boxes(base_n=129, nref=2, nmin=24, nmax=48, smin=17, smax=97, seed=7)
    print(f"case amr129 boxes={[len(l) for l in levels]}", flush=True)
    m = O.AmrMG(129, levels, op=1)
    a = pm.Amr(129, levels, op=1)
    assert a.nlev == m.nlev
    fields = {O.OA_RHO: pm.PMG_F_RHO, O.OA_A: pm.PMG_F_A, O.OA_C: pm.PMG_F_C}
    for L in range(m.base, m.nlev + 1):
        for g in range(m.ngrids(L)):
            shape, bb = m.grid(L, g)
            vals = {O.OA_RHO: -H.gaussian_on(shape, bb, seed=1234), O.OA_A: H.lognormal_on(shape, bb),
                    O.OA_C: H.uniform_on(shape, bb)}
            for w, v in vals.items():
                m.field(L, g, w)[:] = v
                a.upload(L, g, fields[w], v)
    for x in (m, a):
        x.set_params()
        x.init()
    for c in range(2):
        rm, nm, cm = m.vcycle()
        ra, na, ca = a.vcycle()
        print(f"cycle {c}: oracle {rm:.15e} gpu {ra:.15e}", flush=True)
        assert abs(ra - rm) <= 1e-12 * rm
        assert np.allclose(na, nm, rtol=1e-12, atol=0)
        assert ca == cm
    for L in range(1, m.nlev + 1):
        for g in range(m.ngrids(L)):
            assert np.array_equal(a.download(L, g, pm.PMG_F_U), m.field(L, g, O.OA_U)), (L, g)
    a.close()


def test_poisson257_sequence_modes(torch_cuda):
    torch = torch_cuda
    n = 257
    seq = SQ.Sequence(n=n, seed=1234, speed=0.002)
    print("case sequence 257: collect on 4 solves", flush=True)
    X, Y = seq.collect(nsolves=4, pairs=512, side=16, cycles=4)
    net = pm.Net(seed=3)
    SQ.train(net, X, Y, side=16, steps=300, batch=32, lr=2e-3, margin=3)
    seq.S.set_net(net)
    L = pm.lib()
    u1 = torch.empty((n, n, n), dtype=torch.float64, device="cuda")
    u2 = torch.empty_like(u1)
    for s in range(4, 7):
        results = {}
        for mode in (0, 1, 2):
            before = seq.S.net_stats()
            rref = seq.prepare(s)
            r0, hist = seq.solve(mode, 1e-10, rref)
            assert len(hist) and hist[-1] <= 1e-10 * rref, (s, mode, hist)
            results[mode] = hist
            after = seq.S.net_stats()
            if mode == 1:
                SQ._copy_grid(pm.tptr(u1), seq.S.field_ptr(seq.Lf, pm.PMG_F_U), n)
            if mode == 2:
                SQ._copy_grid(pm.tptr(u2), seq.S.field_ptr(seq.Lf, pm.PMG_F_U), n)
                if after["accepted"] == before["accepted"]:
                    # every learned step of this solve was rejected (or not tried):
                    # it is the FP32 solve, bit for bit
                    assert torch.equal(u1, u2), s
            print(f"solve {s} mode {mode}: {len(hist)} iterations, final {hist[-1] / rref:.3e}, "
                  f"net {after}", flush=True)
        seq.push_solution()
    net.close()


def excised_solver(torch, n):
    """the excised variable-coefficient problem of the 513^3 configuration at
    n^3: three seeded spheres, a lognormal, c uniform in [0, 1]"""
    spheres = H.excision_spheres(seed=21)
    S = pm.Solver(n, op=1, do_ex=1, ex=1.0)
    S.set_params(pre=2, pst=2, max_citer=5000, crtol=1e-3, rop_u=3, rop_res=3, rop_src=3)
    Lf = S.nlev
    for lev in range(1, Lf + 1):
        S.set_chr(lev, H.chr_on(S.shape(lev), BOX, spheres))
    shape = (n, n, n)
    a = H.lognormal_on_device(shape, BOX, torch)
    SQ._copy_grid(S.field_ptr(Lf, pm.PMG_F_A), pm.tptr(a), n)
    del a
    gen = torch.Generator(device="cuda").manual_seed(5)
    c = torch.rand(shape, generator=gen, dtype=torch.float64, device="cuda")
    SQ._copy_grid(S.field_ptr(Lf, pm.PMG_F_C), pm.tptr(c), n)
    del c
    ch = torch.from_numpy(H.chr_on(shape, BOX, spheres)).cuda()
    exc = ch == 1.0
    u0 = torch.where(exc, torch.from_numpy(H.excision_data(shape, BOX)).cuda(), torch.zeros_like(ch))
    return S, exc, u0


def test_excised513_two_rhs(torch_cuda):
    torch = torch_cuda
    n = 513
    S, exc, u0 = excised_solver(torch, n)
    Lf = S.nlev
    L = pm.lib()
    u = torch.empty_like(u0)
    for r in range(2):
        src = SQ.DriftingSource(seed=1234 + r, speed=0.0)
        rho = src.rho(n, 0, torch)
        SQ._copy_grid(S.field_ptr(Lf, pm.PMG_F_RHO), pm.tptr(rho), n)
        del rho
        for how in ("fgmres", "mixed"):
            print(f"case excised513 rhs {r} {how}", flush=True)
            L.pmg_set_solution_lincomb(S.h, pm.tptr(u0), None, 1.0, 0.0)
            S.init()
            if how == "fgmres":
                r0, hist = S.solve_fgmres(8, 1e-10, 400)
            else:
                r0, hist = S.solve_rel(1, 1e-10, 400)
            assert len(hist) and hist[-1] <= 1e-10 * r0, (r, how, len(hist), hist[-3:])
            SQ._copy_grid(pm.tptr(u), S.field_ptr(Lf, pm.PMG_F_U), n)
            assert torch.equal(u[exc], u0[exc]), (r, how)
            print(f"rhs {r} {how}: {len(hist)} iterations, final {hist[-1] / r0:.3e}", flush=True)
    S.close()


def test_excised_concurrent_matches_alone(torch_cuda):
    """five right-hand sides of the excised 129^3 problem on two solvers,
    each on its own stream and host thread: every solution and residual
    history equals the same right-hand side solved alone, bit for bit"""
    from pamr_amrd_amd import multirhs as MR
    torch = torch_cuda
    n = 129
    solvers = []
    for _ in range(3):
        S, exc, u0 = excised_solver(torch, n)
        solvers.append(S)
    alone, conc = solvers[0], solvers[1:]
    rhs = [SQ.DriftingSource(seed=1234 + r, speed=0.0).rho(n, 0, torch) for r in range(5)]
    torch.cuda.synchronize()
    outs = [torch.empty_like(u0) for _ in rhs]
    print("case excised concurrent 129: 5 rhs on 2 streams", flush=True)
    res = MR.solve_fgmres_concurrent(conc, [pm.tptr(r) for r in rhs], pm.tptr(u0), 8, 1e-10, 200,
                                     out=[pm.tptr(o) for o in outs])
    L = pm.lib()
    u = torch.empty_like(u0)
    for k, r in enumerate(rhs):
        SQ._copy_grid(alone.field_ptr(alone.nlev, pm.PMG_F_RHO), pm.tptr(r), n)
        L.pmg_set_solution_lincomb(alone.h, pm.tptr(u0), None, 1.0, 0.0)
        alone.init()
        r0, h = alone.solve_fgmres(8, 1e-10, 200)
        SQ._copy_grid(pm.tptr(u), alone.field_ptr(alone.nlev, pm.PMG_F_U), n)
        print(f"rhs {k}: {len(h)} steps alone, {len(res[k][1])} concurrent", flush=True)
        assert res[k][0] == r0 and np.array_equal(res[k][1], h), k
        assert torch.equal(outs[k], u), k
        assert torch.equal(outs[k][exc], u0[exc]), k
    for S in solvers:
        S.close()
