"""The learned band correction of excised problems on the GPU: the band plan,
the correction bit for bit against numpy with the same network outputs, the
defect it leaves equal to a full recomputation, the weighted training loss
against torch's FP64 autograd, the per-iteration fallback, and what the
trained network gains (held-out right-hand sides, time to 1e-10)."""
import ctypes as C
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BOX = (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)


@pytest.fixture(scope="module")
def pm(torch_cuda):
    import pamr_amrd_amd as pm
    pm.lib()
    return pm


def _problem(pm, n, op=1, seed=21):
    from pamr_amrd_amd import hierarchy as H
    spheres = H.excision_spheres(seed=seed)
    S = pm.Solver(n, op=op, do_ex=1, ex=1.0)
    S.set_params(pre=2, pst=2, max_citer=5000, crtol=1e-3, rop_u=3, rop_res=3, rop_src=3)
    for lev in range(1, S.nlev + 1):
        S.set_chr(lev, H.chr_on(S.shape(lev), BOX, spheres))
    shape = (n, n, n)
    ch = H.chr_on(shape, BOX, spheres)
    if op == 1:
        S.upload(S.nlev, pm.PMG_F_A, H.lognormal_on(shape, BOX))
        S.upload(S.nlev, pm.PMG_F_C, H.uniform_on(shape, BOX))
    u0 = np.where(ch == 1.0, H.excision_data(shape, BOX), 0.0)
    return S, ch, u0


def _rho(n, seed):
    from pamr_amrd_amd import hierarchy as H
    rho = -H.gaussian_on((n, n, n), BOX, seed=seed)
    rho[0], rho[-1] = 0, 0
    rho[:, 0], rho[:, -1], rho[:, :, 0], rho[:, :, -1] = 0, 0, 0, 0
    return rho


def _rand_weights(rng):
    return {"w1": rng.standard_normal((16, 1, 27)) * 0.3, "b1": rng.standard_normal(16) * 0.1,
            "w2": rng.standard_normal((16, 16, 27)) * 0.08, "b2": rng.standard_normal(16) * 0.1,
            "w3": rng.standard_normal((1, 16, 27)) * 0.1, "b3": rng.standard_normal(1) * 0.1}


def _host_cores(pm, mask, core, margin=3):
    nz, ny, nx = mask.shape
    cap = (-(-nx // core)) * (-(-ny // core)) * (-(-nz // core))
    org = np.zeros(3 * cap, dtype=np.int32)
    m = pm.lib().pmg_band_cores(np.ascontiguousarray(mask).ctypes.data_as(C.POINTER(C.c_ubyte)), nx, ny, nz,
                                core, margin, org.ctypes.data_as(C.POINTER(C.c_int)), cap)
    return org[:3 * m].reshape(m, 3)


def _crops(src64, org, side):
    """crop c, zero outside the grid ((z, y, x) arrays, origins (i, j, k))"""
    nz, ny, nx = src64.shape
    out = np.zeros((len(org), side, side, side), dtype=src64.dtype)
    for c, (i0, j0, k0) in enumerate(org):
        zs, ys, xs = slice(max(k0, 0), min(k0 + side, nz)), slice(max(j0, 0), min(j0 + side, ny)), \
            slice(max(i0, 0), min(i0 + side, nx))
        out[c, zs.start - k0:zs.stop - k0, ys.start - j0:ys.stop - j0, xs.start - i0:xs.stop - i0] = \
            src64[zs, ys, xs]
    return out


def test_band_plan_matches_host_mask(pm, torch_cuda):
    n = 65
    S, ch, _ = _problem(pm, n)
    nc = S.band_setup(3, 26)
    mask = S.band_mask()
    band = np.zeros(ch.shape, dtype=np.uint8)
    cnt = pm.lib().pmg_band_mask(np.ascontiguousarray(ch).ctypes.data_as(C.POINTER(C.c_double)), n, n, n, 1.0, 3,
                                 band.ctypes.data_as(C.POINTER(C.c_ubyte)))
    npts, nc2, side = S.band_info()
    assert np.array_equal(mask, band) and npts == cnt > 0
    assert nc == nc2 == len(_host_cores(pm, band, 26)) and side == 32
    S.close()


@pytest.mark.parametrize("op", [0, 1])
def test_dc_correction_bit_exact(pm, torch_cuda, op):
    """u += ts ||r|| N(r/||r||) at the band points and nowhere else, bit for
    bit against numpy with the network run on numpy-gathered crops; the
    defect -r it leaves in the FP32 source equals a full recomputation."""
    torch = torch_cuda
    n = 33
    S, ch, u0 = _problem(pm, n, op=op)
    S.upload(S.nlev, pm.PMG_F_RHO, _rho(n, 1234))
    S.upload(S.nlev, pm.PMG_F_U, u0)
    S.init()
    core = 12
    nc = S.band_setup(3, core)
    side = core + 6
    net = pm.Net(seed=5)
    net.set_weights(_rand_weights(np.random.default_rng(1)))
    ts = 0.37
    S.set_band_net(net, ts, use_dc=True)
    _, h = S.solve_rel(1, 0.0, 2)
    r = float(h[-1])
    Lf = S.nlev
    u_before = S.download(Lf, pm.PMG_F_U)
    rho32 = S.download32(Lf, 1)
    mask = S.band_mask()
    org = _host_cores(pm, mask, core)
    assert len(org) == nc > 0
    x = (-1.0 / r * _crops(rho32.astype(np.float64), org, side)).astype(np.float32)
    xd = torch.from_numpy(x).cuda()
    yd = torch.zeros_like(xd)
    net.infer(xd, yd, (side, side, side), nc, scale=1.0)
    y = yd.cpu().numpy().astype(np.float64)
    S.band_correct(0)
    u_after = S.download(Lf, pm.PMG_F_U)
    want = u_before.copy()
    for c, (i0, j0, k0) in enumerate(org):
        sl = tuple(slice(max(o + 3, 0), min(o + 3 + core, n)) for o in (k0, j0, i0))
        ys = y[c][tuple(slice(s.start - o, s.stop - o) for s, o in zip(sl, (k0, j0, i0)))]
        m = mask[sl] == 1
        region = want[sl]
        region[m] = u_before[sl][m] + (ts * r) * ys[m]
    changed = u_after != u_before
    print(f"op {op}: {nc} cores, {mask.sum()} band points, {changed.sum()} changed", flush=True)
    assert changed.sum() > 0 and not changed[mask == 0].any()
    assert np.array_equal(u_after, want)
    got32 = S.download32(Lf, 1)
    S.solve_rel(1, 0.0, 0)  # recomputes the FP64 defect of the current u everywhere
    assert np.array_equal(got32, S.download32(Lf, 1))
    S.close()


def test_preconditioner_correction(pm, torch_cuda):
    """The preconditioner form: the FP32 z changes at band points only, by
    ts ||r|| N(r/||r||) of its FP32 residual (norm folded on the device)."""
    torch = torch_cuda
    n = 33
    S, ch, _ = _problem(pm, n, op=1)
    core = 12
    nc = S.band_setup(3, core)
    side = core + 6
    net = pm.Net(seed=5)
    net.set_weights(_rand_weights(np.random.default_rng(2)))
    ts = 0.5
    S.set_band_net(net, ts, use_dc=False, use_fgmres=True)
    rng = np.random.default_rng(4)
    upd = np.zeros(ch.shape, dtype=bool)
    upd[1:-1, 1:-1, 1:-1] = True
    upd &= ch != 1.0
    S.upload32(S.nlev, 1, np.where(upd, rng.standard_normal(ch.shape), 0.0))
    z0 = np.where(upd, 1e-3 * rng.standard_normal(ch.shape), 0.0).astype(np.float32)
    S.upload32(S.nlev, 0, z0)
    S.band_correct(1)
    res = S.download32(S.nlev, 3).astype(np.float64)
    z1 = S.download32(S.nlev, 0)
    rn = np.sqrt((res[upd] ** 2).mean())
    mask = S.band_mask()
    org = _host_cores(pm, mask, core)
    x = (_crops(res, org, side) / rn).astype(np.float32)
    xd = torch.from_numpy(x).cuda()
    yd = torch.zeros_like(xd)
    net.infer(xd, yd, (side, side, side), nc, scale=1.0)
    y = yd.cpu().numpy().astype(np.float64)
    dz = np.zeros(ch.shape)
    for c, (i0, j0, k0) in enumerate(org):
        sl = [slice(max(o + 3, 0), min(o + 3 + core, n)) for o in (k0, j0, i0)]
        ys = y[c][tuple(slice(s.start - o, s.stop - o) for s, o in zip(sl, (k0, j0, i0)))]
        dz[tuple(sl)] = np.where(mask[tuple(sl)] == 1, ts * rn * ys, 0.0)
    got = z1.astype(np.float64) - z0.astype(np.float64)
    assert not (z1 != z0)[mask == 0].any()
    err = np.abs(got - dz).max() / np.abs(dz).max()
    print(f"preconditioner correction rel err {err:.2e}", flush=True)
    # the device folds the norm in another order: an input may round to the
    # neighbouring BF16 value inside the network
    assert err < 1e-2
    S.close()


def test_weighted_loss_matches_torch(pm, torch_cuda):
    """pmg_net_loss_grad_w: the mean over points with w != 0 past the margin,
    and its gradients, against torch's FP64 autograd on the same BF16-rounded
    inputs and weights (tanh GELU, as the training kernels)."""
    torch = torch_cuda
    F = torch.nn.functional
    rng = np.random.default_rng(9)
    shape, nb, margin = (10, 12, 11), 2, 1
    ws = _rand_weights(rng)

    def bf16(a):
        u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
        u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
        return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)

    wsb = {k: (bf16(v) if k.startswith("w") else v) for k, v in ws.items()}
    x = bf16(rng.standard_normal((nb,) + shape))
    t = rng.standard_normal((nb,) + shape)
    w = (rng.random((nb,) + shape) < 0.3).astype(np.float32)
    net = pm.Net(seed=3)
    net.set_weights(ws)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()
    loss, g = net.loss_grad_w(dev(x), dev(t), dev(w), shape, nb, margin)
    tw = [torch.tensor(wsb[k], dtype=torch.float64, requires_grad=True) for k in ("w1", "b1", "w2", "b2", "w3", "b3")]
    xt = torch.tensor(x, dtype=torch.float64)[:, None]
    hh = F.gelu(F.conv3d(xt, tw[0].reshape(16, 1, 3, 3, 3), tw[1], padding=1), approximate="tanh")
    hh = F.gelu(F.conv3d(hh, tw[2].reshape(16, 16, 3, 3, 3), tw[3], padding=1), approximate="tanh")
    y = F.conv3d(hh, tw[4].reshape(1, 16, 3, 3, 3), tw[5], padding=1)[:, 0]
    on = np.zeros((nb,) + shape, dtype=bool)
    on[:, margin:-margin, margin:-margin, margin:-margin] = True
    on &= w != 0
    ont = torch.tensor(on)
    lref = 0.5 * ((y - torch.tensor(t)) ** 2)[ont].sum() / on.sum()
    lref.backward()
    assert abs(loss - float(lref)) <= 2e-2 * abs(float(lref))
    for name, tt in zip(("w1", "b1", "w2", "b2", "w3", "b3"), tw):
        gr = tt.grad.numpy().reshape(g[name].shape)
        e = np.linalg.norm(g[name] - gr) / np.linalg.norm(gr)
        print(f"grad {name} rel err {e:.3e}", flush=True)
        assert e < 5e-2, name


def test_bad_network_falls_back(pm, torch_cuda):
    """An untrained network with large weights makes the first corrected
    iteration worse than a plain one: mode 3 drops it for the rest of the
    solve, which still converges."""
    n = 33
    S, ch, u0 = _problem(pm, n, op=1)
    S.upload(S.nlev, pm.PMG_F_RHO, _rho(n, 1234))
    S.band_setup(3, 12)
    net = pm.Net(seed=5)
    ws = _rand_weights(np.random.default_rng(3))
    ws["w3"] *= 200.0
    net.set_weights(ws)
    S.set_band_net(net, 1.0, use_dc=True)
    for _ in range(2):
        S.upload(S.nlev, pm.PMG_F_U, u0)
        S.init()
        r0, h = S.solve_rel(3, 1e-10, 200)
        assert h[-1] <= 1e-10 * r0
    st = S.band_stats()
    print(st, flush=True)
    assert st["rejected"] >= 1 and st["rho_plain"] > 0
    got = S.download(S.nlev, pm.PMG_F_U)
    assert np.array_equal(got[ch == 1.0], u0[ch == 1.0])
    S.close()


def test_trained_band_net_gains(pm, torch_cuda):
    """Train on (residual, error) crops of two right-hand sides' plain mixed
    iterations (65^3, variable coefficient, three excised spheres), then
    solve two held-out right-hand sides to 1e-10: defect correction with the
    band correction needs fewer iterations and less time than without, and
    flexible GMRES with it in the preconditioner needs no more applications."""
    torch = torch_cuda
    from pamr_amrd_amd import band as B
    n = 65
    S, ch, u0 = _problem(pm, n, op=1)
    S.band_setup(3, 26)
    u0d = torch.from_numpy(u0).cuda()
    rhos = [torch.from_numpy(_rho(n, 1234 + k)).cuda() for k in range(4)]
    ustars = [B.converged_solution(S, rhos[k], u0d, torch) for k in range(2)]
    x, t, w = B.collect_pairs(S, rhos[:2], ustars, u0d, torch, iters=10, crops=48)
    net, ts, info = B.train(x, t, w, torch, steps=1500, log=print)
    print(info, flush=True)
    del x, t, w
    S.set_band_net(net, ts, use_dc=True, use_fgmres=True)
    res = {}
    for mode in (1, 3, 1, 3):
        for k in (2, 3):
            B.set_state(S, rhos[k], u0d)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r0, h = S.solve_rel(mode, 1e-10, 400)
            dt = time.perf_counter() - t0
            assert h[-1] <= 1e-10 * r0, (mode, k, h[-3:])
            res.setdefault(mode, []).append((len(h), dt))
    it1 = np.mean([a for a, _ in res[1]])
    it3 = np.mean([a for a, _ in res[3]])
    t1 = np.median([b for _, b in res[1]])
    t3 = np.median([b for _, b in res[3]])
    st = S.band_stats()
    print(f"mixed: {it1} it {t1 * 1e3:.1f} ms; with band net: {it3} it {t3 * 1e3:.1f} ms; {st}", flush=True)
    assert st["accepted"] > 0
    assert it3 <= 0.8 * it1
    assert t3 < t1
    # flexible GMRES: the first solve measures the plain rate, the next use the net
    apps = []
    for k in (2, 3, 2):
        B.set_state(S, rhos[k], u0d)
        r0, h = S.solve_fgmres(8, 1e-10, 200)
        assert h[-1] <= 1e-10 * r0
        apps.append(len(h))
    print(f"fgmres applications (plain, band, band): {apps} {S.band_stats()}", flush=True)
    assert apps[2] <= apps[0]
    S.close()


def test_gather_res32_bit_exact(pm, torch_cuda):
    """The in-cycle network input: the FP32 residual of u (source rho) at
    every crop point, 0 off the updatable points and outside the grid, over
    the device norm -- bit for bit against numpy's FP32 arithmetic in the
    operator's order."""
    torch = torch_cuda
    from pamr_amrd_amd import hierarchy as H
    n, core = 33, 12
    side = core + 6
    ch = H.chr_on((n, n, n), BOX, H.excision_spheres(seed=21))
    band = np.zeros(ch.shape, dtype=np.uint8)
    pm.lib().pmg_band_mask(np.ascontiguousarray(ch).ctypes.data_as(C.POINTER(C.c_double)), n, n, n, 1.0, 3,
                           band.ctypes.data_as(C.POINTER(C.c_ubyte)))
    org = _host_cores(pm, band, core)
    nc = len(org)
    rng = np.random.default_rng(12)
    u = rng.standard_normal((n, n, n)).astype(np.float32)
    rho = rng.standard_normal((n, n, n)).astype(np.float32)
    norm = 2.5
    ud, rd = torch.from_numpy(u).cuda(), torch.from_numpy(rho).cuda()
    chd = torch.from_numpy((ch == 1.0).astype(np.uint8)).cuda()
    orgd = torch.from_numpy(np.ascontiguousarray(org.reshape(-1))).cuda()
    nd = torch.tensor([norm], dtype=torch.float64, device="cuda")
    out = torch.zeros((nc, side, side, side), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    h = 1.0 / (n - 1)
    rc = pm.lib().pmg_band_gather_res32(0, 1, pm.tptr(ud), pm.tptr(rd), None, None, pm.tptr(chd), n, n, n,
                                        pm.tptr(orgd), nc, side, 1.0, pm.tptr(nd), h, h, h, pm.tptr(out), None)
    assert rc == 0
    torch.cuda.synchronize()
    f = np.float32
    dx2 = f(h) * f(h)
    idx2 = f(1) / dx2
    c = u[1:-1, 1:-1, 1:-1]
    lx = ((u[1:-1, 1:-1, 2:] - f(2) * c) + u[1:-1, 1:-1, :-2]) * idx2
    ly = ((u[1:-1, 2:, 1:-1] - f(2) * c) + u[1:-1, :-2, 1:-1]) * idx2
    lz = ((u[2:, 1:-1, 1:-1] - f(2) * c) + u[:-2, 1:-1, 1:-1]) * idx2
    r = np.zeros((n, n, n), dtype=np.float32)
    r[1:-1, 1:-1, 1:-1] = ((lx + ly) + lz) - rho[1:-1, 1:-1, 1:-1]
    r[ch == 1.0] = 0
    want = ((1.0 / norm) * _crops(r.astype(np.float64), org, side)).astype(np.float32)
    assert np.array_equal(out.cpu().numpy(), want)


def test_in_cycle_band_net_is_checked(pm, torch_cuda):
    """The correction placed inside the FP32 V-cycle, after the finest
    coarse-grid correction (trained on crops taken there), on held-out
    right-hand sides (65^3, variable coefficient, three excised spheres).
    Measured here it does not beat the plain iteration, and the
    per-iteration check drops it: every solve converges, within two
    iterations of the plain one, with the excised data untouched."""
    torch = torch_cuda
    from pamr_amrd_amd import band as B
    n = 65
    S, ch, u0 = _problem(pm, n, op=1)
    S.band_setup(3, 26)
    u0d = torch.from_numpy(u0).cuda()
    rhos = [torch.from_numpy(_rho(n, 1234 + k)).cuda() for k in range(4)]
    ustars = [B.converged_solution(S, rhos[k], u0d, torch) for k in range(2)]
    x, t, w = B.collect_pairs_inner(S, rhos[:2], ustars, u0d, torch, iters=10, crops=48)
    net, ts, info = B.train(x, t, w, torch, steps=1500, log=print)
    print(info, flush=True)
    del x, t, w
    S.set_band_net_inner(net, ts)
    S.band_set_inner(1)
    S.set_band_net(None, 1.0, use_dc=True, use_fgmres=True)
    S.band_enable(True, True)
    its = {}
    for mode in (1, 3):
        for k in (2, 3):
            B.set_state(S, rhos[k], u0d)
            r0, h = S.solve_rel(mode, 1e-10, 400)
            assert h[-1] <= 1e-10 * r0, (mode, k, h[-3:])
            its.setdefault(mode, []).append(len(h))
    st = S.band_stats()
    print(f"in-cycle: mixed {its[1]} iterations, with the correction {its[3]}; {st}", flush=True)
    assert st["accepted"] + st["rejected"] > 0
    if st["accepted"] == 0:
        assert max(its[3]) <= max(its[1]) + 2
    apps = []
    for k in (2, 3, 2):
        B.set_state(S, rhos[k], u0d)
        r0, h = S.solve_fgmres(8, 1e-10, 200)
        assert h[-1] <= 1e-10 * r0
        apps.append(len(h))
    print(f"in-cycle fgmres applications (plain, band, band): {apps} {S.band_stats()}", flush=True)
    got = S.download(S.nlev, pm.PMG_F_U)
    assert np.array_equal(got[ch == 1.0], u0[ch == 1.0])
    S.close()
