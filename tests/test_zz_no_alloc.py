"""pmg_hook_relax_noalloc (partial-norm scratch in a fixed device array, no
allocation) against pmg_hook_relax on the same inputs: two calls in a row,
the second on a smaller grid, both bit-identical in V and in the norm; an
oversized grid is refused on the host."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(torch, shape, op, seed):
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    f = {
        "V": rng.standard_normal((nz, ny, nx)),
        "rho": rng.standard_normal((nz, ny, nx)),
        "rhs": rng.standard_normal((nz, ny, nx)),
        "a": np.exp(0.5 * rng.standard_normal((nz, ny, nx))),
        "c": rng.random((nz, ny, nx)),
        "cm": np.where(rng.random((nz, ny, nx)) < 0.9, 1.0, 0.0),
        "chr": np.where(rng.random((nz, ny, nx)) < 0.1, 2.0, 0.0),
    }
    return {k: torch.from_numpy(v).to("cuda") for k, v in f.items()}


def _relax(pm, fn, d, shape, op, ex, nrm):
    nx, ny, nz = shape
    P = pm.tptr
    return fn(op, P(d["V"]), P(d["rhs"]), P(d["rho"]), P(d["a"]) if op else None, P(d["c"]) if op else None,
              P(d["cm"]), P(d["chr"]) if ex else None, 2.0, nx, ny, nz, 1.0 / (nx - 1), 0.5 / (ny - 1),
              1.0 / (nz - 1), P(nrm), None)


@pytest.mark.parametrize("op,ex", [(0, 0), (1, 1)])
def test_noalloc_relax_matches_allocating_hook(torch_cuda, op, ex):
    torch = torch_cuda
    import pamr_amrd_amd as pm
    L = pm.lib()
    for call, shape in enumerate([(129, 97, 65), (33, 17, 21)]):
        print(f"case noalloc op={op} ex={ex} call={call} shape={shape}", flush=True)
        d_ref = _case(torch, shape, op, 100 * call + op)
        d_new = {k: v.clone() for k, v in d_ref.items()}
        n_ref = torch.zeros(1, dtype=torch.float64, device="cuda")
        n_new = torch.zeros(1, dtype=torch.float64, device="cuda")
        for sweep in range(2):
            assert _relax(pm, L.pmg_hook_relax, d_ref, shape, op, ex, n_ref) == 0
            assert _relax(pm, L.pmg_hook_relax_noalloc, d_new, shape, op, ex, n_new) == 0
            torch.cuda.synchronize()
            assert torch.equal(d_ref["V"], d_new["V"]), (call, sweep)
            assert n_ref.item() == n_new.item() and n_ref.item() > 0, (call, sweep, n_ref.item(), n_new.item())


def test_noalloc_relax_refuses_oversized_grid(torch_cuda):
    torch = torch_cuda
    import pamr_amrd_amd as pm
    dummy = torch.zeros(8, dtype=torch.float64, device="cuda")
    p = pm.tptr(dummy)
    # 513^3 interior points exceed the scratch: refused before any launch
    rc = pm.lib().pmg_hook_relax_noalloc(0, p, p, p, None, None, p, None, 1.0, 515, 515, 515, 1.0, 1.0, 1.0,
                                         p, None)
    assert rc == -2
