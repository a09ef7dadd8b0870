"""Many right-hand sides of one problem, solved concurrently.

The reference solves one right-hand side per ``solve_elliptics`` call
(amrd/mg.c:20-79).  Here a batch of right-hand sides of the same operator
(same grid, coefficients and excision mask) is spread over a few solver
instances, each with its own HIP stream and driven by its own host thread
(ctypes releases the interpreter lock during library calls).  Every solve is
the ordinary single-RHS solve -- the same kernels, the same arithmetic, so
each right-hand side gets the same bits as when solved alone -- but the
streams overlap: one solver's launch gaps, host-side Krylov steps and small
bottom levels run while another's large levels stream through HBM.

Nothing here touches torch's streams: sources, initial guess and outputs
are device pointers, and every copy runs on the owning solver's stream.
"""
from __future__ import annotations

import ctypes as C
import threading

from . import PMG_F_RHO, PMG_F_U, lib

_hip = None


def _stream_sync(stream):
    global _hip
    if _hip is None:
        lib()  # loads the HIP runtime the library links
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipStreamSynchronize.argtypes = [C.c_void_p]
    if _hip.hipStreamSynchronize(stream) != 0:
        raise RuntimeError("hipStreamSynchronize failed")


def _copy(dst, src, n, stream):
    rc = lib().pmg_dmcopy3d(src, dst, n, n, n, n, n, n, 1, 1, 1, 1, 1, 1, n, n, n, 1, 1, 1, 1, 1, 1, stream)
    if rc != 0:
        raise RuntimeError(f"pmg_dmcopy3d returned {rc}")


def solve_fgmres_concurrent(solvers, rhs, u0, m, rtol, max_iter, out=None):
    """Flexible GMRES(m) with the FP32 V-cycle as preconditioner for every
    right-hand side (solve_concurrent with that solve)."""
    return solve_concurrent(solvers, rhs, u0, lambda S: S.solve_fgmres(m, rtol, max_iter), rtol, out)


def solve_concurrent(solvers, rhs, u0, solve, rtol, out=None):
    """Solve finest-level sources ``rhs`` (device pointers to n^3 FP64 grids)
    with ``solve(S)`` (a relative solve of the solver's current source and
    initial guess returning (reference norm, residual history)), each to
    ||r|| <= rtol ||r(u0)||.  Right-hand side k runs on ``solvers[k %
    len(solvers)]``; all solvers must hold the same problem (shape,
    coefficients, excision mask, parameters).  ``u0``: device pointer of the
    initial guess (the excised Dirichlet data), or None for zero.  ``out``:
    optional device pointers receiving each solution.  Returns the list of
    (reference norm, residual history) per right-hand side; raises if any
    solve fails or does not converge."""
    if not solvers:
        raise ValueError("no solvers")
    n = solvers[0].shape(solvers[0].nlev)[0]
    for S in solvers:
        if S.shape(S.nlev) != (n, n, n):
            raise ValueError("all solvers must have the same finest shape")
    if out is not None and len(out) != len(rhs):
        raise ValueError("out must have one pointer per right-hand side")
    L = lib()
    results = [None] * len(rhs)
    errors = []

    def work(si):
        S = solvers[si]
        st = C.c_void_p(S.stream())
        try:
            for k in range(si, len(rhs), len(solvers)):
                _copy(S.field_ptr(S.nlev, PMG_F_RHO), rhs[k], n, st)
                if u0 is None:
                    S.zero_solution()
                elif L.pmg_set_solution_lincomb(S.h, u0, None, 1.0, 0.0) != 0:
                    raise RuntimeError("pmg_set_solution_lincomb failed")
                S.init()
                r0, h = solve(S)
                if len(h) == 0 or h[-1] > rtol * r0:
                    raise RuntimeError(f"right-hand side {k} did not converge in {len(h)} steps")
                if out is not None:
                    _copy(out[k], S.field_ptr(S.nlev, PMG_F_U), n, st)
                results[k] = (r0, h)
            _stream_sync(st)
        except Exception as e:  # reported after every thread has finished
            errors.append((si, e))

    threads = [threading.Thread(target=work, args=(i,)) for i in range(len(solvers))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    if errors:
        si, e = errors[0]
        raise RuntimeError(f"solver {si}: {e}") from e
    return results
