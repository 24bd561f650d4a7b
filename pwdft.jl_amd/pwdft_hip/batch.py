"""diag_LOBPCG! over all k-points at once (diag_LOBPCG.jl:1-225 run in lockstep over k).

The reference diagonalises k-point after k-point; with many small k-points (hundreds of
k-points of a few hundred plane waves and a handful of bands) every step of that loop is a
tiny launch.  The library's pw_diag_lobpcg_batch_dev iterates all k-points that have not
converged yet together: every operation of an iteration is one batched launch (op_H of all
their bands in one fused launch where the grid fits a workgroup's LDS, the Gram products,
the column updates, Kprec), the small dense problems of all k-points are solved together
in stacked device buffers, each phase of an iteration is replayed from a captured hipGraph
once it has run, and only the eigenvalues of the convergence test reach the host.  A
k-point leaves the batch when it meets the reference's convergence test; one that reaches
the reference's lock branch takes it on its own.  The arithmetic per k-point is that of
diag_LOBPCG_."""
import numpy as np

from ._lib import ptr
from .solvers import EPS, MIXED_TOL_FLOOR


def _pp(vals):
    """Device pointers as a contiguous uint64 array: the layout of the C ABI's void*[]
    arguments.  The caller keeps the array alive until the call returns."""
    return np.ascontiguousarray(vals, dtype=np.uint64)


def diag_LOBPCG_batch(Ham, psiks, tol=1e-5, NiterMax=100, Nstates_conv=0, mixed=None, iks=None):
    """diag_LOBPCG!(Ham, psiks) with all k-points iks (default: all) iterated together, through
    the C ABI (pw_diag_lobpcg_batch_dev): one batched launch per operation for the k-points
    still iterating, each phase replayed from a captured hipGraph once it has run, one
    synchronisation per iteration.  psiks: list of DeviceMatrix (Ngw[iks[i]] x Nst),
    overwritten with the Ritz vectors.  Returns (evals Nst x nk, stats per k-point
    {"iter", "ncols", "locks"})."""
    from ._lib import PW_MIXED, PW_TUNE_LOBPCG_LOCK
    mixed = Ham.mixed if mixed is None else mixed
    if mixed:
        tol = max(tol, MIXED_TOL_FLOOR)
    iks = list(range(Ham.pw.Nkpt)) if iks is None else list(iks)
    nk = len(iks)
    Nst = psiks[0].cols
    ik = np.ascontiguousarray(iks, dtype=np.int32)
    X = _pp([psiks[i].p.value for i in range(nk)])
    ld = np.ascontiguousarray([psiks[i].rows for i in range(nk)], dtype=np.int64)
    ev = np.zeros((nk, Nst))          # column-major Nst x nk
    st = np.zeros((nk, 3), dtype=np.int32)
    # EPS (the 10 eps lock threshold's unit) may be raised to exercise the lock branch
    lock = 0 if EPS == np.finfo(float).eps else int(round(-np.log10(EPS)))
    Ham.ctx.call("pw_tune_set", PW_TUNE_LOBPCG_LOCK, lock)
    Ham.ctx.call("pw_diag_lobpcg_batch_dev", nk, ptr(ik), Ham.ispin, Nst, ptr(X), ptr(ld), float(tol), int(NiterMax),
                 int(Nstates_conv), PW_MIXED if mixed else 0, ptr(ev), ptr(st))
    stats = [{"iter": int(st[k, 0]), "ncols": int(st[k, 1]), "locks": int(st[k, 2])} for k in range(nk)]
    return ev.T.copy(), stats
