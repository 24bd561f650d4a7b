"""diag_LOBPCG! over all k-points at once (diag_LOBPCG.jl:1-225 run in lockstep over k).

The reference diagonalises k-point after k-point; with many small k-points (hundreds of
k-points of a few hundred plane waves and a handful of bands) every step of that loop is a
tiny launch.  Here every step of one LOBPCG iteration is one batched launch for all k-points
that have not converged yet (pw_op_H_batch_dev, pw_zgemm_batch_dev, pw_kprec_batch_dev,
pw_copy_cols_batch_dev), and the small dense problems of all k-points (Nst..3Nst) are
solved together on the device, in stacked buffers that never leave HBM: the Cholesky
factors and their inverses in one call (pw_chol_inv_batch_dev), the Hermitian eigenproblems
of the Rayleigh-Ritz steps in another (pw_heevd_batch_dev).  Only eigenvalues reach the
host, for the convergence test.  A k-point leaves the batch when it meets the
reference's convergence test; one that reaches the reference's lock branch finishes that
iteration through the per-k code (solvers._lobpcg_lock_step).  The arithmetic per k-point is
that of solvers.diag_LOBPCG_."""
import ctypes

import numpy as np

from ._lib import DeviceMatrix, PosDefException, PWError, PW_ERR_NOT_POSDEF, PW_TERM_ALL, ptr
from .solvers import EPS, MIXED_TOL_FLOOR, _Dense, _LockScratch, _lobpcg_lock_step

_ONE = np.array([1.0, 0.0])
_ZERO = np.array([0.0, 0.0])
_MONE = np.array([-1.0, 0.0])


def _pp(vals):
    """Device pointers as a contiguous uint64 array: the layout of the C ABI's void*[]
    arguments.  The caller keeps the array alive until the call returns."""
    return np.ascontiguousarray(vals, dtype=np.uint64)


def _ints(v, n, dtype):
    """Per-product sizes: a scalar repeated n times, or one value per product."""
    if np.ndim(v) == 0:
        return np.full(n, v, dtype=dtype)
    return np.ascontiguousarray(v, dtype=dtype)


class _Slab:
    """One device allocation holding, for every k-point, an Ngw[k] x ncols column-major
    block (leading dimension Ngw[k])."""

    def __init__(self, ctx, rows, ncols):
        self.rows = [int(r) for r in rows]
        self.rows_a = np.asarray(self.rows, dtype=np.uint64)
        self.ncols = ncols
        self.off = np.concatenate([[0], np.cumsum([16 * r * ncols for r in self.rows])]).astype(np.int64)
        self.mem = DeviceMatrix(ctx, int(self.off[-1] // 16) or 1, 1)
        self.base = self.mem.p.value
        self.start = np.uint64(self.base) + self.off[:-1].astype(np.uint64)

    def col(self, k, j):
        return self.base + int(self.off[k]) + 16 * self.rows[k] * int(j)

    def cols(self, ks, j):
        """Column j of the blocks of k-points ks (index array), as device pointers."""
        return self.start[ks] + np.uint64(16 * int(j)) * self.rows_a[ks]

    def view(self, k):
        return _View(self.col(k, 0), self.rows[k], self.ncols)

    def free(self):
        self.mem.free()


class _View:
    """Column-block view with the DeviceMatrix interface the per-k lock step uses."""

    def __init__(self, p, rows, cols):
        self.p = ctypes.c_void_p(p)
        self.rows, self.cols = rows, cols

    def col(self, j):
        return ctypes.c_void_p(self.p.value + 16 * self.rows * int(j))


class _Small:
    """Stacked small matrices (nk x r x c) in one device buffer, column-major per k."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.buf = None

    def ptrs(self, nk, r, c):
        need = nk * r * c
        if self.buf is None or self.buf.rows < need:
            if self.buf is not None:
                self.buf.free()
            self.buf = DeviceMatrix(self.ctx, max(need, 1), 1)
        return np.uint64(self.buf.p.value) + np.uint64(16 * r * c) * np.arange(nk, dtype=np.uint64)

    def download(self, nk, r, c):
        a = np.empty((nk, c, r), dtype=np.complex128)
        self.ctx.call("pw_memcpy_d2h", ptr(a), ctypes.c_void_p(self.buf.p.value), a.nbytes)
        return np.transpose(a, (0, 2, 1))

    def free(self):
        if self.buf is not None:
            self.buf.free()


class _Infos:
    """Device ints that the asynchronous batched Cholesky and eigensolver calls of one
    iteration write their per-matrix status to; read back and checked once per iteration,
    so that those calls do not wait for the device."""

    def __init__(self, ctx, nslots):
        self.ctx = ctx
        self.cap = nslots
        self.mem = DeviceMatrix(ctx, nslots // 4 + 1, 1)
        self.used = []
        self.next = 0

    def take(self, count, what, iks):
        if self.next + count > self.cap:
            self.check()
        off = self.next
        self.next += count
        self.used.append((off, count, what, list(iks)))
        return ctypes.c_void_p(self.mem.p.value + 4 * off)

    def check(self):
        """Read every status written since the last check; raise for the first failure
        (PosDefException for a Cholesky factor, PWError for an eigensolver)."""
        if not self.used:
            return
        h = np.empty(self.next, dtype=np.int32)
        self.ctx.call("pw_memcpy_d2h", ptr(h), self.mem.p, h.nbytes)
        used, self.used, self.next = self.used, [], 0
        for off, cnt, what, iks in used:
            bad = np.flatnonzero(h[off:off + cnt])
            if bad.size == 0:
                continue
            i = int(bad[0])
            if what == "cholesky":
                nb = cnt // 2
                kind = "leading minor" if i < nb else "singular diagonal of the factor"
                raise PosDefException(PW_ERR_NOT_POSDEF, f"cholesky (k-point {iks[i % nb]}): {kind} "
                                                         f"{int(h[off + i])}")
            raise PWError(-1, f"eigen: no convergence for k-point {iks[i]} ({int(h[off + i])})")

    def free(self):
        self.mem.free()


class _Batch:
    """Batched device operations on the active k-points."""

    def __init__(self, Ham, mixed):
        self.Ham = Ham
        self.c = Ham.ctx
        self.mixed = mixed

    def gemm(self, transa, m, n, k, A, lda, B, ldb, C, ldc, alpha=_ONE, beta=_ZERO):
        nb = len(A)
        ms, ns, ks = _ints(m, nb, np.int32), _ints(n, nb, np.int32), _ints(k, nb, np.int32)
        la, lb, lc = _ints(lda, nb, np.int64), _ints(ldb, nb, np.int64), _ints(ldc, nb, np.int64)
        pa, pb, pc = _pp(A), _pp(B), _pp(C)
        self.c.call("pw_zgemm_batch_dev", transa, nb, ptr(ms), ptr(ns), ptr(ks), ptr(alpha), ptr(pa), ptr(la),
                    ptr(pb), ptr(lb), ptr(beta), ptr(pc), ptr(lc))

    def gram(self, small, A, B, na, nbc, rows):
        """A_k^H B_k for column blocks into the stacked buffer `small` (na x nbc per k-point);
        returns the device pointers of the blocks."""
        out = small.ptrs(len(A), na, nbc)
        self.gemm(1, na, nbc, rows, A, rows, B, rows, out, na)
        return out

    def herm(self, p, n, mode):
        """pw_herm_fix_dev on every n x n block at p (mode 0: Hermitian(); 2: (A + A^H) / 2)."""
        self.c.call("pw_herm_fix_batch_dev", n, len(p), ctypes.c_void_p(int(p[0])), n, n * n, mode)

    def chol_inv(self, p, n, iks, infos=None):
        """Every n x n block at p <- inv(cholesky(Hermitian(block)).U), strict lower zeroed.
        With infos the call does not wait; the status is checked at infos.check()."""
        if infos is not None:
            d = infos.take(2 * len(p), "cholesky", iks)
            self.c.call("pw_chol_inv_batch_async_dev", n, len(p), ctypes.c_void_p(int(p[0])), n, n * n, d)
            return
        try:
            self.c.call("pw_chol_inv_batch_dev", n, len(p), ctypes.c_void_p(int(p[0])), n, n * n)
        except PosDefException as e:
            raise PosDefException(e.code, f"{e} (batch of k-points {list(iks)})")

    def eigh(self, p, n):
        """eigen(Hermitian(block)) of every n x n block at p: eigenvectors in place,
        eigenvalues returned (len(p) x n, ascending)."""
        w = np.empty((len(p), n))
        self.c.call("pw_heevd_batch_dev", n, len(p), ctypes.c_void_p(int(p[0])), n, n * n, ptr(w))
        return w

    def eigh_async(self, p, n, w_dev, infos, iks):
        """eigen(Hermitian(block)) of every n x n block at p, eigenvectors in place and
        eigenvalues into the device array w_dev, without waiting (status at infos.check())."""
        d = infos.take(len(p), "eigen", iks)
        self.c.call("pw_heevd_batch_async_dev", n, len(p), ctypes.c_void_p(int(p[0])), n, n * n, w_dev, d)

    def copy(self, rows, ncols, src, dst):
        nb = len(src)
        r = np.ascontiguousarray(rows, dtype=np.int32)
        cl = np.full(nb, ncols, dtype=np.int32)
        ld = np.ascontiguousarray(rows, dtype=np.int64)
        ps, pd = _pp(src), _pp(dst)
        self.c.call("pw_copy_cols_batch_dev", nb, ptr(r), ptr(cl), ptr(ps), ptr(ld), ptr(pd), ptr(ld))

    def opH(self, iks, src, dst, ncols, rows):
        from ._lib import PW_MIXED
        ik = np.ascontiguousarray(iks, dtype=np.int32)
        ld = np.ascontiguousarray(rows, dtype=np.int64)
        ps, pd = _pp(src), _pp(dst)
        self.c.call("pw_op_H_batch_dev", len(iks), ptr(ik), self.Ham.ispin, ncols, ptr(ps), ptr(ld), ptr(pd), ptr(ld),
                    PW_TERM_ALL | (PW_MIXED if self.mixed else 0))

    def kprec(self, iks, src, dst, ncols, rows):
        ik = np.ascontiguousarray(iks, dtype=np.int32)
        ld = np.ascontiguousarray(rows, dtype=np.int64)
        ps, pd = _pp(src), _pp(dst)
        self.c.call("pw_kprec_batch_dev", len(iks), ptr(ik), ncols, ptr(ps), ptr(ld), ptr(pd), ptr(ld))


def _ritz_coefs(B, sM, T, G, nq, Nst, iks, infos, w_dev):
    """Lowest Nst eigenvectors of eigen(Hermitian(T), Hermitian(G)) for every stacked pair
    (diag_LOBPCG.jl:176-190), on the device: G = U^H U, C = U^-H T U^-1, C y = l y,
    x = U^-1 y (the reduction zhegv performs).  T and G (nq x nq blocks) are overwritten;
    returns the pointers of the nq x Nst coefficient blocks (leading dimension nq) in sM."""
    B.herm(T, nq, 0)
    B.chol_inv(G, nq, iks, infos)
    M = sM.ptrs(len(T), nq, nq)
    B.gemm(0, nq, nq, nq, T, nq, G, nq, M, nq)      # T Ui
    B.gemm(1, nq, nq, nq, G, nq, M, nq, T, nq)      # Ui^H T Ui
    B.eigh_async(T, nq, w_dev, infos, iks)
    B.gemm(0, nq, Nst, nq, G, nq, T, nq, M, nq)     # Ui y, lowest Nst
    return M


def diag_LOBPCG_batch(Ham, psiks, tol=1e-5, NiterMax=100, Nstates_conv=0, mixed=None, iks=None):
    """diag_LOBPCG!(Ham, psiks) with all k-points iks (default: all) iterated together.
    psiks: list of DeviceMatrix (Ngw[ik] x Nst), overwritten with the Ritz vectors.
    Returns (evals Nst x nk, stats per k-point {"iter", "ncols", "locks"})."""
    mixed = Ham.mixed if mixed is None else mixed
    if mixed:
        tol = max(tol, MIXED_TOL_FLOOR)
    iks = list(range(Ham.pw.Nkpt)) if iks is None else list(iks)
    nk = len(iks)
    Nst = psiks[0].cols
    if Nstates_conv == 0:
        Nstates_conv = Nst
    rows = [psiks[i].rows for i in range(nk)]
    c = Ham.ctx
    B = _Batch(Ham, mixed)
    Q = _Slab(c, rows, 3 * Nst)
    HQ = _Slab(c, rows, 3 * Nst)
    Tmp = _Slab(c, rows, 4 * Nst)
    sX, sC, sT, sG, sM = (_Small(c) for _ in range(5))   # stacked small matrices in HBM
    infos = _Infos(c, 8 * nk)                             # statuses of one iteration's dense calls
    w_dev = DeviceMatrix(c, (3 * Nst * nk + 1) // 2, 1)   # Rayleigh-Ritz eigenvalues (unused on the host)
    allk = list(range(nk))
    rows_a = np.asarray(rows, dtype=np.int64)
    cols = lambda S, ks, j: S.cols(np.asarray(ks, dtype=np.int64), j)
    rws = lambda ks: rows_a[np.asarray(ks, dtype=np.int64)]
    B.copy(rws(allk), Nst, [psiks[k].p.value for k in allk], cols(Q, allk, 0))
    B.opH([iks[k] for k in allk], cols(Q, allk, 0), cols(HQ, allk, 0), Nst, rws(allk))
    ncols = np.full(nk, Nst)
    iters = np.zeros(nk, dtype=int)
    nlocks = np.zeros(nk, dtype=int)
    evals_old = np.zeros((nk, Nst))
    SMALL = 10.0 * EPS
    lock_ws = None
    active = list(allk)
    for it in range(1, NiterMax + 1):
        if not active:
            break
        a = active
        ka = [iks[k] for k in a]
        iters[a] = it
        x, hx, w, hw = cols(Q, a, 0), cols(HQ, a, 0), cols(Q, a, Nst), cols(HQ, a, Nst)
        pX = B.gram(sX, x, hx, Nst, Nst, rws(a))
        B.herm(pX, Nst, 0)
        infos.check()                                   # the previous iteration's factorisations
        evals = np.linalg.eigvalsh(sX.download(len(a), Nst, Nst))   # eigvals(Hermitian(XHX))
        devals = np.abs(evals - evals_old[a])
        evals_old[a] = evals
        # R = HX - X*XHX (into the W slot)
        B.copy(rws(a), Nst, hx, w)
        B.gemm(0, rws(a), Nst, Nst, x, rws(a), pX, Nst, w, rws(a), alpha=_MONE, beta=_ONE)
        nconv = np.count_nonzero(devals < tol, axis=1)
        done = nconv >= Nstates_conv
        lock = (~done) & np.any(devals <= SMALL, axis=1)
        keep = [k for k, d, l in zip(a, done, lock) if not d and not l]
        locked = [k for k, d, l in zip(a, done, lock) if l]
        if locked:
            for k in locked:
                B.kprec([iks[k]], [Q.col(k, Nst)], [Q.col(k, Nst)], Nst, [rows[k]])
                D = _Dense(Ham, mixed)
                Ham.ik = iks[k]
                dv = devals[a.index(k)]
                ilock = np.flatnonzero(dv <= SMALL)
                iact = np.flatnonzero(dv > SMALL)
                nlocks[k] += 1
                if lock_ws is None:
                    lock_ws = _LockScratch(c, max(rows), Nst)
                ncols[k] += _lobpcg_lock_step(D, Q.view(k), HQ.view(k), Tmp.view(k), rows[k], Nst, ilock, iact,
                                              scratch=lock_ws)
        active = keep + locked
        active.sort()
        a = keep
        if not a:
            continue
        ka = [iks[k] for k in a]
        ra = rws(a)
        w, hw = cols(Q, a, Nst), cols(HQ, a, Nst)
        B.kprec(ka, w, w, Nst, ra)
        B.opH(ka, w, hw, Nst, ra)
        ncols[a] += Nst
        pC = B.gram(sC, w, w, Nst, Nst, ra)
        B.chol_inv(pC, Nst, ka, infos)                   # inv(cholesky(Hermitian(W'W)).U)
        t0, t1 = cols(Tmp, a, 0), cols(Tmp, a, Nst)
        B.gemm(0, ra, Nst, Nst, w, ra, pC, Nst, t0, ra)
        B.gemm(0, ra, Nst, Nst, hw, ra, pC, Nst, t1, ra)
        B.copy(ra, Nst, t0, w)
        B.copy(ra, Nst, t1, hw)
        nq = 3 * Nst if it > 1 else 2 * Nst
        qa, hqa = cols(Q, a, 0), cols(HQ, a, 0)
        pT = B.gram(sT, qa, hqa, nq, nq, ra)
        pG = B.gram(sG, qa, qa, nq, nq, ra)
        pU = _ritz_coefs(B, sM, pT, pG, nq, Nst, ka, infos, w_dev.p)
        B.gemm(0, ra, Nst, nq, qa, ra, pU, nq, t0, ra)
        B.gemm(0, ra, Nst, nq, hqa, ra, pU, nq, t1, ra)
        x, hx, p, hp = cols(Q, a, 0), cols(HQ, a, 0), cols(Q, a, 2 * Nst), cols(HQ, a, 2 * Nst)
        if it > 1:
            pU23 = pU + np.uint64(16 * Nst)             # rows Nst:3Nst of the coefficients
            t2, t3 = cols(Tmp, a, 2 * Nst), cols(Tmp, a, 3 * Nst)
            B.gemm(0, ra, Nst, 2 * Nst, w, ra, pU23, nq, t2, ra)
            B.gemm(0, ra, Nst, 2 * Nst, hw, ra, pU23, nq, t3, ra)
            B.copy(ra, Nst, t0, x)
            B.copy(ra, Nst, t1, hx)
            pC = B.gram(sC, t2, t2, Nst, Nst, ra)
            B.chol_inv(pC, Nst, ka, infos)
            B.gemm(0, ra, Nst, Nst, t2, ra, pC, Nst, p, ra)
            B.gemm(0, ra, Nst, Nst, t3, ra, pC, Nst, hp, ra)
        else:
            B.copy(ra, Nst, w, p)
            B.copy(ra, Nst, hw, hp)
            B.copy(ra, Nst, t0, x)
            B.copy(ra, Nst, t1, hx)
    infos.check()
    # final Rayleigh-Ritz of every k-point
    x, hx = cols(Q, allk, 0), cols(HQ, allk, 0)
    pS = B.gram(sX, x, hx, Nst, Nst, rws(allk))
    B.herm(pS, Nst, 2)
    ev = B.eigh(pS, Nst)
    B.gemm(0, rws(allk), Nst, Nst, x, rws(allk), pS, Nst, [psiks[k].p.value for k in allk], rws(allk))
    c.call("pw_synchronize")
    for s in (Q, HQ, Tmp, sX, sC, sT, sG, sM, infos, w_dev):
        s.free()
    if lock_ws is not None:
        lock_ws.free()
    stats = [{"iter": int(iters[k]), "ncols": int(ncols[k]), "locks": int(nlocks[k])} for k in allk]
    return ev.T.copy(), stats
