"""Block eigensolvers of PWDFT.jl on the device: diag_LOBPCG! (diag_LOBPCG.jl:1-225),
diag_davidson! (diag_davidson.jl:1-191) and diag_davidson_qe! (diag_davidson_qe_v2.jl:1-216),
plus ortho_sqrt (ortho_sqrt.jl:1-9).

Everything Ngw-sized (X, W, P, their H-images, residuals, the Gram/overlap products and
the Q*U updates) stays in device memory and runs through libpwdft_hip; only the small
dense FP64 problems the reference hands to LAPACK (Hermitian eigensolves of size
Nst..3Nst, Cholesky of Nst x Nst, triangular inverses) run on the host, as the reference
also runs them serially.  X, W and P are column blocks of one buffer Q = [X W P] so that
Q'HQ and Q'Q are single products."""
import ctypes

import numpy as np
import scipy.linalg as sla

from ._lib import PosDefException, DeviceMatrix, PW_MIXED, PW_TERM_ALL, PW_TERM_K, PW_TUNE_LOBPCG_LOCK, ptr
from .hamiltonian import op_H

EPS = np.finfo(float).eps

_ONE = np.array([1.0, 0.0])
_ZERO = np.array([0.0, 0.0])
_MONE = np.array([-1.0, 0.0])


def herm_u(A):
    """Julia Hermitian(A): upper triangle mirrored, real diagonal."""
    U = np.triu(A, 1)
    return U + U.conj().T + np.diag(np.real(np.diag(A)))


class _Dense:
    """Device products for one k-point: gemm wrappers and small-matrix staging."""

    def __init__(self, Ham, mixed):
        self.Ham = Ham
        self.c = Ham.ctx
        self.flags = PW_MIXED if mixed else 0
        self.small = None

    def gemm(self, transa, m, n, k, A, lda, B, ldb, C, ldc, alpha=_ONE, beta=_ZERO):
        self.c.call("pw_zgemm_dev", transa, m, n, k, ptr(alpha), A, lda, B, ldb, ptr(beta), C, ldc,
                    self.flags)

    def gram(self, A, B, ncolA, ncolB, rows):
        """A' B for column blocks (device pointers) -> host ncolA x ncolB."""
        out = self._small(ncolA, ncolB)
        self.gemm(1, ncolA, ncolB, rows, A, rows, B, rows, out.p, ncolA)
        a = np.empty((ncolA, ncolB), dtype=np.complex128, order="F")
        self.c.call("pw_memcpy_d2h", ptr(a), out.p, 16 * ncolA * ncolB)
        return a

    def _small(self, r, c):
        need = r * c
        if self.small is None or self.small.rows * self.small.cols < need:
            self.small = DeviceMatrix(self.c, max(need, 1), 1)
        return self.small

    def upload_small(self, M):
        """Upload a host matrix into a fresh device matrix (column-major)."""
        return DeviceMatrix.from_host(self.c, np.asfortranarray(M, dtype=np.complex128))

    def copy(self, rows, ncols, src, dst):
        self.c.call("pw_copy_cols_dev", rows, ncols, src, rows, dst, rows)

    def opH(self, src, dst, ncols, rows):
        self.c.call("pw_op_H_dev", self.Ham.ik, self.Ham.ispin, ncols, src, rows, dst, rows,
                    PW_TERM_ALL | (PW_MIXED if self.Ham.mixed else 0))

    def kprec(self, src, dst, ncols, rows):
        self.c.call("pw_kprec_dev", self.Ham.ik, ncols, src, rows, dst, rows)


def _chol_upper_inv(C):
    """cholesky(Hermitian(C)).U and its inverse (PosDefException on failure)."""
    try:
        U = sla.cholesky(herm_u(C), lower=False)
    except np.linalg.LinAlgError as e:
        raise PosDefException(-4, f"cholesky: {e}")
    return U, sla.solve_triangular(U, np.eye(U.shape[0], dtype=complex), lower=False)


def _colptr(M, j):
    return M.col(j)


def _lobpcg_device(Ham, X, tol, NiterMax, Nstates_conv, mixed):
    """diag_LOBPCG! through the C ABI (pw_diag_lobpcg_dev): every block, Gram matrix, small
    dense problem and Ritz coefficient stays in HBM; the host receives the Ritz values once
    per iteration for the |dlambda| test."""
    Nst = X.cols
    ev = np.zeros(Nst)
    st = np.zeros(3, dtype=np.int32)
    # EPS (the 10 eps lock threshold's unit) may be raised to exercise the lock branch
    lock = 0 if EPS == np.finfo(float).eps else int(round(-np.log10(EPS)))
    Ham.ctx.call("pw_tune_set", PW_TUNE_LOBPCG_LOCK, lock)
    Ham.ctx.call("pw_diag_lobpcg_dev", Ham.ik, Ham.ispin, Nst, X.p, X.rows, float(tol), int(NiterMax),
                 int(Nstates_conv), PW_MIXED if mixed else 0, ev.ctypes.data_as(ctypes.c_void_p),
                 st.ctypes.data_as(ctypes.c_void_p))
    return ev, {"iter": int(st[0]), "ncols": int(st[1]), "locks": int(st[2])}


DENSE = "device"   # where diag_LOBPCG_ solves its small dense problems: "device" or "host"
# In mixed precision the FP32 products resolve eigenvalue changes to about 1e-7 Ha; below
# that the |dlambda| test cannot be met and the block vectors turn to noise, so the
# tolerance is not taken lower (the FP64 iterations that follow reach the final accuracy).
MIXED_TOL_FLOOR = 1e-6


def diag_LOBPCG_(Ham, X, tol=1e-5, NiterMax=100, Nstates_conv=0, mixed=None, dense=None):
    """diag_LOBPCG! for the current Ham.ik (diag_LOBPCG.jl:51-225).  X is a DeviceMatrix
    (Ngw x Nst, orthonormal) overwritten with the Ritz vectors.  Returns (evals, stats)
    where stats = {"iter": iterations, "ncols": op_H columns applied, "locks": iterations
    that took the lock branch}.  dense = "device" (default) runs the whole solver in the
    library (pw_diag_lobpcg_dev), its small dense problems on the device; "host" keeps the
    Ngw-sized products on the device and solves the small dense problems with LAPACK on
    the host, as the reference does."""
    mixed = Ham.mixed if mixed is None else mixed
    if Nstates_conv == 0:
        Nstates_conv = X.cols
    if mixed:
        tol = max(tol, MIXED_TOL_FLOOR)
    if (DENSE if dense is None else dense) == "device":
        return _lobpcg_device(Ham, X, tol, NiterMax, Nstates_conv, mixed)
    D = _Dense(Ham, mixed)
    D64 = _Dense(Ham, False)
    Nb, Nst = X.rows, X.cols
    if Nstates_conv == 0:
        Nstates_conv = Nst
    c = Ham.ctx
    Q = DeviceMatrix(c, Nb, 3 * Nst)
    HQ = DeviceMatrix(c, Nb, 3 * Nst)
    Tmp = DeviceMatrix(c, Nb, 4 * Nst)
    x, w, p = Q.col(0), Q.col(Nst), Q.col(2 * Nst)
    hx, hw, hp = HQ.col(0), HQ.col(Nst), HQ.col(2 * Nst)
    D.copy(Nb, Nst, X.p, x)
    D.opH(x, hx, Nst, Nb)
    ncols = Nst
    nlock = 0
    lock_ws = None
    evals_old = np.zeros(Nst)
    SMALL = 10.0 * EPS
    it_done = 0
    for it in range(1, NiterMax + 1):
        it_done = it
        XHX = herm_u(D.gram(x, hx, Nst, Nst, Nb))
        evals = np.linalg.eigvalsh(XHX)
        devals = np.abs(evals - evals_old)
        evals_old = evals.copy()
        # R = HX - X*XHX  (into the W slot)
        D.copy(Nb, Nst, hx, w)
        dX = D.upload_small(XHX)
        D.gemm(0, Nb, Nst, Nst, x, Nb, dX.p, Nst, w, Nb, alpha=_MONE, beta=_ONE)
        nconv = int(np.count_nonzero(devals < tol))
        ilock = np.flatnonzero(devals <= SMALL)
        iact = np.flatnonzero(devals > SMALL)
        if nconv >= Nstates_conv:
            break
        D.kprec(w, w, Nst, Nb)
        if len(ilock) > 0:
            nlock += 1
            if lock_ws is None:
                lock_ws = _LockScratch(c, Nb, Nst)
            ncols += _lobpcg_lock_step(D, Q, HQ, Tmp, Nb, Nst, ilock, iact, scratch=lock_ws)
            continue
        D.opH(w, hw, Nst, Nb)
        ncols += Nst
        U, Ui = _chol_upper_inv(D64.gram(w, w, Nst, Nst, Nb))
        dUi = D.upload_small(Ui)
        D.gemm(0, Nb, Nst, Nst, w, Nb, dUi.p, Nst, Tmp.col(0), Nb)
        D.gemm(0, Nb, Nst, Nst, hw, Nb, dUi.p, Nst, Tmp.col(Nst), Nb)
        D.copy(Nb, Nst, Tmp.col(0), w)
        D.copy(Nb, Nst, Tmp.col(Nst), hw)
        nq = 3 * Nst if it > 1 else 2 * Nst
        T = herm_u(D.gram(Q.p, HQ.p, nq, nq, Nb))
        G = herm_u(D64.gram(Q.p, Q.p, nq, nq, Nb))
        try:
            _, S = sla.eigh(T, G)
        except np.linalg.LinAlgError as e:
            raise PosDefException(-4, f"eigen(T, G): {e}")
        Um = np.asfortranarray(S[:, :Nst])
        dU = D.upload_small(Um)
        D.gemm(0, Nb, Nst, nq, Q.p, Nb, dU.p, nq, Tmp.col(0), Nb)
        D.gemm(0, Nb, Nst, nq, HQ.p, Nb, dU.p, nq, Tmp.col(Nst), Nb)
        if it > 1:
            dU23 = D.upload_small(Um[Nst:3 * Nst, :])
            D.gemm(0, Nb, Nst, 2 * Nst, w, Nb, dU23.p, 2 * Nst, Tmp.col(2 * Nst), Nb)
            D.gemm(0, Nb, Nst, 2 * Nst, hw, Nb, dU23.p, 2 * Nst, Tmp.col(3 * Nst), Nb)
            D.copy(Nb, Nst, Tmp.col(0), x)
            D.copy(Nb, Nst, Tmp.col(Nst), hx)
            U, Ui = _chol_upper_inv(D64.gram(Tmp.col(2 * Nst), Tmp.col(2 * Nst), Nst, Nst, Nb))
            dUi = D.upload_small(Ui)
            D.gemm(0, Nb, Nst, Nst, Tmp.col(2 * Nst), Nb, dUi.p, Nst, p, Nb)
            D.gemm(0, Nb, Nst, Nst, Tmp.col(3 * Nst), Nb, dUi.p, Nst, hp, Nb)
        else:
            D.copy(Nb, Nst, w, p)
            D.copy(Nb, Nst, hw, hp)
            D.copy(Nb, Nst, Tmp.col(0), x)
            D.copy(Nb, Nst, Tmp.col(Nst), hx)
    S = D.gram(x, hx, Nst, Nst, Nb)
    S = herm_u(0.5 * (S + S.conj().T))
    evals, Qm = np.linalg.eigh(S)
    dQ = D.upload_small(Qm)
    D.gemm(0, Nb, Nst, Nst, x, Nb, dQ.p, Nst, X.p, Nb)
    c.call("pw_synchronize")
    for m in (Q, HQ, Tmp):
        m.free()
    if lock_ws is not None:
        lock_ws.free()
    return evals, {"iter": it_done, "ncols": ncols, "locks": nlock}


def _gather(D, src_cols, cols, dst, Nb):
    for j, cj in enumerate(cols):
        D.copy(Nb, 1, src_cols(cj), _col(dst, j, Nb))


def _col(M, j, Nb):
    from ctypes import c_void_p
    return c_void_p(M.value + 16 * Nb * int(j)) if isinstance(M, c_void_p) else M.col(j)


class _Block:
    """Column block inside a larger device allocation (leading dimension rows), with the
    DeviceMatrix members the dense helpers use."""

    def __init__(self, p, rows, cols):
        self.p = ctypes.c_void_p(int(p))
        self.rows, self.cols = rows, cols

    def col(self, j):
        return ctypes.c_void_p(self.p.value + 16 * self.rows * int(j))


class _LockScratch:
    """Device memory of the lock step, allocated once per solver call instead of at every
    lock: the Xlock, [Xact W], [HXact HW] and staging blocks (Nb x (nl + 5 na) <= Nb x 5 Nst
    columns for nl + na = Nst) and one buffer for the small coefficient matrices.  Uploads
    into that buffer are ordered on the context stream after the products that read its
    previous contents."""

    def __init__(self, ctx, rows_max, Nst):
        self.ctx = ctx
        self.rows_max = rows_max
        self.mem = DeviceMatrix(ctx, rows_max, 5 * Nst)
        self.small = DeviceMatrix(ctx, 4 * Nst * Nst, 1)

    def blocks(self, Nb, nl, na):
        assert Nb <= self.rows_max and nl + 5 * na <= self.mem.cols
        b, s = self.mem.p.value, 16 * Nb
        return (_Block(b, Nb, nl), _Block(b + s * nl, Nb, 2 * na), _Block(b + s * (nl + 2 * na), Nb, 2 * na),
                _Block(b + s * (nl + 4 * na), Nb, na))

    def upload(self, M):
        a = np.asfortranarray(M, dtype=np.complex128)
        assert a.size <= self.small.rows
        self.ctx.call("pw_memcpy_h2d", self.small.p, ptr(a), a.nbytes)
        return _Block(self.small.p.value, a.shape[0], a.shape[1] if a.ndim > 1 else 1)

    def free(self):
        self.mem.free()
        self.small.free()


def _lobpcg_lock_step(D, Q, HQ, Tmp, Nb, Nst, ilock, iact, scratch=None):
    """Lock branch of diag_LOBPCG.jl:125-164 (the W projection with Ylock at :130-133 is
    a rebinding of a local in the reference and has no effect; it is not performed).
    scratch: a _LockScratch to take the device blocks from (else they are allocated)."""
    c = D.c
    nl, na = len(ilock), len(iact)
    x, w, hx = Q.col(0), Q.col(Nst), HQ.col(0)
    if scratch is None:
        own = [DeviceMatrix(c, Nb, nl), DeviceMatrix(c, Nb, 2 * na), DeviceMatrix(c, Nb, 2 * na),
               DeviceMatrix(c, Nb, na)]
        Xl, A, HA, T2 = own                 # Xlock, [Xact Wact], [HXact HW], staging
        up = D.upload_small
    else:
        own = []
        Xl, A, HA, T2 = scratch.blocks(Nb, nl, na)
        up = scratch.upload
    _gather(D, lambda j: Q.col(j), ilock, Xl, Nb)
    _gather(D, lambda j: Q.col(j), iact, A, Nb)
    _gather(D, lambda j: Q.col(Nst + j), iact, T2, Nb)
    D.copy(Nb, na, T2.p, A.col(na))

    def project(block_ptr, ncols):
        # block -= Xlock (Xlock' block)
        M = D.gram(Xl.p, block_ptr, nl, ncols, Nb)
        dM = up(M)
        D.gemm(0, Nb, ncols, nl, Xl.p, Nb, dM.p, nl, block_ptr, Nb, alpha=_MONE, beta=_ONE)

    project(A.p, 2 * na)
    D.opH(A.p, HA.p, 2 * na, Nb)
    project(HA.p, 2 * na)
    U, Ui = _chol_upper_inv(D.gram(A.col(na), A.col(na), na, na, Nb))
    dUi = up(Ui)
    D.gemm(0, Nb, na, na, A.col(na), Nb, dUi.p, na, T2.p, Nb)
    D.copy(Nb, na, T2.p, A.col(na))
    D.gemm(0, Nb, na, na, HA.col(na), Nb, dUi.p, na, T2.p, Nb)
    D.copy(Nb, na, T2.p, HA.col(na))
    T = herm_u(D.gram(A.p, HA.p, 2 * na, 2 * na, Nb))
    G = herm_u(D.gram(A.p, A.p, 2 * na, 2 * na, Nb))
    _, S = sla.eigh(T, G)
    dU = up(np.asfortranarray(S[:, :na]))
    # X = [Xlock Xact]
    D.copy(Nb, nl, Xl.p, x)
    D.gemm(0, Nb, na, 2 * na, A.p, Nb, dU.p, 2 * na, Q.col(nl), Nb)
    D.opH(x, Tmp.p, Nst, Nb)
    T = herm_u(D.gram(x, Tmp.p, Nst, Nst, Nb))
    _, ev = np.linalg.eigh(T)
    dE = up(ev)
    D.gemm(0, Nb, Nst, Nst, x, Nb, dE.p, Nst, Tmp.p, Nb)
    D.copy(Nb, Nst, Tmp.p, x)
    D.opH(x, hx, Nst, Nb)
    for m in own:
        m.free()
    return 2 * na + 2 * Nst


def _davidson_device(Ham, X, tol, NiterMax, Nstates_conv, mixed):
    """diag_davidson! through the C ABI (pw_diag_davidson_dev): the search space, its
    H-image, the reduced matrices and their generalised eigenproblem stay in HBM; the host
    receives the Ritz values once per iteration for the |dlambda| test."""
    Nst = X.cols
    ev = np.zeros(Nst)
    st = np.zeros(3, dtype=np.int32)
    Ham.ctx.call("pw_diag_davidson_dev", Ham.ik, Ham.ispin, Nst, X.p, X.rows, float(tol), int(NiterMax),
                 int(Nstates_conv), PW_MIXED if mixed else 0, ev.ctypes.data_as(ctypes.c_void_p),
                 st.ctypes.data_as(ctypes.c_void_p))
    return ev, {"iter": int(st[0]), "ncols": int(st[1])}


def diag_davidson_(Ham, X, tol=1e-5, NiterMax=100, Nstates_conv=0, mixed=None, dense=None):
    """diag_davidson! for the current Ham.ik (diag_davidson.jl:50-191).  Q = [X R] and
    HQ = [HX HR] give the reduced matrices of the 2Nst search space in two products.
    dense = "device" (default) runs the whole solver in the library
    (pw_diag_davidson_dev); "host" solves the reduced eigenproblem with LAPACK on the host
    as the reference does, the Ngw-sized products staying on the device."""
    mixed = Ham.mixed if mixed is None else mixed
    if (DENSE if dense is None else dense) == "device":
        if Nstates_conv == 0:
            Nstates_conv = X.cols
        if mixed:
            tol = max(tol, MIXED_TOL_FLOOR)
        return _davidson_device(Ham, X, tol, NiterMax, Nstates_conv, mixed)
    D = _Dense(Ham, mixed)
    Nb, Nst = X.rows, X.cols
    if Nstates_conv == 0:
        Nstates_conv = Nst
    c = Ham.ctx
    Q = DeviceMatrix(c, Nb, 2 * Nst)
    HQ = DeviceMatrix(c, Nb, 2 * Nst)
    Tmp = DeviceMatrix(c, Nb, 2 * Nst)
    x, r, hx, hr = Q.col(0), Q.col(Nst), HQ.col(0), HQ.col(Nst)
    D.copy(Nb, Nst, X.p, x)
    D.opH(x, hx, Nst, Nb)
    ncols = Nst

    def residual(evals):
        # R = X diag(evals) - HX ; res = column norms
        D.copy(Nb, Nst, hx, r)
        dE = D.upload_small(np.diag(evals.astype(complex)))
        D.gemm(0, Nb, Nst, Nst, x, Nb, dE.p, Nst, r, Nb, alpha=_ONE, beta=_MONE)
        RR = D.gram(r, r, Nst, Nst, Nb)
        return np.sqrt(np.maximum(np.real(np.diag(RR)), 0.0))

    evals = np.real(np.diag(D.gram(x, hx, Nst, Nst, Nb))).copy()
    res = residual(evals)
    evals_old = evals.copy()
    it_done = 0
    for it in range(1, NiterMax + 1):
        it_done = it
        rn = np.where(EPS < res, 1.0 / np.where(res > 0, res, 1.0), 1.0)
        dN = D.upload_small(np.diag(rn.astype(complex)))
        D.gemm(0, Nb, Nst, Nst, r, Nb, dN.p, Nst, Tmp.p, Nb)
        D.kprec(Tmp.p, r, Nst, Nb)
        D.opH(r, hr, Nst, Nb)
        ncols += Nst
        Hred = D.gram(Q.p, HQ.p, 2 * Nst, 2 * Nst, Nb)
        Sred = D.gram(Q.p, Q.p, 2 * Nst, 2 * Nst, Nb)
        if it > 1:
            Hred[:Nst, :Nst] = np.diag(evals.astype(complex))
        Hred[Nst:, :Nst] = Hred[:Nst, Nst:].conj().T
        Sred[:Nst, :Nst] = np.eye(Nst)
        Sred[Nst:, :Nst] = Sred[:Nst, Nst:].conj().T
        Hred = 0.5 * (Hred + Hred.conj().T)
        Sred = 0.5 * (Sred + Sred.conj().T)
        lam, Xr = sla.eigh(herm_u(Hred), herm_u(Sred))
        evals = lam[:Nst].copy()
        devals = np.abs(evals - evals_old)
        nconv = int(np.count_nonzero(devals < tol))
        dU = D.upload_small(np.asfortranarray(Xr[:, :Nst]))
        D.gemm(0, Nb, Nst, 2 * Nst, Q.p, Nb, dU.p, 2 * Nst, Tmp.col(0), Nb)
        D.gemm(0, Nb, Nst, 2 * Nst, HQ.p, Nb, dU.p, 2 * Nst, Tmp.col(Nst), Nb)
        D.copy(Nb, Nst, Tmp.col(0), x)
        D.copy(Nb, Nst, Tmp.col(Nst), hx)
        res = residual(evals)
        if nconv >= Nstates_conv:
            break
        evals_old = evals.copy()
    D.copy(Nb, Nst, x, X.p)
    c.call("pw_synchronize")
    for m in (Q, HQ, Tmp):
        m.free()
    return evals, {"iter": it_done, "ncols": ncols}


def diag_davidson_qe_(Ham, X, EBANDS_THR=1e-8, NiterMax=40, mixed=None):
    """diag_davidson_qe! for the current Ham.ik (diag_davidson_qe_v2.jl:40-216), QE's
    cegterg.  The search space P (Ngw x 2Nst) grows by the preconditioned, normalised
    corrections (H - e) P vc of the unconverged roots only; HP holds its H-image.  The
    reduced Hc, Sc (at most 2Nst square) and their Ritz vectors vc live on the host, as the
    reference diagonalises them with LAPACK.  op_S! of a GTH Hamiltonian is the identity
    (op_S.jl:82-89), so S P is P itself: the products with S P use P, which gives the same
    numbers as the reference's separate Spsi copy.  X holds the Ritz vectors on return.
    Returns (evals, {"iter", "ncols"})."""
    mixed = Ham.mixed if mixed is None else mixed
    D = _Dense(Ham, mixed)
    c = Ham.ctx
    Nb, Nvec = X.rows, X.cols
    Nvecx = 2 * Nvec
    P = DeviceMatrix(c, Nb, Nvecx)
    HP = DeviceMatrix(c, Nb, Nvecx)
    T = DeviceMatrix(c, Nb, Nvec)
    Hc = np.zeros((Nvecx, Nvecx), dtype=complex)
    Sc = np.zeros((Nvecx, Nvecx), dtype=complex)
    vc = np.zeros((Nvecx, Nvecx), dtype=complex)
    ew = np.zeros(Nvecx)
    is_conv = np.zeros(Nvec, dtype=bool)
    notcnv = nbase = Nvec
    D.copy(Nb, Nvec, X.p, P.p)
    D.opH(P.p, HP.p, Nvec, Nb)
    ncols = Nvec
    Hc[:Nvec, :Nvec] = D.gram(P.p, HP.p, Nvec, Nvec, Nb)
    Sc[:Nvec, :Nvec] = D.gram(P.p, P.p, Nvec, Nvec, Nb)
    w, v = sla.eigh(herm_u(Hc[:Nvec, :Nvec]), herm_u(Sc[:Nvec, :Nvec]), lower=False)
    ew[:Nvec] = w
    vc[:Nvec, :Nvec] = v
    evals = w.copy()
    kter = 1
    dav_iter = 0
    small = []
    while kter <= NiterMax or notcnv == 0:
        dav_iter = kter
        k = 0
        for ist in range(Nvec):
            if not is_conv[ist]:
                if k != ist:
                    vc[:, k] = vc[:, ist]
                ew[nbase + k] = evals[ist]
                k += 1
        # T = P vc (-e) + HP vc over the unconverged roots, then P_new = K T
        V = vc[:nbase, :notcnv]
        dVe = D.upload_small(V * (-ew[nbase:nbase + notcnv])[None, :])
        dV = D.upload_small(V)
        small += [dVe, dV]
        D.gemm(0, Nb, notcnv, nbase, P.p, Nb, dVe.p, nbase, T.p, Nb)
        D.gemm(0, Nb, notcnv, nbase, HP.p, Nb, dV.p, nbase, T.p, Nb, beta=_ONE)
        new = P.col(nbase)
        D.kprec(T.p, new, notcnv, Nb)
        nrm = np.real(np.diag(D.gram(new, new, notcnv, notcnv, Nb)))
        ew[:notcnv] = nrm
        inv = np.ascontiguousarray(1.0 / np.sqrt(nrm))
        c.call("pw_scale_cols_dev", Nb, notcnv, ptr(inv), new, Nb)
        D.opH(new, HP.col(nbase), notcnv, Nb)
        ncols += notcnv
        m = nbase + notcnv
        Hc[:m, nbase:m] = D.gram(P.p, HP.col(nbase), m, notcnv, Nb)
        Sc[:m, nbase:m] = D.gram(P.p, new, m, notcnv, Nb)
        nbase = m
        Hc[:nbase, :nbase] = herm_u(Hc[:nbase, :nbase])
        Sc[:nbase, :nbase] = herm_u(Sc[:nbase, :nbase])
        w, v = sla.eigh(Hc[:nbase, :nbase], Sc[:nbase, :nbase])
        ew[:nbase] = w
        vc[:nbase, :nbase] = v
        vc[:, Nvec:] = 0.0
        is_conv = np.abs(ew[:Nvec] - evals) < EBANDS_THR
        notcnv = int(np.count_nonzero(~is_conv))
        evals = ew[:Nvec].copy()
        if notcnv == 0 or nbase + notcnv > Nvecx or dav_iter == NiterMax:
            dU = D.upload_small(vc[:nbase, :Nvec])
            small.append(dU)
            D.gemm(0, Nb, Nvec, nbase, P.p, Nb, dU.p, nbase, X.p, Nb)
            if notcnv == 0 or dav_iter == NiterMax:
                break
            D.gemm(0, Nb, Nvec, nbase, HP.p, Nb, dU.p, nbase, T.p, Nb)
            D.copy(Nb, Nvec, X.p, P.p)
            D.copy(Nb, Nvec, T.p, HP.p)
            nbase = Nvec
            Hc[:, :nbase] = 0.0
            Sc[:, :nbase] = 0.0
            vc[:, :nbase] = 0.0
            Hc[np.arange(nbase), np.arange(nbase)] = evals
            Sc[np.arange(nbase), np.arange(nbase)] = 1.0
            vc[np.arange(nbase), np.arange(nbase)] = 1.0
        if len(small) > 8:
            c.call("pw_synchronize")
            for s in small:
                s.free()
            small = []
        kter += 1
    c.call("pw_synchronize")
    for m_ in [P, HP, T] + small:
        m_.free()
    return evals, {"iter": dav_iter, "ncols": ncols}


def diag_davidson_qe(Ham, psiks, tol=1e-5, NiterMax=100):
    """diag_davidson_qe!(Ham, psiks::BlochWavefunc) (diag_davidson_qe_v2.jl:6-37): per
    k-point with EBANDS_THR = tol; evals Nst x Nkpt and the per-k statistics."""
    Nk = Ham.pw.Nkpt
    evals = np.zeros((psiks[0].cols, Nk))
    stats = []
    for ik in range(Nk):
        Ham.ik, Ham.ispin = ik, 0
        evals[:, ik], st = diag_davidson_qe_(Ham, psiks[ik], EBANDS_THR=tol, NiterMax=NiterMax)
        stats.append(st)
    return evals, stats


def diag_LOBPCG(Ham, psiks, tol=1e-5, NiterMax=100, Nstates_conv=0):
    """diag_LOBPCG!(Ham, psiks::BlochWavefunc) (diag_LOBPCG.jl:1-28): evals Nst x Nkpt."""
    return _over_k(Ham, psiks, diag_LOBPCG_, tol, NiterMax, Nstates_conv)


def diag_davidson(Ham, psiks, tol=1e-5, NiterMax=100, Nstates_conv=0):
    """diag_davidson!(Ham, psiks::BlochWavefunc) (diag_davidson.jl:12-39)."""
    return _over_k(Ham, psiks, diag_davidson_, tol, NiterMax, Nstates_conv)


def _over_k(Ham, psiks, fn, tol, NiterMax, Nstates_conv):
    Nk = Ham.pw.Nkpt
    evals = np.zeros((psiks[0].cols, Nk))
    stats = []
    for ik in range(Nk):
        Ham.ik, Ham.ispin = ik, 0
        evals[:, ik], st = fn(Ham, psiks[ik], tol=tol, NiterMax=NiterMax, Nstates_conv=Nstates_conv)
        stats.append(st)
    return evals, stats


def ortho_sqrt_(Ham, X):
    """ortho_sqrt! (ortho_sqrt.jl:6-9) on a DeviceMatrix: X <- X (X'X)^(-1/2)."""
    D = _Dense(Ham, False)
    Nb, Nst = X.rows, X.cols
    O = herm_u(D.gram(X.p, X.p, Nst, Nst, Nb))
    wv, U = np.linalg.eigh(O)
    M = U @ np.diag(1.0 / np.sqrt(wv)) @ U.conj().T
    dM = D.upload_small(M)
    T = DeviceMatrix(Ham.ctx, Nb, Nst)
    D.gemm(0, Nb, Nst, Nst, X.p, Nb, dM.p, Nst, T.p, Nb)
    D.copy(Nb, Nst, T.p, X.p)
    Ham.ctx.call("pw_synchronize")
    T.free()
    return X


def rand_BlochWavefunc(Ham, seed=1234):
    """rand_BlochWavefunc (gen_wavefunc.jl:37-92) with the project RNG (numpy PCG64):
    U[0,1) + i U[0,1) per coefficient, Loewdin-orthonormalised on the device."""
    rng = np.random.default_rng(seed)
    out = []
    for ik in range(Ham.pw.Nkpt):
        a = rng.random((Ham.pw.Ngw[ik], Ham.electrons.Nstates)) + 1j * rng.random((Ham.pw.Ngw[ik], Ham.electrons.Nstates))
        X = DeviceMatrix.from_host(Ham.ctx, a)
        Ham.ik = ik
        out.append(ortho_sqrt_(Ham, X))
    return out
