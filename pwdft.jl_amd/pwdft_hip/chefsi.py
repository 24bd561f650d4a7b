"""Chebyshev-filtered subspace iteration of PWDFT.jl on the device: diag_CheFSI!
(CheFSI.jl:1-31), chebyfilt (:36-62) and get_ub_lb_lanczos (:65-130).

The filter is op_H on the whole block plus three-term real linear combinations of
Ngw x Nst blocks (pw_vec_lincomb_dev over the interleaved storage); the Lanczos bound
estimate applies op_H to one vector per step and keeps the basis V, its H-image and the
reorthogonalisation products in device memory (only h, s and |f| come back per step to fill
the small T).  diag_CheFSI! binds `ub, lb = get_ub_lb_lanczos(...)` to a function that
returns (lb, ub), so the filter receives the two bounds exchanged; this is kept, as the
scaled recurrence (sigma1 = -1/2 either way) then only flips the block's overall sign."""
import ctypes

import numpy as np

from ._lib import DeviceMatrix, PW_MIXED, PW_TERM_ALL, ptr
from .solvers import _Dense, herm_u, ortho_sqrt_

_ONE = np.array([1.0, 0.0])
_ZERO = np.array([0.0, 0.0])
_MONE = np.array([-1.0, 0.0])


def _opH(Ham, src, dst, ncols, rows):
    Ham.ctx.call("pw_op_H_dev", Ham.ik, Ham.ispin, ncols, src, rows, dst, rows,
                 PW_TERM_ALL | (PW_MIXED if Ham.mixed else 0))


def _lin(ctx, n, y, ay, terms):
    """y = ay y + sum c v on raw device pointers, n doubles each."""
    c = np.ascontiguousarray([float(a) for a, _ in terms], dtype=float)
    x = (ctypes.c_void_p * len(terms))(*[v.value for _, v in terms]) if terms else None
    ctx.call("pw_vec_lincomb_dev", n, len(terms), ptr(c) if terms else None, x, float(ay), y)


def chebyfilt(Ham, X, degree, lb, ub):
    """chebyfilt (CheFSI.jl:36-62) of the DeviceMatrix X in place: X <- Y_degree of
    Y_1 = (H X - c X) sigma1 / e,  Y_i = (H Y_{i-1} - c Y_{i-1}) 2 sigma_i / e - sigma sigma_i Y_{i-2}."""
    ctx = Ham.ctx
    Nb, Nst = X.rows, X.cols
    n = 2 * Nb * Nst
    ee = (ub - lb) / 2
    c = (ub + lb) / 2
    sigma = ee / (lb - ub)
    sigma1 = sigma
    HY = DeviceMatrix(ctx, Nb, Nst)
    work = [DeviceMatrix(ctx, Nb, Nst) for _ in range(3)]
    prev, cur = X.p, work[0].p
    _opH(Ham, prev, HY.p, Nst, Nb)
    _lin(ctx, n, cur, 0.0, [(sigma1 / ee, HY.p), (-c * sigma1 / ee, prev)])
    for _ in range(2, degree + 1):
        sigma2 = 1 / (2 / sigma1 - sigma)
        nxt = next(w.p for w in work if w.p.value not in (cur.value, prev.value))
        _opH(Ham, cur, HY.p, Nst, Nb)
        _lin(ctx, n, nxt, 0.0, [(2 * sigma2 / ee, HY.p), (-2 * c * sigma2 / ee, cur), (-sigma * sigma2, prev)])
        prev, cur = cur, nxt
        sigma = sigma2
    ctx.call("pw_copy_cols_dev", Nb, Nst, cur, Nb, X.p, Nb)
    ctx.call("pw_synchronize")
    for b in work + [HY]:
        b.free()
    return X


def get_ub_lb_lanczos(Ham, nlancz, v0=None, rng=None):
    """get_ub_lb_lanczos (CheFSI.jl:65-130) for k-point Ham.ik: returns (lb, ub) with lb the
    nlancz/2-th eigenvalue of the Lanczos T (sorted by real part, as Julia's eigvals of a
    general matrix) and ub = |T 1/sqrt(n)| + |f|.  v0: start vector (default: complex
    normal from `rng`, the reference's randn)."""
    ctx = Ham.ctx
    Nb = Ham.pw.Ngw[Ham.ik]
    if v0 is None:
        rng = np.random.default_rng() if rng is None else rng
        v0 = rng.standard_normal(Nb) + 1j * rng.standard_normal(Nb)
    D = _Dense(Ham, False)
    V = DeviceMatrix(ctx, Nb, nlancz)
    HV = DeviceMatrix(ctx, Nb, nlancz)
    f = DeviceMatrix(ctx, Nb, 1)
    hs = DeviceMatrix(ctx, 2 * nlancz, 1)          # h (first nlancz) and s
    T = np.zeros((nlancz, nlancz))
    hb = np.empty(2 * nlancz, dtype=complex)
    v0 = np.asarray(v0, dtype=complex)
    v1 = np.ascontiguousarray(v0 / np.linalg.norm(v0))
    ctx.call("pw_memcpy_h2d", V.p, ptr(v1), 16 * Nb)
    col = lambda M, j: ctypes.c_void_p(M.p.value + 16 * Nb * j)
    sp = ctypes.c_void_p(hs.p.value + 16 * nlancz)
    # first step: h = real(v' H v), f = Hv - v h, s = v' f, f -= v s
    _opH(Ham, V.p, HV.p, 1, Nb)
    h0 = float(np.real(D.gram(V.p, HV.p, 1, 1, Nb)[0, 0]))
    T[0, 0] = h0
    _lin(ctx, 2 * Nb, f.p, 0.0, [(1.0, HV.p), (-h0, V.p)])
    D.gemm(1, 1, 1, Nb, V.p, Nb, f.p, Nb, sp, 1)
    D.gemm(0, Nb, 1, 1, V.p, Nb, sp, 1, f.p, Nb, alpha=_MONE, beta=_ONE)
    for j in range(1, nlancz):
        beta = float(np.sqrt(np.real(D.gram(f.p, f.p, 1, 1, Nb)[0, 0])))
        T[j, j - 1] = beta
        _lin(ctx, 2 * Nb, col(V, j), 0.0, [(1.0 / beta, f.p)])
        _opH(Ham, col(V, j), col(HV, j), 1, Nb)
        m = j + 1
        D.gemm(1, m, 1, Nb, V.p, Nb, col(HV, j), Nb, hs.p, m)                 # h = V' H v_j
        ctx.call("pw_copy_cols_dev", Nb, 1, col(HV, j), Nb, f.p, Nb)
        D.gemm(0, Nb, 1, m, V.p, Nb, hs.p, m, f.p, Nb, alpha=_MONE, beta=_ONE)  # f = H v_j - V h
        D.gemm(1, m, 1, Nb, V.p, Nb, f.p, Nb, sp, m)                          # s = V' f
        D.gemm(0, Nb, 1, m, V.p, Nb, sp, m, f.p, Nb, alpha=_MONE, beta=_ONE)   # f -= V s
        ctx.call("pw_memcpy_d2h", ptr(hb), hs.p, 16 * 2 * nlancz)
        T[:m, j] = np.real(hb[:m] + hb[nlancz:nlancz + m])
    fn = float(np.sqrt(np.real(D.gram(f.p, f.p, 1, 1, Nb)[0, 0])))
    for b in (V, HV, f, hs):
        b.free()
    ev = np.linalg.eigvals(T)
    ev = ev[np.lexsort((ev.imag, ev.real))]
    lb = float(ev[nlancz // 2 - 1].real)
    ub = float(np.linalg.norm(T @ (np.ones(nlancz) / np.sqrt(nlancz))) + fn)
    return lb, ub


def diag_CheFSI_(Ham, psiks, cheby_degree=8, v0s=None, rng=None):
    """diag_CheFSI! (CheFSI.jl:1-31) over the k-points: per k, Lanczos bounds with 2 Nstates
    steps, the filter, ortho_sqrt!, and eigvals(Hermitian(psi' H psi)).  psiks (list of
    DeviceMatrix) are replaced by the filtered, orthonormalised blocks.  Returns
    (evals Nstates x Nkpt, {"ncols": op_H columns applied})."""
    Nst = Ham.electrons.Nstates
    nk = len(psiks)
    evals = np.zeros((Nst, nk))
    ncols = 0
    for ik in range(nk):
        Ham.ik, Ham.ispin = ik, 0
        X = psiks[ik]
        ub, lb = get_ub_lb_lanczos(Ham, 2 * Nst, None if v0s is None else v0s[ik], rng=rng)
        chebyfilt(Ham, X, cheby_degree, lb, ub)
        ortho_sqrt_(Ham, X)
        D = _Dense(Ham, False)
        HX = DeviceMatrix(Ham.ctx, X.rows, X.cols)
        _opH(Ham, X.p, HX.p, X.cols, X.rows)
        Hr = D.gram(X.p, HX.p, X.cols, X.cols, X.rows)
        HX.free()
        evals[:, ik] = np.linalg.eigvalsh(herm_u(Hr))
        ncols += 2 * Nst + cheby_degree * Nst + Nst
    return evals, {"ncols": ncols}
