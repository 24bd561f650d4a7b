"""All-band preconditioned conjugate-gradient eigensolver of PWDFT.jl on the device:
diag_Emin_PCG! (diag_Emin_PCG.jl:56-195) with calc_grad_evals (calc_grad.jl:60-82).

Every Ngw x Nst block (X, the gradients, their preconditioned images, the directions, the
trial point) stays in device memory.  The real inner products Re sum conj(a) b that set the
conjugate-direction factor and the line step are ordered device reductions over the
interleaved storage (pw_vec_dots_dev); the block updates are pw_vec_lincomb_dev; only the
Nst x Nst Rayleigh quotient matrix comes back for its eigenvalues, as the reference also
diagonalises it on the host."""
import ctypes

import numpy as np

from ._lib import DeviceMatrix, ptr
from .solvers import _Dense, herm_u, ortho_sqrt_
from .chefsi import _lin, _opH

_ONE = np.array([1.0, 0.0])
_MONE = np.array([-1.0, 0.0])


def _rdots(ctx, n, pairs):
    """[Re sum conj(a) b for (a, b) in pairs] over n doubles of interleaved storage."""
    out = np.zeros(len(pairs))
    a = (ctypes.c_void_p * len(pairs))(*[x.p.value for x, _ in pairs])
    b = (ctypes.c_void_p * len(pairs))(*[y.p.value for _, y in pairs])
    ctx.call("pw_vec_dots_dev", n, len(pairs), a, b, ptr(out))
    return out


def calc_grad_evals_(Ham, X, G, S):
    """G <- H X - X (X' H X) (calc_grad.jl:60-82); S: device Nst x Nst scratch."""
    D = _Dense(Ham, False)
    Nb, Nst = X.rows, X.cols
    _opH(Ham, X.p, G.p, Nst, Nb)
    D.gemm(1, Nst, Nst, Nb, X.p, Nb, G.p, Nb, S.p, Nst)
    D.gemm(0, Nb, Nst, Nst, X.p, Nb, S.p, Nst, G.p, Nb, alpha=_MONE, beta=_ONE)


def _rayleigh(Ham, X, HX):
    D = _Dense(Ham, False)
    _opH(Ham, X.p, HX.p, X.cols, X.rows)
    return herm_u(D.gram(X.p, HX.p, X.cols, X.cols, X.rows))


def diag_Emin_PCG_(Ham, X, tol=1e-5, NiterMax=100, Nstates_conv=0, tol_ebands=1e-4, alpha_t=3e-5, I_CG_BETA=2):
    """diag_Emin_PCG!(Ham, X) for k-point Ham.ik on a DeviceMatrix X (orthonormal on
    entry; the Ritz vectors on return).  Returns (evals, {"iterations", "ncols"})."""
    ctx = Ham.ctx
    Nb, Nst = X.rows, X.cols
    n = 2 * Nb * Nst
    if Nstates_conv == 0:
        Nstates_conv = Nst
    new = lambda: DeviceMatrix(ctx, Nb, Nst)
    g, g_old, Kg, Kg_old, d, d_old, Xc, gt = (new() for _ in range(8))
    S = DeviceMatrix(ctx, Nst, Nst)
    ncols = Nst
    evals = np.linalg.eigvalsh(_rayleigh(Ham, X, gt))
    evals_old = evals.copy()
    Ebands_old = 0.0
    beta = 0.0
    it_done = 0
    for it in range(1, NiterMax + 1):
        it_done = it
        calc_grad_evals_(Ham, X, g, S)
        ctx.call("pw_kprec_dev", Ham.ik, Nst, g.p, Nb, Kg.p, Nb)
        ncols += Nst
        if it != 1:
            if I_CG_BETA == 1:
                a, b = _rdots(ctx, n, [(g, Kg), (g_old, Kg_old)])
                beta = a / b
            elif I_CG_BETA == 2:
                a, a2, b = _rdots(ctx, n, [(g, Kg), (g_old, Kg), (g_old, Kg_old)])
                beta = (a - a2) / b
            elif I_CG_BETA == 3:
                # the reference's `d` here is the previous direction, held in d_old
                a, a2, c1, c2 = _rdots(ctx, n, [(g, Kg), (g_old, Kg), (g, d_old), (g_old, d_old)])
                beta = (a - a2) / (c1 - c2)
            else:
                a, c1, c2 = _rdots(ctx, n, [(g, Kg), (d_old, g), (d_old, g_old)])
                beta = a / (c1 - c2)
        if beta < 0.0:
            beta = 0.0
        # d = -Kg + beta d_old (d_old is still the previous direction here)
        _lin(ctx, n, d.p, 0.0, [(-1.0, Kg.p)] + ([(beta, d_old.p)] if it != 1 else []))
        _lin(ctx, n, Xc.p, 0.0, [(1.0, X.p), (alpha_t, d.p)])
        ortho_sqrt_(Ham, Xc)
        calc_grad_evals_(Ham, Xc, gt, S)
        ncols += Nst
        gd, gtd = _rdots(ctx, n, [(g, d), (gt, d)])
        denum = gd - gtd
        alpha = abs(alpha_t * gd / denum) if denum != 0.0 else 0.0
        _lin(ctx, n, X.p, 1.0, [(alpha, d.p)])
        ortho_sqrt_(Ham, X)
        evals = np.linalg.eigvalsh(_rayleigh(Ham, X, gt))
        ncols += Nst
        Ebands = float(np.sum(evals))
        devals = np.abs(evals - evals_old)
        evals_old = evals.copy()
        nconv = int(np.sum(devals < tol))
        if nconv >= Nstates_conv or abs(Ebands - Ebands_old) <= tol_ebands * Nst:
            break
        g, g_old = g_old, g
        Kg, Kg_old = Kg_old, Kg
        d, d_old = d_old, d
        Ebands_old = Ebands
    ortho_sqrt_(Ham, X)
    Hr = _rayleigh(Ham, X, gt)
    ncols += Nst
    evals, evecs = np.linalg.eigh(Hr)
    D = _Dense(Ham, False)
    U = D.upload_small(evecs)
    D.gemm(0, Nb, Nst, Nst, X.p, Nb, U.p, Nst, Xc.p, Nb)
    D.copy(Nb, Nst, Xc.p, X.p)
    ctx.call("pw_synchronize")
    for m in (g, g_old, Kg, Kg_old, d, d_old, Xc, gt, S, U):
        m.free()
    return evals, {"iterations": it_done, "ncols": ncols}
