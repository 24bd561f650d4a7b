"""Direct minimisation of the Kohn-Sham energy on the device: KS_solve_Emin_PCG!
(KS_solve_Emin_PCG.jl:27-270) with calc_grad (calc_grad.jl:1-57).

Each iteration takes the preconditioned conjugate direction of the energy gradient per
k-point, a trial step alpha_t to the orthonormalised psi_c, the density and potential of
psi_c, the gradient there, and the line-minimising step; then the density, potential and
energies of the new psi.  All blocks, densities and potentials stay in device memory; the
host gets the Nst x Nst products that set the occupation term and the scalars."""
import ctypes

import numpy as np

from ._lib import DeviceMatrix, DeviceVector, ptr
from .chefsi import _lin, _opH
from .hamiltonian import calc_rhoe_
from .pcg import _rdots
from .scf import calc_energies
from .solvers import _Dense, diag_LOBPCG_, herm_u, ortho_sqrt_, rand_BlochWavefunc

_ONE = np.array([1.0, 0.0])
_MONE = np.array([-1.0, 0.0])


def calc_grad_(Ham, psi, G, S, Focc_k, wk):
    """calc_grad (calc_grad.jl:1-57) for k-point Ham.ik into G:
    G[:, i] = wk Focc_i (H psi - psi (psi' H psi))[:, i], plus psi Q with
    Q = (Hr F - F Hr) / 2, Hr = psi' H psi, F = diag(Focc), unless all Focc are 2 or all 1."""
    D = _Dense(Ham, False)
    Nb, Nst = psi.rows, psi.cols
    _opH(Ham, psi.p, G.p, Nst, Nb)
    D.gemm(1, Nst, Nst, Nb, psi.p, Nb, G.p, Nb, S.p, Nst)                      # Hr = psi' H psi
    constant = bool(np.all(Focc_k == 2.0) or np.all(Focc_k == 1.0))
    Hr = None
    if not constant:
        Hr = np.empty((Nst, Nst), dtype=complex, order="F")
        Ham.ctx.call("pw_memcpy_d2h", ptr(Hr), S.p, 16 * Nst * Nst)
    D.gemm(0, Nb, Nst, Nst, psi.p, Nb, S.p, Nst, G.p, Nb, alpha=_MONE, beta=_ONE)
    w = np.ascontiguousarray(wk * Focc_k, dtype=float)
    Ham.ctx.call("pw_scale_cols_dev", Nb, Nst, ptr(w), G.p, Nb)
    if not constant:
        F = np.diag(Focc_k)
        Q = D.upload_small(0.5 * (Hr @ F - F @ Hr))
        D.gemm(0, Nb, Nst, Nst, psi.p, Nb, Q.p, Nst, G.p, Nb, alpha=_ONE, beta=_ONE)
        Q.free()


def KS_solve_Emin_PCG_(Ham, psiks=None, skip_initial_diag=False, alpha_t=3e-5, NiterMax=200, i_cg_beta=2,
                       etot_conv_thr=1e-6, alpha_max=2.1, restrict_linmin=False, seed=1234, verbose=False):
    """KS_solve_Emin_PCG! (KS_solve_Emin_PCG.jl:27-270), Nspin = 1, starting from the
    Gaussian density guess; returns the same summary as KS_solve_SCF_."""
    import time
    pw, el, ctx = Ham.pw, Ham.electrons, Ham.ctx
    if psiks is None:
        psiks = rand_BlochWavefunc(Ham, seed=seed)
    nk = len(psiks)
    Focc = el.Focc
    rho = DeviceVector.from_host(ctx, Ham.guess_rhoe())
    Ham.update(rho)
    if not skip_initial_diag:
        for ik in range(nk):
            Ham.ik, Ham.ispin = ik, 0
            diag_LOBPCG_(Ham, psiks[ik], NiterMax=10)
    shape = [(p.rows, p.cols) for p in psiks]
    blocks = lambda: [DeviceMatrix(ctx, r, c) for r, c in shape]
    g, g_old, Kg, Kg_old, d, d_old, psic, gt = (blocks() for _ in range(8))
    S = [DeviceMatrix(ctx, c, c) for _, c in shape]
    beta = np.zeros(nk)
    alpha = np.zeros(nk)
    e = calc_energies(Ham, psiks, rho, Focc, E_local=Ham.update(rho))
    Etot_old = 0.0
    nconv = 0
    history = []
    t0 = time.perf_counter()
    for it in range(1, NiterMax + 1):
        for i in range(nk):
            Ham.ik, Ham.ispin = i, 0
            n = 2 * shape[i][0] * shape[i][1]
            calc_grad_(Ham, psiks[i], g[i], S[i], Focc[:, i], pw.wk[i])
            ctx.call("pw_kprec_dev", i, shape[i][1], g[i].p, shape[i][0], Kg[i].p, shape[i][0])
            if it != 1:
                if i_cg_beta == 1:
                    a, b = _rdots(ctx, n, [(g[i], Kg[i]), (g_old[i], Kg_old[i])])
                    beta[i] = a / b
                    if abs(beta[i] - 1.0) < 0.2:
                        beta[i] = 0.0
                elif i_cg_beta == 2:
                    a, a2, b = _rdots(ctx, n, [(g[i], Kg[i]), (g_old[i], Kg[i]), (g_old[i], Kg_old[i])])
                    beta[i] = (a - a2) / b
                elif i_cg_beta == 3:
                    # the reference's `d` here still holds the previous direction (d_old)
                    a, a2, c1, c2 = _rdots(ctx, n, [(g[i], Kg[i]), (g_old[i], Kg[i]), (g[i], d_old[i]),
                                                    (g_old[i], d_old[i])])
                    beta[i] = (a - a2) / (c1 - c2)
                else:
                    a, c1, c2 = _rdots(ctx, n, [(g[i], Kg[i]), (d_old[i], g[i]), (d_old[i], g_old[i])])
                    beta[i] = a / (c1 - c2)
                    if abs(beta[i] - 1.0) < 0.2:
                        beta[i] = 0.0
            if beta[i] < 0.0 or np.isnan(beta[i]):
                beta[i] = 0.0
            _lin(ctx, n, d[i].p, 0.0, [(-1.0, Kg[i].p)] + ([(beta[i], d_old[i].p)] if it != 1 else []))
            _lin(ctx, n, psic[i].p, 0.0, [(1.0, psiks[i].p), (alpha_t, d[i].p)])
            ortho_sqrt_(Ham, psic[i])
        calc_rhoe_(Ham, psic, rho, Focc=Focc)
        Ham.update(rho)
        for i in range(nk):
            Ham.ik, Ham.ispin = i, 0
            n = 2 * shape[i][0] * shape[i][1]
            calc_grad_(Ham, psic[i], gt[i], S[i], Focc[:, i], pw.wk[i])
            gd, gtd = _rdots(ctx, n, [(g[i], d[i]), (gt[i], d[i])])
            denum = gd - gtd
            alpha[i] = abs(alpha_t * gd / denum) if denum != 0.0 else 0.0
            if restrict_linmin and alpha[i] > alpha_max:
                alpha[i] = alpha_max
            _lin(ctx, n, psiks[i].p, 1.0, [(alpha[i], d[i].p)])
            ortho_sqrt_(Ham, psiks[i])
        calc_rhoe_(Ham, psiks, rho, Focc=Focc)
        e = calc_energies(Ham, psiks, rho, Focc, E_local=Ham.update(rho))
        Etot = e.total()
        diffE = Etot_old - Etot
        history.append(Etot)
        if verbose:
            print(f"Emin_PCG step {it:8d} = {Etot:18.10f} {diffE:10.7e}", flush=True)
        nconv = nconv + 1 if abs(diffE) < etot_conv_thr else 0
        if nconv >= 2:
            break
        g, g_old = g_old, g
        d, d_old = d_old, d
        Kg, Kg_old = Kg_old, Kg
        Etot_old = Etot
    evals = np.zeros((el.Nstates, nk))
    for i in range(nk):
        Ham.ik, Ham.ispin = i, 0
        X = psiks[i]
        ortho_sqrt_(Ham, X)
        D = _Dense(Ham, False)
        _opH(Ham, X.p, gt[i].p, X.cols, X.rows)
        Hr = herm_u(D.gram(X.p, gt[i].p, X.cols, X.cols, X.rows))
        ev, U = np.linalg.eigh(Hr)
        evals[:, i] = ev
        dU = D.upload_small(U)
        D.gemm(0, X.rows, X.cols, X.cols, X.p, X.rows, dU.p, X.cols, psic[i].p, X.rows)
        D.copy(X.rows, X.cols, psic[i].p, X.p)
        dU.free()
    ctx.call("pw_synchronize")
    for lst in (g, g_old, Kg, Kg_old, d, d_old, psic, gt, S):
        for m in lst:
            m.free()
    el.ebands = evals
    Ham.energies = e
    return {"energies": e.as_dict(), "evals": evals, "psiks": psiks, "rho": rho.download(), "history": history,
            "seconds": time.perf_counter() - t0, "iterations": len(history)}
