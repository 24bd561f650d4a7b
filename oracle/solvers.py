"""Oracle: eigensolvers restated from diag_LOBPCG.jl, diag_davidson.jl, ortho_sqrt.jl,
gen_wavefunc.jl.  Test infrastructure only.

The solvers take `opH(X)` and `prec(R)` callables so the same restatement runs on the
oracle Hamiltonian (ops.op_H / ops.Kprec)."""
import numpy as np
import scipy.linalg as sla

EPS = np.finfo(float).eps


def herm_u(A):
    """Julia's Hermitian(A) (uplo = :U): upper triangle mirrored, real diagonal."""
    U = np.triu(A, 1)
    return U + U.conj().T + np.diag(np.real(np.diag(A)))


def ortho_sqrt(psi):
    """ortho_sqrt.jl:1-4: psi * inv(sqrt(psi'psi)) (Loewdin)."""
    O = herm_u(psi.conj().T @ psi)
    w, U = np.linalg.eigh(O)
    return psi @ (U @ np.diag(1.0 / np.sqrt(w)) @ U.conj().T)


def rand_wavefunc(rng, Nbasis, Nstates):
    """gen_wavefunc.jl:90-92 with the project RNG: ortho_sqrt(U[0,1) + i U[0,1))."""
    X = rng.random((Nbasis, Nstates)) + 1j * rng.random((Nbasis, Nstates))
    return ortho_sqrt(X)


def _rdiv_chol(W, C):
    """W / cholesky(Hermitian(C)).U  (diag_LOBPCG.jl:172-174) -> (W U^{-1}, U)."""
    U = sla.cholesky(herm_u(C), lower=False)
    return sla.solve_triangular(U, W.T, trans="T", lower=False).T, U


def _rdiv_U(W, U):
    return sla.solve_triangular(U, W.T, trans="T", lower=False).T


def diag_LOBPCG(opH, prec, X, tol=1e-5, NiterMax=100, Nstates_conv=0, stats=None):
    """diag_LOBPCG.jl:51-225 for one k-point.  X (Nbasis x Nstates, orthonormal) is
    overwritten in place; returns (evals, iterations, number of opH columns applied).
    stats (a dict, optional) receives "locks": the iterations that took the lock branch."""
    Nbasis, Nstates = X.shape
    if Nstates_conv == 0:
        Nstates_conv = Nstates
    ncols = Nstates
    HX = opH(X)
    P = np.zeros_like(X)
    HP = np.zeros_like(X)
    evals_old = np.zeros(Nstates)
    SMALL = 10.0 * EPS
    it_done = 0
    for it in range(1, NiterMax + 1):
        it_done = it
        XHX = herm_u(X.conj().T @ HX)
        evals = np.linalg.eigvalsh(XHX)
        devals = np.abs(evals - evals_old)
        evals_old = evals.copy()
        R = HX - X @ XHX
        nconv = int(np.count_nonzero(devals < tol))
        ilock = np.nonzero(devals <= SMALL)[0]
        iact = np.nonzero(devals > SMALL)[0]
        Nlock, Nact = len(ilock), len(iact)
        if nconv >= Nstates_conv:
            break
        W = prec(R)
        if Nlock > 0:
            if stats is not None:
                stats["locks"] = stats.get("locks", 0) + 1
            # lock branch, diag_LOBPCG.jl:125-164 (the `ww = ww - ...` update at :130-133
            # rebinds a local in the reference and changes nothing; it is omitted)
            Xlock = X[:, ilock].copy()
            X[:, iact] = X[:, iact] - Xlock @ (Xlock.conj().T @ X[:, iact])
            W[:, iact] = W[:, iact] - Xlock @ (Xlock.conj().T @ W[:, iact])
            HX[:, iact] = opH(X[:, iact])
            HW = opH(W[:, iact])
            ncols += 2 * Nact
            HX[:, iact] = HX[:, iact] - Xlock @ (Xlock.conj().T @ HX[:, iact])
            HW = HW - Xlock @ (Xlock.conj().T @ HW)
            Wa, U = _rdiv_chol(W[:, iact], W[:, iact].conj().T @ W[:, iact])
            W[:, iact] = Wa
            HW = _rdiv_U(HW, U)
            Q = np.hstack([X[:, iact], W[:, iact]])
            HQ = np.hstack([HX[:, iact], HW])
            T = herm_u(Q.conj().T @ HQ)
            G = herm_u(Q.conj().T @ Q)
            sd, S = sla.eigh(T, G, lower=False)
            Xact = Q @ S[:, :Nact]
            X[:, :] = np.hstack([Xlock, Xact])
            T = herm_u(X.conj().T @ opH(X))
            ncols += Nstates
            evT, evecT = np.linalg.eigh(T)
            X[:, :] = X @ evecT
            HX = opH(X)
            ncols += Nstates
        else:
            HW = opH(W)
            ncols += Nstates
            W, U = _rdiv_chol(W, W.conj().T @ W)
            HW = _rdiv_U(HW, U)
            if it > 1:
                Q = np.hstack([X, W, P])
                HQ = np.hstack([HX, HW, HP])
            else:
                Q = np.hstack([X, W])
                HQ = np.hstack([HX, HW])
            T = herm_u(Q.conj().T @ HQ)
            G = herm_u(Q.conj().T @ Q)
            sd, S = sla.eigh(T, G, lower=False)
            Um = S[:, :Nstates]
            X[:, :] = Q @ Um
            HX = HQ @ Um
            if it > 1:
                s2 = slice(Nstates, 2 * Nstates)
                s3 = slice(2 * Nstates, 3 * Nstates)
                P = W @ Um[s2, :] + P @ Um[s3, :]
                HP = HW @ Um[s2, :] + HP @ Um[s3, :]
                P, U = _rdiv_chol(P, P.conj().T @ P)
                HP = _rdiv_U(HP, U)
            else:
                P = W.copy()
                HP = HW.copy()
    S = X.conj().T @ HX
    S = herm_u(0.5 * (S + S.conj().T))
    evals, Qm = np.linalg.eigh(S)
    X[:, :] = X @ Qm
    return evals, it_done, ncols


def diag_davidson(opH, prec, X, tol=1e-5, NiterMax=100, Nstates_conv=0):
    """diag_davidson.jl:50-191 for one k-point (block Davidson on [X R]).  X is
    overwritten in place; returns (evals, iterations, number of opH columns applied)."""
    Nbasis, Nstates = X.shape
    if Nstates_conv == 0:
        Nstates_conv = Nstates
    HX = opH(X)
    ncols = Nstates
    evals = np.real(np.einsum("gi,gi->i", X.conj(), HX))
    R = evals[None, :] * X - HX
    res = np.sqrt(np.real(np.einsum("gi,gi->i", R.conj(), R)))
    s1 = slice(0, Nstates)
    s2 = slice(Nstates, 2 * Nstates)
    evals_old = evals.copy()
    Hred = np.zeros((2 * Nstates, 2 * Nstates), dtype=complex)
    Sred = np.zeros((2 * Nstates, 2 * Nstates), dtype=complex)
    it_done = 0
    for it in range(1, NiterMax + 1):
        it_done = it
        res_norm = np.where(EPS < res, 1.0 / np.where(res > 0, res, 1.0), 1.0)
        R = R * res_norm[None, :]
        R = prec(R)
        HR = opH(R)
        ncols += Nstates
        if it == 1:
            Hred[s1, s1] = X.conj().T @ HX
        else:
            Hred[s1, s1] = np.diag(evals.astype(complex))
        Hred[s1, s2] = X.conj().T @ HR
        Hred[s2, s2] = R.conj().T @ HR
        Hred[s2, s1] = Hred[s1, s2].conj().T
        Sred[s1, s1] = np.eye(Nstates)
        Sred[s1, s2] = X.conj().T @ R
        Sred[s2, s2] = R.conj().T @ R
        Sred[s2, s1] = Sred[s1, s2].conj().T
        Hred = 0.5 * (Hred + Hred.conj().T)
        Sred = 0.5 * (Sred + Sred.conj().T)
        lam, Xr = sla.eigh(herm_u(Hred), herm_u(Sred), lower=False)
        evals = lam[:Nstates].copy()
        devals = np.abs(evals - evals_old)
        nconv = int(np.count_nonzero(devals < tol))
        X[:, :] = X @ Xr[s1, s1] + R @ Xr[s2, s1]
        HX = HX @ Xr[s1, s1] + HR @ Xr[s2, s1]
        R = evals[None, :] * X - HX
        res = np.sqrt(np.real(np.einsum("gi,gi->i", R.conj(), R)))
        if nconv >= Nstates_conv:
            break
        evals_old = evals.copy()
    return evals, it_done, ncols


def diag_davidson_qe(opH, prec, evc, EBANDS_THR=1e-8, NiterMax=40, opS=None):
    """diag_davidson_qe! (diag_davidson_qe_v2.jl:40-216) for one k-point: QE's cegterg.
    The search space psi grows by the preconditioned, normalised corrections
    (H - e S) psi vc of the unconverged roots only, up to Nvecx = 2 Nvec columns; the
    reduced Hc, Sc are extended by their new columns; the basis is refreshed to the Ritz
    vectors when all roots have converged (|ew - evals| < EBANDS_THR), when it would
    overflow, or at iteration NiterMax.  opS defaults to the identity (op_S! of a GTH or
    norm-conserving Hamiltonian, op_S.jl:82-89).  evc is overwritten in place; returns
    (evals, iterations, number of opH columns applied)."""
    if opS is None:
        opS = lambda x: x.copy()
    N, Nvec = evc.shape
    Nvecx = 2 * Nvec
    evals = np.zeros(Nvec)
    psi = np.zeros((N, Nvecx), dtype=complex)
    Hpsi = np.zeros((N, Nvecx), dtype=complex)
    Spsi = np.zeros((N, Nvecx), dtype=complex)
    Sc = np.zeros((Nvecx, Nvecx), dtype=complex)
    Hc = np.zeros((Nvecx, Nvecx), dtype=complex)
    vc = np.zeros((Nvecx, Nvecx), dtype=complex)
    ew = np.zeros(Nvecx)
    is_conv = np.zeros(Nvec, dtype=bool)
    notcnv = Nvec
    nbase = Nvec
    psi[:, :Nvec] = evc
    Hpsi[:, :Nvec] = opH(psi[:, :Nvec])
    Spsi[:, :Nvec] = opS(psi[:, :Nvec])
    ncols = Nvec
    Hc[:nbase, :nbase] = psi[:, :nbase].conj().T @ Hpsi[:, :nbase]
    Sc[:nbase, :nbase] = psi[:, :nbase].conj().T @ Spsi[:, :nbase]
    w, v = sla.eigh(herm_u(Hc[:Nvec, :Nvec]), herm_u(Sc[:Nvec, :Nvec]), lower=False)
    ew[:Nvec] = w
    vc[:Nvec, :Nvec] = v
    evals[:] = w
    kter = 1
    dav_iter = 0
    while kter <= NiterMax or notcnv == 0:
        dav_iter = kter
        # unconverged roots first, with their eigenvalue estimates beyond nbase
        np_ = 0
        for ist in range(Nvec):
            if not is_conv[ist]:
                if np_ != ist:
                    vc[:, np_] = vc[:, ist]
                ew[nbase + np_] = evals[ist]
                np_ += 1
        nb1 = nbase
        new = slice(nb1, nb1 + notcnv)
        # correction vectors (H - e S) psi vc, preconditioned and normalised
        psi[:, new] = Spsi[:, :nbase] @ vc[:nbase, :notcnv]
        psi[:, new] *= -ew[nbase:nbase + notcnv][None, :]
        psi[:, new] += Hpsi[:, :nbase] @ vc[:nbase, :notcnv]
        psi[:, new] = prec(psi[:, new])
        ew[:notcnv] = np.real(np.einsum("gi,gi->i", psi[:, new].conj(), psi[:, new]))
        psi[:, new] /= np.sqrt(ew[:notcnv])[None, :]
        Hpsi[:, new] = opH(psi[:, new])
        Spsi[:, new] = opS(psi[:, new])
        ncols += notcnv
        Hc[:nbase + notcnv, new] = psi[:, :nbase + notcnv].conj().T @ Hpsi[:, new]
        Sc[:nbase + notcnv, new] = psi[:, :nbase + notcnv].conj().T @ Spsi[:, new]
        nbase += notcnv
        for n in range(nbase):
            Hc[n, n] = Hc[n, n].real
            Sc[n, n] = Sc[n, n].real
            Hc[n + 1:nbase, n] = Hc[n, n + 1:nbase].conj()
            Sc[n + 1:nbase, n] = Sc[n, n + 1:nbase].conj()
        w, v = sla.eigh(herm_u(Hc[:nbase, :nbase]), herm_u(Sc[:nbase, :nbase]), lower=False)
        ew[:nbase] = w
        vc[:nbase, :nbase] = v
        vc[:, Nvec:] = 0.0
        is_conv = np.abs(ew[:Nvec] - evals) < EBANDS_THR
        notcnv = int(np.count_nonzero(~is_conv))
        evals[:] = ew[:Nvec]
        if notcnv == 0 or nbase + notcnv > Nvecx or dav_iter == NiterMax:
            evc[:, :] = psi[:, :nbase] @ vc[:nbase, :Nvec]
            if notcnv == 0 or dav_iter == NiterMax:
                break
            psi[:, :Nvec] = evc
            Snew = Spsi[:, :nbase] @ vc[:nbase, :Nvec]
            Hnew = Hpsi[:, :nbase] @ vc[:nbase, :Nvec]
            Spsi[:, :Nvec] = Snew
            Hpsi[:, :Nvec] = Hnew
            nbase = Nvec
            Hc[:, :nbase] = 0.0
            Sc[:, :nbase] = 0.0
            vc[:, :nbase] = 0.0
            for n in range(nbase):
                Hc[n, n] = evals[n]
                Sc[n, n] = 1.0
                vc[n, n] = 1.0
        kter += 1
    return evals, dav_iter, ncols


# ---------------------------------------------------------------------------- CheFSI


def chebyfilt(opH, X, degree, lb, ub):
    """CheFSI.jl:36-62: the scaled three-term Chebyshev recurrence, which returns
    T_m((H - c)/e) X / T_m(sigma1^-1) with e = (ub - lb)/2, c = (ub + lb)/2 and
    sigma1 = e / (lb - ub) = -1/2."""
    ee = (ub - lb) / 2
    c = (ub + lb) / 2
    sigma = ee / (lb - ub)
    sigma1 = sigma
    Y = (opH(X) - X * c) * sigma1 / ee
    for _ in range(2, degree + 1):
        sigma2 = 1 / (2 / sigma1 - sigma)
        Y1 = (opH(Y) - Y * c) * 2 * sigma2 / ee - X * (sigma * sigma2)
        X = Y
        Y = Y1
        sigma = sigma2
    return Y


def get_ub_lb_lanczos(opH, v0, nlancz):
    """CheFSI.jl:65-130: nlancz Lanczos steps from v0 with one reorthogonalisation per step;
    returns (lb, ub) = (the nlancz/2-th eigenvalue of T, |T 1/sqrt(n)| + |f|).  T is the
    reference's upper-Hessenberg T (sub-diagonal beta, upper part real(h)); its eigenvalues
    are real to rounding and are sorted by real part as Julia's eigvals sorts them."""
    n = v0.size
    V = np.zeros((n, nlancz), dtype=complex)
    HV = np.zeros((n, nlancz), dtype=complex)
    T = np.zeros((nlancz, nlancz))
    h = np.zeros(nlancz, dtype=complex)
    s = np.zeros(nlancz, dtype=complex)
    V[:, 0] = v0 / np.linalg.norm(v0)
    HV[:, 0] = opH(V[:, :1])[:, 0]
    h[0] = np.real(np.vdot(V[:, 0], HV[:, 0]))
    T[0, 0] = h[0].real
    f = HV[:, 0] - V[:, 0] * h[0]
    s[0] = np.vdot(V[:, 0], f)
    h[0] += s[0]
    f = f - V[:, 0] * s[0]
    for j in range(1, nlancz):
        beta = np.linalg.norm(f)
        T[j, j - 1] = beta
        V[:, j] = f / beta
        HV[:, j] = opH(V[:, j:j + 1])[:, 0]
        h[:j + 1] = V[:, :j + 1].conj().T @ HV[:, j]
        f = HV[:, j] - V[:, :j + 1] @ h[:j + 1]
        s[:j + 1] = V[:, :j + 1].conj().T @ f
        h[:j + 1] += s[:j + 1]
        f = f - V[:, :j + 1] @ s[:j + 1]
        T[:j + 1, j] = np.real(h[:j + 1])
    ev = np.linalg.eigvals(T)
    ev = ev[np.lexsort((ev.imag, ev.real))]
    lb = ev[nlancz // 2 - 1].real
    ub = np.linalg.norm(T @ (np.ones(nlancz) / np.sqrt(nlancz))) + np.linalg.norm(f)
    return lb, ub


def diag_CheFSI(opH, X, degree, v0):
    """diag_CheFSI! (CheFSI.jl:1-31) for one k-point; returns (evals, filtered X).  The
    reference binds `ub, lb = get_ub_lb_lanczos(...)` to a function returning (lb, ub), so
    the filter receives the two bounds exchanged; that is kept (with sigma1 = -1/2 it only
    changes the filtered block's overall sign, which ortho_sqrt and eigvals do not see)."""
    ub, lb = get_ub_lb_lanczos(opH, v0, 2 * X.shape[1])
    Y = ortho_sqrt(chebyfilt(opH, X, degree, lb, ub))
    Hr = Y.conj().T @ opH(Y)
    return np.linalg.eigvalsh(herm_u(Hr)), Y


# ---------------------------------------------------------------------------- PCG


def calc_grad_evals(opH, psi):
    """calc_grad.jl:60-82: H psi - psi (psi' H psi) (the gradient of the band-energy sum)."""
    Hpsi = opH(psi)
    return Hpsi - psi @ (psi.conj().T @ Hpsi)


def diag_Emin_PCG(opH, prec, X, tol=1e-5, NiterMax=100, Nstates_conv=0, tol_ebands=1e-4, alpha_t=3e-5,
                  I_CG_BETA=2):
    """diag_Emin_PCG! (diag_Emin_PCG.jl:56-195): all-band preconditioned conjugate gradient
    on the band-energy sum, trial step alpha_t and the line-minimising step from the gradient
    at the trial point; convergence on |d evals| < tol for Nstates_conv bands or on the
    band-energy change.  X (orthonormal) is overwritten with the Ritz vectors; returns
    (evals, iterations, number of opH columns applied)."""
    Nstates = X.shape[1]
    if Nstates_conv == 0:
        Nstates_conv = Nstates
    rdot = lambda a, b: float(np.real(np.sum(np.conj(a) * b)))
    ncols = 0
    g_old = d_old = Kg_old = d = None
    beta = 0.0
    Ebands_old = 0.0
    Hr = herm_u(X.conj().T @ opH(X))
    ncols += Nstates
    evals = np.linalg.eigvalsh(Hr)
    evals_old = evals.copy()
    it_done = 0
    for it in range(1, NiterMax + 1):
        it_done = it
        g = calc_grad_evals(opH, X)
        Kg = prec(g)
        ncols += Nstates
        if it != 1:
            if I_CG_BETA == 1:
                beta = rdot(g, Kg) / rdot(g_old, Kg_old)
            elif I_CG_BETA == 2:
                beta = rdot(g - g_old, Kg) / rdot(g_old, Kg_old)
            elif I_CG_BETA == 3:
                beta = rdot(g - g_old, Kg) / rdot(g - g_old, d)
            else:
                beta = rdot(g, Kg) / float(np.real(np.sum((g - g_old) * np.conj(d_old))))
        if beta < 0.0:
            beta = 0.0
        d = -Kg + beta * d_old if it != 1 else -Kg + beta * np.zeros_like(Kg)
        Xc = ortho_sqrt(X + alpha_t * d)
        gt = calc_grad_evals(opH, Xc)
        ncols += Nstates
        denum = rdot(g - gt, d)
        alpha = abs(alpha_t * rdot(g, d) / denum) if denum != 0.0 else 0.0
        X[:] = ortho_sqrt(X + alpha * d)
        Hr = herm_u(X.conj().T @ opH(X))
        ncols += Nstates
        evals = np.linalg.eigvalsh(Hr)
        Ebands = float(np.sum(evals))
        devals = np.abs(evals - evals_old)
        evals_old = evals.copy()
        nconv = int(np.sum(devals < tol))
        diffE = abs(Ebands - Ebands_old)
        if nconv >= Nstates_conv or diffE <= tol_ebands * Nstates:
            break
        g_old, d_old, Kg_old = g.copy(), d.copy(), Kg.copy()
        Ebands_old = Ebands
    X[:] = ortho_sqrt(X)
    Hr = herm_u(X.conj().T @ opH(X))
    ncols += Nstates
    evals, evecs = np.linalg.eigh(Hr)
    X[:] = X @ evecs
    return evals, it_done, ncols


def calc_grad(opH, psi, Focc_k, wk):
    """calc_grad.jl:1-57: wk Focc_i (H psi - psi psi' H psi)_i, plus psi Q with
    Q = (Hr F - F Hr)/2 (Hr = psi' H psi, F = diag(Focc)) unless the occupations are all 2
    or all 1."""
    Hpsi = opH(psi)
    grad = (Hpsi - psi @ (psi.conj().T @ Hpsi)) * (np.asarray(Focc_k) * wk)[None, :]
    if np.all(Focc_k == 2.0) or np.all(Focc_k == 1.0):
        return grad
    F = np.diag(Focc_k)
    Hr = psi.conj().T @ Hpsi
    return grad + psi @ (0.5 * (Hr @ F - F @ Hr))
