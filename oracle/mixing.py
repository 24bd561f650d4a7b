"""Oracle: the density mixers of KS_solve_SCF! (KS_solve_SCF.jl:240-275), test
infrastructure only.  Each function restates the reference function of the same name on
numpy arrays, in place, with the reference's 1-based iteration counter and history layout:
x / deltain / vin / mu is the input density (overwritten with the mixed one), gx / deltaout
/ vout / nu the output density of the iteration."""
import numpy as np


def mix_broyden(deltain, deltaout_, alphamix, it, n_iter, df, dv):
    """mix_broyden.jl:4-70 (PWSCF's modified Broyden); df, dv: N x n_iter."""
    deltaout = deltaout_.copy()
    maxter = 8
    wg0 = 0.01
    wg = np.ones(maxter)
    deltainsave = deltain.copy()
    iter_used = min(it - 1, n_iter)
    ipos = it - 1 - int(np.floor((it - 2) / n_iter)) * n_iter
    deltaout -= deltain
    if it > 1:
        df[:, ipos - 1] = deltaout - df[:, ipos - 1]
        dv[:, ipos - 1] = deltain - dv[:, ipos - 1]
        nrm = np.sqrt(np.linalg.norm(df[:, ipos - 1]) ** 2)
        df[:, ipos - 1] /= nrm
        dv[:, ipos - 1] /= nrm
    beta = np.zeros((maxter, maxter))
    for i in range(iter_used):
        for j in range(i + 1, iter_used):
            beta[i, j] = wg[i] * wg[j] * np.dot(df[:, j], df[:, i])
            beta[j, i] = beta[i, j]
        beta[i, i] = wg0 ** 2 + wg[i] ** 2
    beta_inv = np.linalg.inv(beta[:iter_used, :iter_used]) if iter_used > 0 else np.zeros((0, 0))
    work = np.array([np.dot(df[:, i], deltaout) for i in range(iter_used)])
    deltain += alphamix * deltaout
    for i in range(iter_used):
        gammamix = sum(beta_inv[j, i] * wg[j] * work[j] for j in range(iter_used))
        deltain -= wg[i] * gammamix * (alphamix * df[:, i] + dv[:, i])
    inext = it - int(np.floor((it - 1) / n_iter)) * n_iter
    df[:, inext - 1] = deltaout
    dv[:, inext - 1] = deltainsave


def mix_pulay(x, gx, beta, X, F, it, MIXDIM, x_old, f_old):
    """mix_pulay.jl:1-42; X, F: N x MIXDIM."""
    f = gx - x
    if it == 1:
        dx, dfv = x.copy(), f.copy()
    else:
        dx, dfv = x - x_old, f - f_old
    if it <= MIXDIM:
        X[:, it - 1] = dx
        F[:, it - 1] = dfv
    else:
        X[:, :-1] = X[:, 1:].copy()
        F[:, :-1] = F[:, 1:].copy()
        X[:, MIXDIM - 1] = dx
        F[:, MIXDIM - 1] = dfv
    x_old[:] = x
    f_old[:] = f
    if it > MIXDIM:
        addv = (X + beta * F) @ np.linalg.inv(F.T @ F) @ (F.T @ f)
        x += beta * f - addv
    else:
        x += beta * f


def mix_rpulay(x, gx, beta, X, F, it, MIXDIM, x_old, f_old):
    """mix_rpulay.jl:11-94 (restarted Pulay, Pratapa and Suryanarayana 2015)."""
    f = gx - x
    if it > 1:
        dx = x - x_old
        dfv = f - f_old
    active_dim = (it - 1) % (MIXDIM + 1)
    if it > 1:
        if active_dim == 0:
            X[:, 0] = X[:, MIXDIM - 1]
            F[:, 0] = F[:, MIXDIM - 1]
            X[:, 1:] = 0.0
            F[:, 1:] = 0.0
            Xk, Fk = X[:, :1], F[:, :1]
        else:
            for i in range(active_dim, 1, -1):
                X[:, i - 1] = X[:, i - 2]
                F[:, i - 1] = F[:, i - 2]
            X[:, 0] = dx
            F[:, 0] = dfv
            Xk, Fk = X[:, :active_dim], F[:, :active_dim]
        addv = (Xk + beta * Fk) @ np.linalg.inv(Fk.T @ Fk) @ (Fk.T @ f)
        x_old[:] = x
        f_old[:] = f
        x += beta * f - addv
    else:
        X[:, 0] = x
        F[:, 0] = f
        x_old[:] = x
        f_old[:] = f
        x += beta * f


def mix_ppulay(x, gx, beta, X, F, it, MIXDIM, k, x_old, f_old):
    """mix_ppulay.jl:9-58 (periodic Pulay, Banerjee, Suryanarayana and Pask 2016)."""
    f = gx - x
    if it == 1:
        dx, dfv = x.copy(), f.copy()
    else:
        dx, dfv = x - x_old, f - f_old
    is_pulay = False if it < MIXDIM else it % k == 0
    if it <= MIXDIM:
        X[:, it - 1] = dx
        F[:, it - 1] = dfv
    else:
        X[:, :-1] = X[:, 1:].copy()
        F[:, :-1] = F[:, 1:].copy()
    if it > MIXDIM:
        X[:, MIXDIM - 1] = dx
        F[:, MIXDIM - 1] = dfv
    x_old[:] = x
    f_old[:] = f
    if is_pulay:
        addv = (X + beta * F) @ np.linalg.inv(F.T @ F) @ (F.T @ f)
        x += beta * f - addv
    else:
        x += beta * f


def mix_anderson(vin, vout, beta, df, dv, it, mixdim):
    """mix_anderson.jl:1-40 (KSSOLV's Anderson mixing); df, dv: N x mixdim."""
    dvout = vout - vin
    iterused = min(it - 1, mixdim)
    ipos = it - 1 - int(np.floor((it - 2) / mixdim)) * mixdim
    if it > 1:
        df[:, ipos - 1] -= dvout
        dv[:, ipos - 1] -= vin
    vinsave = vin.copy()
    dvoutsave = dvout.copy()
    if it > 1:
        gammas = np.linalg.pinv(df[:, :iterused]) @ dvout
        for i in range(iterused):
            vin -= gammas[i] * dv[:, i]
            dvout -= gammas[i] * df[:, i]
    inext = it - int(np.floor((it - 1) / mixdim)) * mixdim
    df[:, inext - 1] = dvoutsave
    dv[:, inext - 1] = vinsave
    vin += beta * dvout


def mix_adaptive(mu, nu, beta0, beta, f, betamax=0.8):
    """mix_adaptive.jl:34-54."""
    for i in range(mu.size):
        t = nu[i] - mu[i]
        if t * f[i] >= 0.0:
            beta[i] = min(beta[i] + beta0, betamax)
        else:
            beta[i] = 0.5 * (beta[i] + beta0)
        f[i] = t
        mu[i] = beta[i] * nu[i] + (1.0 - beta[i]) * mu[i]
