"""Oracle: Hellmann-Feynman forces restated from calc_forces_Ps_loc.jl,
calc_forces_Ps_nloc.jl (GTH) and calc_forces_NN.jl.  Test infrastructure only."""
import math
import numpy as np
from scipy.special import erfc

from . import ops
from .setup import R_to_G, eval_Vloc_G


def calc_forces_Ps_loc(ham, rho):
    """calc_forces_Ps_loc.jl:13-71:
    F[:, a] += Re( i G V_sp(|G|) conj(rho(G)) exp(-i G.tau_a) ), G != 0, rho(G) = FFT(rho)/N."""
    pw, atoms = ham.pw, ham.atoms
    rhoG = R_to_G(pw, rho.astype(complex)) / pw.Npoints
    F = np.zeros((3, atoms.Natoms))
    G = pw.G[:, 1:]
    cr = np.conj(rhoG[pw.idx_g2r[1:]])
    for ia in range(atoms.Natoms):
        isp = atoms.atm2species[ia]
        Vgl = eval_Vloc_G(ham.pspots[isp], pw.G2_shells)
        V = Vgl[pw.idx_g2shells[1:]]
        X = atoms.positions[:, ia]
        GX = G[0] * X[0] + G[1] * X[1] + G[2] * X[2]
        Sf = np.cos(GX) - 1j * np.sin(GX)
        F[:, ia] = np.real(1j * G @ (V * cr * Sf))
    return F


def calc_forces_Ps_nloc(ham, psiks, Focc):
    """calc_forces_Ps_nloc.jl:291-343 with calc_dbetaNL (:347-391) = beta * (i g) and
    calc_dbetaNL_psi (:393-411):
    F[:, a] += 2 wk Focc Re( h_ij conj(<beta_i|psi>) <dbeta_j|psi> )."""
    pw, atoms, nl = ham.pw, ham.atoms, ham.pspotNL
    F = np.zeros((3, atoms.Natoms))
    if nl.NbetaNL == 0:
        return F
    for ik, psi in enumerate(psiks):
        beta = nl.betaNL[ik]
        g = pw.gk(ik)
        bp = ops.calc_betaNL_psi(beta, psi)                          # Nst x Nb
        dbp = np.stack([(np.conj(1j * g[d][:, None] * beta)).T @ psi for d in range(3)])  # 3 x Nb x Nst
        w = pw.wk[ik] * Focc[:, ik]
        for ia, l, m, ids in nl.blocks:
            if ia is None:
                continue
            for ib in ids:
                for jb in ids:
                    hij = nl.D[ib, jb]
                    fnl = hij * np.conj(bp[:, ib])[None, :] * dbp[:, jb, :]   # 3 x Nst
                    F[:, ia] += np.sum(2 * w[None, :] * np.real(fnl), axis=1)
    return F


def _Herfc(x):
    """calc_forces_NN.jl:11-13."""
    return -2 * np.exp(-x ** 2) / math.sqrt(np.pi) - erfc(x) / x


def calc_forces_NN(ham):
    """calc_forces_NN.jl:15-141: Ewald ion-ion forces."""
    pw, atoms = ham.pw, ham.atoms
    Z = ham.Zvals[atoms.atm2species]
    tau = atoms.positions
    Omega = pw.CellVolume
    charge = float(np.sum(Z))
    G2max = pw.ecutrho
    alpha = 1.0
    ub = charge ** 2 * math.sqrt(alpha / np.pi) * erfc(math.sqrt(G2max / 2.0 / alpha))
    while ub > 1e-6:
        alpha -= 0.1
        if alpha <= 0.0:
            raise RuntimeError("Optimal alpha is not found")
        ub = charge ** 2 * math.sqrt(alpha / np.pi) * erfc(math.sqrt(G2max / 2.0 / alpha))
    G = pw.G[:, 1:]
    G2 = pw.G2[1:]
    aux = np.zeros(G.shape[1], dtype=complex)
    for ia in range(atoms.Natoms):
        GX = tau[0, ia] * G[0] + tau[1, ia] * G[1] + tau[2, ia] * G[2]
        aux += Z[ia] * (np.cos(GX) + 1j * np.sin(GX))
    aux = aux * np.exp(-G2 / alpha / 4.0) / G2
    F_G = np.zeros((3, atoms.Natoms))
    for ia in range(atoms.Natoms):
        GX = G[0] * tau[0, ia] + G[1] * tau[1, ia] + G[2] * tau[2, ia]
        sumnb = np.cos(GX) * np.imag(aux) - np.sin(GX) * np.real(aux)
        F_G[:, ia] = (G @ sumnb) * (-4 * np.pi * Z[ia] / Omega)
    eta = math.sqrt(alpha)
    gexp = -math.log(1e-10)
    tmax = math.sqrt(0.5 * gexp) / eta
    t = [pw.LatVecs[:, i] for i in range(3)]
    mm = [int(round(tmax / math.sqrt(np.dot(v, v)) + 1.5)) for v in t]
    rng = [np.arange(-mm[i], mm[i] + 1) for i in range(3)]
    I, J, K = np.meshgrid(*rng, indexing="ij")
    T = np.outer(t[0], I.ravel()) + np.outer(t[1], J.ravel()) + np.outer(t[2], K.ravel())
    F_R = np.zeros((3, atoms.Natoms))
    for ia in range(atoms.Natoms):
        for ja in range(ia + 1, atoms.Natoms):
            D = (tau[:, ia] - tau[:, ja])[:, None] - T
            D2 = D[0] ** 2 + D[1] ** 2 + D[2] ** 2
            xx = _Herfc(eta * np.sqrt(D2))
            df = (eta * xx / D2 * Z[ia] * Z[ja])[None, :] * D
            s = df.sum(axis=1)
            F_R[:, ia] -= s
            F_R[:, ja] += s
    return F_G + F_R
