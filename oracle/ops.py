"""Oracle: Hamiltonian application, restated from op_H.jl / op_K.jl / op_V_loc.jl /
op_V_Ps_nloc.jl / PsPotNL.jl / Kprec.jl.  Test infrastructure only.

`ham` is an oracle Hamiltonian (energy.Hamiltonian) with pw, V_loc (Npoints x Nspin),
pspotNL, and the cursor ik/ispin of Hamiltonian.jl:16-17.
"""
import numpy as np

from .setup import G_to_R, R_to_G


def op_K(ham, psi):
    """op_K.jl:55-75: Hpsi += 0.5*|G+k|^2 psi (allocating form returns the term)."""
    gk2 = ham.pw.gk2(ham.ik)
    return 0.5 * psi * gk2[:, None]


def op_V_loc(ham, psi, V=None):
    """op_V_loc.jl:33-92 per band: scatter psi to the grid at idx_gw2r, G_to_R
    (normalised IFFT), multiply by V_loc[:, ispin], R_to_G (unnormalised FFT),
    gather at idx_gw2r."""
    pw = ham.pw
    if V is None:
        V = ham.V_loc[:, ham.ispin]
    idx = pw.idx_gw2r[ham.ik]
    out = np.zeros_like(psi)
    ctmp = np.zeros(pw.Npoints, dtype=complex)
    for ist in range(psi.shape[1]):
        ctmp[:] = 0.0
        ctmp[idx] = psi[:, ist]
        r = G_to_R(pw, ctmp) * V
        out[:, ist] = R_to_G(pw, r)[idx]
    return out


def calc_betaNL_psi(beta, psi):
    """PsPotNL.jl:242-254: conj(psi' * betaNL) -> Nstates x Nbeta."""
    return np.conj(psi.conj().T @ beta)


def op_V_Ps_nloc(ham, psi):
    """op_V_Ps_nloc.jl:39-78 (GTH): Hpsi += hij * beta_i * betaNL_psi[:, j] summed over
    the (atom, l, m, iprj, jprj) blocks -- equal to beta @ D @ (beta^H psi)."""
    nl = ham.pspotNL
    if nl.NbetaNL == 0:
        return np.zeros_like(psi)
    beta = nl.betaNL[ham.ik]
    bp = calc_betaNL_psi(beta, psi)            # Nst x Nbeta
    out = np.zeros_like(psi)
    for ia, l, m, ids in nl.blocks:
        for ib in ids:
            for jb in ids:
                hij = nl.D[ib, jb]
                out += hij * np.outer(beta[:, ib], bp[:, jb])
    return out


def op_H(ham, psi):
    """op_H.jl:4-27: op_K + op_V_loc + op_V_Ps_nloc (+ op_Vtau = 0 for LDA/PBE)."""
    out = op_K(ham, psi) + op_V_loc(ham, psi)
    if ham.pspotNL.NbetaNL > 0:
        out = out + op_V_Ps_nloc(ham, psi)
    return out


def Kprec(ham, psi):
    """Kprec.jl:10-28: psi / (1 + |G+k|^2)  (no 1/2, not Teter)."""
    return psi / (1.0 + ham.pw.gk2(ham.ik))[:, None]
