"""Oracle: the nonlocal pseudopotential terms of a norm-conserving UPF Hamiltonian, where
each atom carries a dense nh x nh block Deeq of projector couplings: op_V_Ps_nloc!
(op_V_Ps_nloc.jl:82-126), calc_E_Ps_nloc (calc_energies.jl:61-128) and the nonlocal forces
(calc_forces_Ps_nloc.jl:31-160 with _calc_Deff!, calc_Deff.jl:1-26).  Test infrastructure
only.

Arguments follow PsPotNL_UPF: betaNL[ik] is Ngw[ik] x NbetaNL; Deeq is nhm x nhm x Natoms
(real); nh[isp] projectors per species; indv_ijkb0[ia] is the 0-based offset of atom ia's
first projector; atm2species is 0-based.  For norm-conserving data Deff = Deeq (there is
no augmentation charge, so the qq_at term of _calc_Deff! is absent).
"""
import numpy as np


def _atom_rows(ia, nh, indv_ijkb0, atm2species):
    return slice(indv_ijkb0[ia], indv_ijkb0[ia] + nh[atm2species[ia]])


def op_V_Ps_nloc_upf(betaNL_k, Deeq, nh, indv_ijkb0, atm2species, psi):
    """op_V_Ps_nloc.jl:82-126: ps[idx_a, :] = Deeq[:, :, a] * (beta' psi)[idx_a, :] for each
    atom, then Hpsi += beta ps (returned as the added term)."""
    betaNL_psi = betaNL_k.conj().T @ psi
    ps = np.zeros_like(betaNL_psi)
    for ia in range(len(atm2species)):
        n = nh[atm2species[ia]]
        if n == 0:
            continue
        r = _atom_rows(ia, nh, indv_ijkb0, atm2species)
        ps[r, :] = Deeq[:n, :n, ia] @ betaNL_psi[r, :]
    return betaNL_k @ ps


def calc_E_Ps_nloc_upf(betaNL, Deeq, nh, indv_ijkb0, atm2species, psiks, Focc, wk):
    """calc_energies.jl:61-128: sum_k wk sum_ist Focc Re <psi | beta D beta' | psi>.  The
    reference takes Dvan per species there; for norm-conserving data each atom's Deeq is
    its species' Dvan, which is what is passed here."""
    E = 0.0
    for ik, psi in enumerate(psiks):
        Vpsi = op_V_Ps_nloc_upf(betaNL[ik], Deeq, nh, indv_ijkb0, atm2species, psi)
        for ist in range(psi.shape[1]):
            E += wk[ik] * Focc[ist, ik] * np.real(np.vdot(psi[:, ist], Vpsi[:, ist]))
    return E


def calc_forces_Ps_nloc_upf(betaNL, G, idx_gw2g, Deeq, nh, indv_ijkb0, atm2species, psiks, Focc, wk):
    """calc_forces_Ps_nloc.jl:31-88 and _force_Ps_nloc_k! (:90-160), norm-conserving:
    dbeta = -i G beta (G of the rho sphere, without k, as the reference writes it),
    F[ipol, a] -= 2 wk f [ sum_i D_ii Re(conj(dB_i) B_i)
                          + sum_{i<j} D_ij Re(conj(dB_i) B_j + dB_j conj(B_i)) ]
    with B = beta' psi and dB = dbeta' psi over atom a's projectors."""
    Natoms = len(atm2species)
    F = np.zeros((3, Natoms))
    for ik, psi in enumerate(psiks):
        beta = betaNL[ik]
        B = beta.conj().T @ psi
        Gk = G[:, idx_gw2g[ik]]
        for ipol in range(3):
            dbeta = -1j * beta * Gk[ipol][:, None]
            dB = dbeta.conj().T @ psi
            for ist in range(psi.shape[1]):
                fac = wk[ik] * Focc[ist, ik]
                for ia in range(Natoms):
                    n = nh[atm2species[ia]]
                    o = indv_ijkb0[ia]
                    D = Deeq[:, :, ia]
                    for ih in range(n):
                        F[ipol, ia] -= 2.0 * fac * D[ih, ih] * np.real(np.conj(dB[o + ih, ist]) * B[o + ih, ist])
                    for ih in range(n):
                        for jh in range(ih + 1, n):
                            F[ipol, ia] -= 2.0 * fac * D[ih, jh] * np.real(
                                np.conj(dB[o + ih, ist]) * B[o + jh, ist] + dB[o + jh, ist] * np.conj(B[o + ih, ist]))
    return F


def assemble_D(Deeq, nh, indv_ijkb0, atm2species, NbetaNL):
    """The NbetaNL x NbetaNL block-diagonal matrix whose atom blocks are Deeq[:, :, a]."""
    D = np.zeros((NbetaNL, NbetaNL))
    for ia in range(len(atm2species)):
        n = nh[atm2species[ia]]
        r = _atom_rows(ia, nh, indv_ijkb0, atm2species)
        D[r, r] = Deeq[:n, :n, ia]
    return D
