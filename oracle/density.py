"""Oracle: density and local potentials -- calc_rhoe.jl, guess_rhoe.jl, Poisson_solve.jl,
XC_funcs/XC_x_slater.jl + XC_funcs/XC_c_vwn.jl (the internal Slater+VWN that
test/test_xc_internal.jl shows equal to the reference's Libxc call to 1e-5).
Test infrastructure only."""
import math
import numpy as np

from .setup import G_to_R, R_to_G, ZATOMS, calc_strfact


def calc_rhoe(ham, psiks, Focc=None):
    """calc_rhoe.jl:9-80 (norm-conserving, Nspin = 1, no dual grid, no symmetrisation):
    rho(r) = sum_k sum_ist wk[ik] * Focc[ist,ik] * |(N/sqrt(Omega)) IFFT[psi]|^2."""
    pw = ham.pw
    if Focc is None:
        Focc = ham.Focc
    rho = np.zeros(pw.Npoints)
    scale = pw.Npoints / math.sqrt(pw.CellVolume)
    ctmp = np.zeros(pw.Npoints, dtype=complex)
    for ik in range(pw.Nkpt):
        psi = psiks[ik]
        idx = pw.idx_gw2r[ik]
        for ist in range(psi.shape[1]):
            ctmp[:] = 0.0
            ctmp[idx] = psi[:, ist]
            c = G_to_R(pw, ctmp) * scale
            w = pw.wk[ik] * Focc[ist, ik]
            rho += w * np.real(np.conj(c) * c)
    return rho


def atmlength(zion, znucl):
    """guess_rhoe.jl:230-330 (ABINIT atmlength, densty = 0): Gaussian decay length."""
    coreel = znucl - zion
    nval = int(round(zion))
    if nval == 0:
        return 0.0
    tables = [
        (0.5, [0.6, 0.4, 0.3, 0.25], 0.2),
        (2.5, [1.8, 1.4, 1.0, 0.7, 0.6, 0.5, 0.4, 0.35], 0.3),
        (10.5, [2.0, 1.6, 1.25, 1.1, 1.0, 0.9, 0.8, 0.7, 0.7, 0.7], 0.6),
        (12.5, [1.9, 1.5, 1.15, 1.0, 0.9, 0.8, 0.7, 0.6, 0.6, 0.6], 0.5),
        (18.5, [2.0, 1.8, 1.5, 1.2, 1.0, 0.9, 0.85, 0.8, 0.75, 0.7, 0.65, 0.65], 0.6),
        (28.5, [1.5, 1.25, 1.15, 1.05, 1.00, 0.95, 0.95, 0.9, 0.9, 0.85, 0.85, 0.80, 0.8, 0.75], 0.7),
        (36.5, [2.0, 2.00, 1.60, 1.40, 1.25, 1.10, 1.00, 0.95, 0.90, 0.85, 0.80, 0.75], 0.7),
    ]
    for lim, data, default in tables:
        if coreel < lim:
            return data[nval - 1] if nval <= len(data) else default
    data = [2.0, 2.00, 1.55, 1.25, 1.15, 1.10, 1.05, 1.0, 0.95, 0.9, 0.85, 0.85]
    return data[nval - 1] if nval <= 12 else 0.8


def guess_rhoe(ham):
    """guess_rhoe.jl:1-38: Gaussian superposition, scaled to integrate to Nelectrons.
    (The species structure factor is added once per atom, as the reference does.)"""
    atoms, pw = ham.atoms, ham.pw
    strf = calc_strfact(atoms.positions, atoms.atm2species, atoms.Nspecies, pw.G)
    RhoeG = np.zeros(pw.Npoints, dtype=complex)
    for ia in range(atoms.Natoms):
        isp = atoms.atm2species[ia]
        psp = ham.pspots[isp]
        l = atmlength(psp.zval, ZATOMS[atoms.SpeciesSymbols[isp]])
        np.add.at(RhoeG, pw.idx_g2r, psp.zval * strf[:, isp] * np.exp(-pw.G2 * l ** 2))
    rho = np.real(G_to_R(pw, RhoeG))
    intRhoe = rho.sum() * pw.CellVolume / pw.Npoints
    return rho * ham.Nelectrons / intRhoe


def Poisson_solve(pw, rho):
    """Poisson_solve.jl:22-39 (Poisson_solve!): V_H(r).  Only the G-sphere entries are
    scaled by 4 pi / G^2; G = 0 is zeroed and entries outside the sphere are kept as is."""
    c = R_to_G(pw, rho.astype(complex))
    c[0] = 0.0
    ip = pw.idx_g2r[1:]
    c[ip] = 4 * np.pi * c[ip] / pw.G2[1:]
    return np.real(G_to_R(pw, c))


def xc_slater(rho):
    """XC_funcs/XC_x_slater.jl: returns (ex, vx); zero for rho <= 1e-10."""
    rho = np.asarray(rho, dtype=float)
    ok = rho > 1e-10
    r = np.where(ok, rho, 1.0)
    rs = 0.6203504908994 / r ** (1.0 / 3.0)
    f = -0.687247939924714
    alpha = 2.0 / 3.0
    ex = np.where(ok, f * alpha / rs, 0.0)
    vx = np.where(ok, 4.0 / 3.0 * f * alpha / rs, 0.0)
    return ex, vx


def xc_vwn(rho):
    """XC_funcs/XC_c_vwn.jl (VWN5 paramagnetic).  Densities below the 1e-10 threshold
    that the reference passes to Libxc (LDA_VWN.jl:175-186) give zero."""
    rho = np.asarray(rho, dtype=float)
    ok = rho > 1e-10
    r = np.where(ok, rho, 1.0)
    rs = 0.6203504908994 / r ** (1.0 / 3.0)
    a, b, c, x0 = 0.0310907, 3.72744, 12.9352, -0.10498
    q = math.sqrt(4.0 * c - b * b)
    f1 = 2.0 * b / q
    f2 = b * x0 / (x0 * x0 + b * x0 + c)
    f3 = 2.0 * (2.0 * x0 + b) / q
    rs12 = np.sqrt(rs)
    fx = rs + b * rs12 + c
    qx = np.arctan(q / (2.0 * rs12 + b))
    ec = a * (np.log(rs / fx) + f1 * qx - f2 * (np.log((rs12 - x0) ** 2 / fx) + f3 * qx))
    tx = 2.0 * rs12 + b
    tt = tx * tx + q * q
    vc = ec - rs12 * a / 6.0 * (2.0 / rs12 - tx / fx - 4.0 * b / tt
                                - f2 * (2.0 / (rs12 - x0) - tx / fx - 4.0 * (2.0 * x0 + b) / tt))
    return np.where(ok, ec, 0.0), np.where(ok, vc, 0.0)


def calc_epsxc_Vxc_VWN(rho):
    """LDA_VWN_internal.jl:34-53: eps_xc = ex + ec, V_xc = vx + vc."""
    ex, vx = xc_slater(rho)
    ec, vc = xc_vwn(rho)
    return ex + ec, vx + vc
