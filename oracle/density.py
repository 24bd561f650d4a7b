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


# ----------------------------------------------------------------------------- PBE (GGA)

def op_nabla(pw, f):
    """PWGrid.jl:351-372: grad f = G_to_R(i G f(G)) on the rho sphere (real part)."""
    fG = R_to_G(pw, np.asarray(f, dtype=complex))[pw.idx_g2r]
    out = np.zeros((3, pw.Npoints))
    for d in range(3):
        full = np.zeros(pw.Npoints, dtype=complex)
        full[pw.idx_g2r] = 1j * pw.G[d] * fG
        out[d] = np.real(G_to_R(pw, full))
    return out


def op_nabla_dot(pw, h):
    """PWGrid.jl:375-399: div h = G_to_R(i G . h(G)) on the rho sphere (real part)."""
    full = np.zeros(pw.Npoints, dtype=complex)
    for d in range(3):
        hG = R_to_G(pw, np.asarray(h[d], dtype=complex))[pw.idx_g2r]
        full[pw.idx_g2r] += 1j * pw.G[d] * hG
    return np.real(G_to_R(pw, full))


def xc_x_pbe(rho, grho):
    """XC_funcs/XC_x_pbe.jl: (sx, v1x, v2x); zero for rho <= 1e-6 or |grad rho|^2 <= 1e-10."""
    ok = (rho > 1e-6) & (grho > 1e-10)
    r = np.where(ok, rho, 1.0)
    g = np.where(ok, grho, 1.0)
    third = 1.0 / 3.0
    k, mu = 0.804, 0.2195149727645171
    c1, c2, c5 = 0.75 / np.pi, 3.093667726280136, 4.0 * third
    agrho = np.sqrt(g)
    kf = c2 * r ** third
    dsg = 0.5 / kf
    s1 = agrho * dsg / r
    s2 = s1 * s1
    ds = -c5 * s1
    f1 = s2 * mu / k
    f2 = 1.0 + f1
    f3 = k / f2
    fx = k - f3
    exunif = -c1 * kf
    sx_s = exunif * fx
    dxunif = exunif * third
    dfx = 2.0 * mu * s1 / (f2 * f2)
    v1x = sx_s + dxunif * fx + exunif * dfx * ds
    v2x = exunif * dfx * dsg / agrho
    return (np.where(ok, sx_s * r, 0.0), np.where(ok, v1x, 0.0), np.where(ok, v2x, 0.0))


def xc_c_pw(rho):
    """XC_funcs/XC_c_pw.jl (Perdew-Wang 92 LDA correlation): (ec, vc); zero for rho <= 1e-10."""
    ok = rho > 1e-10
    r = np.where(ok, rho, 1.0)
    rs = 0.6203504908994 / r ** (1.0 / 3.0)
    a, a1, b1, b2, b3, b4 = 0.031091, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294
    rs12 = np.sqrt(rs)
    rs32 = rs * rs12
    rs2 = rs * rs
    om = 2.0 * a * (b1 * rs12 + b2 * rs + b3 * rs32 + b4 * rs2)
    dom = 2.0 * a * (0.5 * b1 * rs12 + b2 * rs + 1.5 * b3 * rs32 + 2.0 * b4 * rs2)
    olog = np.log(1.0 + 1.0 / om)
    ec = -2.0 * a * (1.0 + a1 * rs) * olog
    vc = -2.0 * a * (1.0 + 2.0 / 3.0 * a1 * rs) * olog - 2.0 / 3.0 * a * (1.0 + a1 * rs) * dom / (om * (om + 1.0))
    return np.where(ok, ec, 0.0), np.where(ok, vc, 0.0)


def xc_c_pbe(rho, grho):
    """XC_funcs/XC_c_pbe.jl (gradient part of PBE correlation): (sc, v1c, v2c)."""
    ok = (rho > 1e-6) & (grho > 1e-10)
    r = np.where(ok, rho, 1.0)
    g = np.where(ok, grho, 1.0)
    ga, be = 0.0310906908696548950, 0.06672455060314922
    pi34, xkf, xks = 0.6203504908994, 1.919158292677513, 1.128379167095513
    rs = pi34 / r ** (1.0 / 3.0)
    ec, vc = xc_c_pw(r)
    kf = xkf / rs
    ks = xks * np.sqrt(kf)
    t = np.sqrt(g) / (2.0 * ks * r)
    expe = np.exp(-ec / ga)
    af = be / ga * (1.0 / (expe - 1.0))
    bf = expe * (vc - ec)
    y = af * t * t
    xy = (1.0 + y) / (1.0 + y + y * y)
    qy = y * y * (2.0 + y) / (1.0 + y + y * y) ** 2
    s1 = 1.0 + be / ga * t * t * xy
    h0 = ga * np.log(s1)
    dh0 = be * t * t / s1 * (-7.0 / 3.0 * xy - qy * (af * bf / be - 7.0 / 3.0))
    ddh0 = be / (2.0 * ks * ks * r) * (xy - qy) / s1
    return (np.where(ok, r * h0, 0.0), np.where(ok, h0 + dh0, 0.0), np.where(ok, ddh0, 0.0))


def calc_epsxc_Vxc_PBE(pw, rho):
    """GGA_PBE_internal.jl:29-55 (epsxc) and :119-166 (Vxc), Nspin = 1:
    eps = ex + ec + (sx + sc)/rho, V = vx + vc + v1x + v1c - div((v2x + v2c) grad rho)."""
    rho = np.asarray(rho, dtype=float)
    grad = op_nabla(pw, rho)
    g2 = grad[0] ** 2 + grad[1] ** 2 + grad[2] ** 2
    ex, vx = xc_slater(rho)
    sx, v1x, v2x = xc_x_pbe(rho, g2)
    ec, vc = xc_c_pw(rho)
    sc, v1c, v2c = xc_c_pbe(rho, g2)
    eps = ex + ec + (sx + sc) / rho
    V = vx + vc + v1x + v1c - op_nabla_dot(pw, (v2x + v2c)[None, :] * grad)
    return eps, V
