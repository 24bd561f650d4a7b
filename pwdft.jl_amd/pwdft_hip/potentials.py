"""Potential update and ion-ion terms used around the hot path: guess_rhoe
(guess_rhoe.jl:1-38), Poisson_solve! (Poisson_solve.jl:22-39), Slater + VWN LDA exchange-
correlation (XC_funcs/XC_x_slater.jl, XC_funcs/XC_c_vwn.jl), Ewald E_NN (calc_E_NN.jl:60-172).
Host-side, once per SCF iteration (O(N_grid)), in FP64."""
import math

import numpy as np
from scipy.special import erfc

from .basis import ZATOMS, strfact


def _ifft(pw, f):
    return np.fft.ifftn(np.asarray(f).reshape(pw.Ns, order="F")).reshape(-1, order="F")


def _fft(pw, f):
    return np.fft.fftn(np.asarray(f, dtype=complex).reshape(pw.Ns, order="F")).reshape(-1, order="F")


_ATMLEN = [  # guess_rhoe.jl:230-330: (core-electron limit, table by nval, default)
    (0.5, [0.6, 0.4, 0.3, 0.25], 0.2),
    (2.5, [1.8, 1.4, 1.0, 0.7, 0.6, 0.5, 0.4, 0.35], 0.3),
    (10.5, [2.0, 1.6, 1.25, 1.1, 1.0, 0.9, 0.8, 0.7, 0.7, 0.7], 0.6),
    (12.5, [1.9, 1.5, 1.15, 1.0, 0.9, 0.8, 0.7, 0.6, 0.6, 0.6], 0.5),
    (18.5, [2.0, 1.8, 1.5, 1.2, 1.0, 0.9, 0.85, 0.8, 0.75, 0.7, 0.65, 0.65], 0.6),
    (28.5, [1.5, 1.25, 1.15, 1.05, 1.00, 0.95, 0.95, 0.9, 0.9, 0.85, 0.85, 0.80, 0.8, 0.75], 0.7),
    (36.5, [2.0, 2.00, 1.60, 1.40, 1.25, 1.10, 1.00, 0.95, 0.90, 0.85, 0.80, 0.75], 0.7),
    (1e9, [2.0, 2.00, 1.55, 1.25, 1.15, 1.10, 1.05, 1.0, 0.95, 0.9, 0.85, 0.85], 0.8),
]


def atmlength(zion, znucl):
    nval = int(round(zion))
    if nval == 0:
        return 0.0
    for lim, tab, dflt in _ATMLEN:
        if znucl - zion < lim:
            return tab[nval - 1] if nval <= len(tab) else dflt


def guess_rhoe(atoms, pw, pspots, Nelectrons):
    """guess_rhoe.jl:1-38 (species structure factor added once per atom, as there)."""
    Sf = strfact(atoms.positions, atoms.atm2species, atoms.Nspecies, pw.G)
    rG = np.zeros(pw.Npoints, dtype=complex)
    for ia in range(atoms.Natoms):
        isp = atoms.atm2species[ia]
        l = atmlength(pspots[isp].zval, ZATOMS[atoms.SpeciesSymbols[isp]])
        rG[pw.idx_g2r] += pspots[isp].zval * Sf[:, isp] * np.exp(-pw.G2 * l ** 2)
    rho = np.real(_ifft(pw, rG))
    return rho * Nelectrons / (rho.sum() * pw.CellVolume / pw.Npoints)


def poisson_solve(pw, rho):
    """Poisson_solve.jl:22-39: 4 pi rho(G)/G^2 on the rho sphere, G = 0 zeroed; grid
    entries outside the sphere are left unscaled, as the reference does."""
    c = _fft(pw, rho)
    c[0] = 0.0
    c[pw.idx_g2r[1:]] *= 4 * np.pi / pw.G2[1:]
    return np.real(_ifft(pw, c))


def epsxc_vxc_vwn(rho):
    """Slater exchange + VWN5 correlation (spin-unpolarised); zero below rho = 1e-10,
    the density threshold the reference gives Libxc (LDA_VWN.jl:175-186)."""
    rho = np.asarray(rho, dtype=float)
    ok = rho > 1e-10
    rs = 0.6203504908994 / np.where(ok, rho, 1.0) ** (1.0 / 3.0)
    fa = -0.687247939924714 * 2.0 / 3.0
    ex, vx = fa / rs, 4.0 / 3.0 * fa / rs
    a, b, c, x0 = 0.0310907, 3.72744, 12.9352, -0.10498
    q = math.sqrt(4.0 * c - b * b)
    f1, f2, f3 = 2.0 * b / q, b * x0 / (x0 * x0 + b * x0 + c), 2.0 * (2.0 * x0 + b) / q
    r12 = np.sqrt(rs)
    fx = rs + b * r12 + c
    qx = np.arctan(q / (2.0 * r12 + b))
    ec = a * (np.log(rs / fx) + f1 * qx - f2 * (np.log((r12 - x0) ** 2 / fx) + f3 * qx))
    tx = 2.0 * r12 + b
    tt = tx * tx + q * q
    vc = ec - r12 * a / 6.0 * (2.0 / r12 - tx / fx - 4.0 * b / tt - f2 * (2.0 / (r12 - x0) - tx / fx - 4.0 * (2.0 * x0 + b) / tt))
    return np.where(ok, ex + ec, 0.0), np.where(ok, vx + vc, 0.0)


def _grad(pw, f):
    """op_nabla (PWGrid.jl:351-372): i G f(G) on the rho sphere back to real space."""
    fG = _fft(pw, f)[pw.idx_g2r]
    out = np.empty((3, pw.Npoints))
    full = np.zeros(pw.Npoints, dtype=complex)
    for d in range(3):
        full[pw.idx_g2r] = 1j * pw.G[d] * fG
        out[d] = np.real(_ifft(pw, full))
    return out


def _div(pw, h):
    """op_nabla_dot (PWGrid.jl:375-399): i G . h(G) on the rho sphere back to real space."""
    acc = np.zeros(pw.Ng, dtype=complex)
    for d in range(3):
        acc += 1j * pw.G[d] * _fft(pw, h[d])[pw.idx_g2r]
    full = np.zeros(pw.Npoints, dtype=complex)
    full[pw.idx_g2r] = acc
    return np.real(_ifft(pw, full))


def _pw92(r):
    """Perdew-Wang 92 correlation of the uniform gas (XC_funcs/XC_c_pw.jl), r > 0."""
    rs = 0.6203504908994 / r ** (1.0 / 3.0)
    a, a1, b1, b2, b3, b4 = 0.031091, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294
    q = np.sqrt(rs)
    om = 2.0 * a * (b1 * q + b2 * rs + b3 * rs * q + b4 * rs * rs)
    dom = 2.0 * a * (0.5 * b1 * q + b2 * rs + 1.5 * b3 * rs * q + 2.0 * b4 * rs * rs)
    lg = np.log(1.0 + 1.0 / om)
    ec = -2.0 * a * (1.0 + a1 * rs) * lg
    vc = -2.0 * a * (1.0 + 2.0 / 3.0 * a1 * rs) * lg - 2.0 / 3.0 * a * (1.0 + a1 * rs) * dom / (om * (om + 1.0))
    return ec, vc


def epsxc_vxc_pbe(pw, rho):
    """PBE (GGA_PBE_internal.jl:29-55, 119-166, with XC_funcs/XC_x_slater.jl, XC_x_pbe.jl,
    XC_c_pw.jl, XC_c_pbe.jl), spin-unpolarised: eps_xc and V_xc on the grid.  The density
    thresholds are the reference's: 1e-10 for the LDA parts, 1e-6 (and 1e-10 for
    |grad rho|^2) for the gradient corrections."""
    rho = np.asarray(rho, dtype=float)
    grad = _grad(pw, rho)
    g2 = grad[0] ** 2 + grad[1] ** 2 + grad[2] ** 2
    lda = rho > 1e-10
    r = np.where(lda, rho, 1.0)
    rs = 0.6203504908994 / r ** (1.0 / 3.0)
    ex = np.where(lda, -0.687247939924714 * (2.0 / 3.0) / rs, 0.0)
    vx = 4.0 / 3.0 * ex
    ec, vc = _pw92(r)
    ec, vc = np.where(lda, ec, 0.0), np.where(lda, vc, 0.0)
    gga = (rho > 1e-6) & (g2 > 1e-10)
    rg = np.where(gga, rho, 1.0)
    gg = np.where(gga, g2, 1.0)
    # exchange enhancement
    kf = 3.093667726280136 * rg ** (1.0 / 3.0)
    ag = np.sqrt(gg)
    dsg = 0.5 / kf
    s = ag * dsg / rg
    kap, mu = 0.804, 0.2195149727645171
    f2 = 1.0 + s * s * mu / kap
    fx = kap - kap / f2
    exu = -0.75 / np.pi * kf
    dfx = 2.0 * mu * s / (f2 * f2)
    sx = exu * fx * rg
    v1x = exu * fx + exu / 3.0 * fx + exu * dfx * (-4.0 / 3.0 * s)
    v2x = exu * dfx * dsg / ag
    # correlation gradient part
    ga, be = 0.0310906908696548950, 0.06672455060314922
    ecg, vcg = _pw92(rg)
    ks = 1.128379167095513 * np.sqrt(1.919158292677513 / (0.6203504908994 / rg ** (1.0 / 3.0)))
    t = ag / (2.0 * ks * rg)
    ee = np.exp(-ecg / ga)
    af = be / ga / (ee - 1.0)
    bf = ee * (vcg - ecg)
    y = af * t * t
    den = 1.0 + y + y * y
    xy = (1.0 + y) / den
    qy = y * y * (2.0 + y) / den ** 2
    s1 = 1.0 + be / ga * t * t * xy
    h0 = ga * np.log(s1)
    sc = rg * h0
    v1c = h0 + be * t * t / s1 * (-7.0 / 3.0 * xy - qy * (af * bf / be - 7.0 / 3.0))
    v2c = be / (2.0 * ks * ks * rg) * (xy - qy) / s1
    z = lambda a: np.where(gga, a, 0.0)
    sx, v1x, v2x, sc, v1c, v2c = z(sx), z(v1x), z(v2x), z(sc), z(v1c), z(v2c)
    with np.errstate(divide="ignore", invalid="ignore"):
        eps = ex + ec + (sx + sc) / rho
    V = vx + vc + v1x + v1c - _div(pw, (v2x + v2c)[None, :] * grad)
    return eps, V


def calc_E_NN(atoms, Zvals):
    """calc_E_NN.jl:60-172 (Ewald, eta from gcut = 2 and ebsl = 1e-8)."""
    L = atoms.LatVecs
    Om = abs(np.linalg.det(L))
    Rec = 2 * np.pi * np.linalg.inv(L.T)
    gexp = -math.log(1e-8)
    eta = math.sqrt(4.0 / gexp) / 2
    Z = np.asarray(Zvals, dtype=float)[atoms.atm2species]
    tau = atoms.positions
    E = -2 * eta * np.sum(Z ** 2) / math.sqrt(np.pi) - np.pi * np.sum(Z) ** 2 / (Om * eta ** 2)

    def lattice_sum(vecs, cut):
        mm = [int(round(cut / np.linalg.norm(vecs[:, i]) + 1.5)) for i in range(3)]
        I = np.stack(np.meshgrid(*[np.arange(-m, m + 1) for m in mm], indexing="ij")).reshape(3, -1)
        return vecs @ I, np.abs(I).sum(axis=0) != 0

    T, nzT = lattice_sum(L, math.sqrt(0.5 * gexp) / eta)
    for ia in range(atoms.Natoms):
        for ja in range(atoms.Natoms):
            d = (tau[:, ia] - tau[:, ja])[:, None] - T
            r = np.sqrt((d ** 2).sum(axis=0))
            use = nzT if ia == ja else np.ones_like(nzT)
            E += Z[ia] * Z[ja] * np.sum(erfc(eta * r[use]) / r[use])
    Gv, nzG = lattice_sum(Rec, 2.0)
    Gv = Gv[:, nzG]
    G2 = (Gv ** 2).sum(axis=0)
    x = 4 * np.pi / Om * np.exp(-0.25 * G2 / eta ** 2) / G2
    for ia in range(atoms.Natoms):
        for ja in range(atoms.Natoms):
            d = tau[:, ia] - tau[:, ja]
            E += Z[ia] * Z[ja] * np.sum(x * np.cos(d @ Gv))
    return 0.5 * E


def calc_forces_NN(atoms, Zvals):
    """Ion-ion (Ewald) forces F = -dE_NN/dtau, the exact gradient of calc_E_NN above (same
    eta, real-space and reciprocal lattice sums), so that forces and total energy agree to
    the Ewald sums' convergence.  The reference's own calc_forces_NN.jl:15-141 uses another
    splitting parameter; both converge to the same Ewald force."""
    L = atoms.LatVecs
    Om = abs(np.linalg.det(L))
    Rec = 2 * np.pi * np.linalg.inv(L.T)
    gexp = -math.log(1e-8)
    eta = math.sqrt(4.0 / gexp) / 2
    Z = np.asarray(Zvals, dtype=float)[atoms.atm2species]
    tau = atoms.positions
    F = np.zeros((3, atoms.Natoms))

    def lattice_sum(vecs, cut):
        mm = [int(round(cut / np.linalg.norm(vecs[:, i]) + 1.5)) for i in range(3)]
        I = np.stack(np.meshgrid(*[np.arange(-m, m + 1) for m in mm], indexing="ij")).reshape(3, -1)
        return vecs @ I, np.abs(I).sum(axis=0) != 0

    T, _ = lattice_sum(L, math.sqrt(0.5 * gexp) / eta)
    Gv, nzG = lattice_sum(Rec, 2.0)
    Gv = Gv[:, nzG]
    G2 = (Gv ** 2).sum(axis=0)
    x = 4 * np.pi / Om * np.exp(-0.25 * G2 / eta ** 2) / G2
    for ia in range(atoms.Natoms):
        for ja in range(atoms.Natoms):
            if ja != ia:
                d = (tau[:, ia] - tau[:, ja])[:, None] - T
                r = np.sqrt((d ** 2).sum(axis=0))
                # d/dr [erfc(eta r)/r]
                fp = -2 * eta / math.sqrt(np.pi) * np.exp(-(eta * r) ** 2) / r - erfc(eta * r) / r ** 2
                F[:, ia] -= Z[ia] * Z[ja] * (d @ (fp / r))
            dg = tau[:, ia] - tau[:, ja]
            F[:, ia] += Z[ia] * Z[ja] * (Gv @ (x * np.sin(dg @ Gv)))
    return F
