"""Per-geometry set-up around the hot path, host-side in FP64: the initial density guess_rhoe
(guess_rhoe.jl:1-38) and the Ewald ion-ion energy E_NN (calc_E_NN.jl:60-172) and forces.
The per-iteration potential update (Poisson, exchange-correlation) runs on the device
(Hamiltonian.update)."""
import math

import numpy as np
from scipy.special import erfc

from .basis import ZATOMS, strfact


def _ifft(pw, f):
    return np.fft.ifftn(np.asarray(f).reshape(pw.Ns, order="F")).reshape(-1, order="F")


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
