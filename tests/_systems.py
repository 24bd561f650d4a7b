"""Paired test systems: the same structure built by the product (pwdft_hip) and by the
oracle, with the GTH parameter files shipped in the package."""
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PSP = os.path.join(ROOT, "pwdft.jl_amd", "pwdft_hip", "pseudopotentials")

A_SI = 10.2631


def si_fcc_structure():
    return ["Si", "Si"], [[0, 0, 0], [0.25, 0.25, 0.25]]


def oracle_ham(symbols, frac, LatVecs, pspfile, ecut, **kw):
    from oracle.energy import Hamiltonian
    from oracle.setup import Atoms, PsPot_GTH
    atoms = Atoms(symbols, frac, LatVecs)
    return Hamiltonian(atoms, [PsPot_GTH(os.path.join(PSP, pspfile))], ecut, **kw)


def product_ham(symbols, frac, LatVecs, pspfile, ecut, **kw):
    """Product Hamiltonian; k-points reduced by time reversal only (as the oracle's
    default) unless use_symmetry is passed."""
    import pwdft_hip as ph
    atoms = ph.Atoms(symbols, frac, LatVecs)
    kw.setdefault("use_symmetry", False)
    return ph.Hamiltonian(atoms, [pspfile], ecut, **kw)


def rand_psi(rng, ngw, nst):
    return np.asfortranarray(rng.standard_normal((ngw, nst)) + 1j * rng.standard_normal((ngw, nst)))


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))
