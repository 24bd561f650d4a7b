"""Pins of the oracle's force functions (oracle/forces.py) by what the mathematics fixes:
central finite differences of the oracle's own energy terms (h = 1e-4 bohr), and zero
forces on the undisplaced diamond lattice.  A dropped term, a wrong sign or a swapped
index in calc_forces_Ps_loc / calc_forces_Ps_nloc / calc_forces_NN fails one of these.
Also pins the product's host-side Ewald forces (pwdft_hip.potentials.calc_forces_NN)."""
import copy

import numpy as np
import pytest

from _systems import A_SI, PSP, si_fcc_structure

H_FD = 1e-4


def _displaced_si(ecut=8.0, meshk=(1, 1, 1)):
    import os
    from oracle.energy import Hamiltonian
    from oracle.setup import Atoms, PsPot_GTH, gen_lattice_fcc
    sym, frac = si_fcc_structure()
    atoms = Atoms(sym, frac, gen_lattice_fcc(A_SI))
    atoms.positions[:, 0] += [0.07, -0.04, 0.03]
    psp = [PsPot_GTH(os.path.join(PSP, "Si-q4.lda.gth"))]
    return Hamiltonian(atoms, psp, ecut, meshk=meshk, extra_states=2)


def _moved(atoms, ia, d, h):
    a = copy.deepcopy(atoms)
    a.positions[d, ia] += h
    return a


def _E_Ps_loc(ham, atoms, rho):
    from oracle.setup import calc_V_Ps_loc
    V = calc_V_Ps_loc(atoms, ham.pw, ham.pspots)
    return float(np.dot(V, rho) * ham.pw.CellVolume / ham.pw.Npoints)


def test_forces_Ps_loc_fd():
    from oracle.density import guess_rhoe
    from oracle.forces import calc_forces_Ps_loc
    ham = _displaced_si()
    rho = guess_rhoe(ham)
    rho = rho * (1.0 + 0.1 * np.cos(np.arange(rho.size)))  # break any symmetry of the guess
    F = calc_forces_Ps_loc(ham, rho)
    for ia in range(2):
        for d in range(3):
            ep = _E_Ps_loc(ham, _moved(ham.atoms, ia, d, H_FD), rho)
            em = _E_Ps_loc(ham, _moved(ham.atoms, ia, d, -H_FD), rho)
            fd = -(ep - em) / (2 * H_FD)
            assert abs(F[d, ia] - fd) < 1e-7 * max(1.0, abs(fd)), (ia, d, F[d, ia], fd)


def test_forces_Ps_nloc_fd():
    from oracle.energy import calc_E_Ps_nloc
    from oracle.forces import calc_forces_Ps_nloc
    from oracle.setup import PsPotNL
    ham = _displaced_si(meshk=(2, 2, 2))
    rng = np.random.default_rng(3)
    psiks = [rng.standard_normal((n, ham.Nstates)) + 1j * rng.standard_normal((n, ham.Nstates))
             for n in ham.pw.Ngw]
    psiks = [p / np.linalg.norm(p, axis=0) for p in psiks]
    Focc = rng.uniform(0.5, 2.0, size=ham.Focc.shape)
    F = calc_forces_Ps_nloc(ham, psiks, Focc)
    for ia in range(2):
        for d in range(3):
            vals = []
            for h in (H_FD, -H_FD):
                hm = copy.copy(ham)
                hm.pspotNL = PsPotNL(_moved(ham.atoms, ia, d, h), ham.pw, ham.pspots)
                vals.append(calc_E_Ps_nloc(hm, psiks, Focc))
            fd = -(vals[0] - vals[1]) / (2 * H_FD)
            assert abs(F[d, ia] - fd) < 1e-7 * max(1.0, abs(fd)), (ia, d, F[d, ia], fd)


def test_forces_NN_fd_and_ewald_split():
    """The oracle's calc_forces_NN (calc_forces_NN.jl splitting) against central
    differences of the oracle's calc_E_NN (calc_E_NN.jl splitting): two Ewald sums with
    different parameters that converge to the same force."""
    from oracle.energy import calc_E_NN
    from oracle.forces import calc_forces_NN
    ham = _displaced_si()
    F = calc_forces_NN(ham)
    for ia in range(2):
        for d in range(3):
            ep = calc_E_NN(ham.atoms.LatVecs, _moved(ham.atoms, ia, d, H_FD), ham.Zvals)
            em = calc_E_NN(ham.atoms.LatVecs, _moved(ham.atoms, ia, d, -H_FD), ham.Zvals)
            fd = -(ep - em) / (2 * H_FD)
            assert abs(F[d, ia] - fd) < 1e-6, (ia, d, F[d, ia], fd)


def test_product_forces_NN():
    """Product Ewald forces: exact gradient of the product's calc_E_NN, and equal to the
    oracle's calc_forces_NN."""
    import pwdft_hip.potentials as P
    from oracle.forces import calc_forces_NN
    ham = _displaced_si()
    F = P.calc_forces_NN(ham.atoms, ham.Zvals)
    for ia in range(2):
        for d in range(3):
            ep = P.calc_E_NN(_moved(ham.atoms, ia, d, H_FD), ham.Zvals)
            em = P.calc_E_NN(_moved(ham.atoms, ia, d, -H_FD), ham.Zvals)
            assert abs(F[d, ia] + (ep - em) / (2 * H_FD)) < 1e-8
    assert np.abs(F - calc_forces_NN(ham)).max() < 1e-6


def test_zero_forces_undisplaced_diamond():
    """Site symmetry of diamond: no local or ion-ion force on either atom for a density
    with the crystal's symmetry (the atomic-superposition guess)."""
    import os
    from oracle.density import guess_rhoe
    from oracle.energy import Hamiltonian
    from oracle.forces import calc_forces_NN, calc_forces_Ps_loc
    from oracle.setup import Atoms, PsPot_GTH, gen_lattice_fcc
    sym, frac = si_fcc_structure()
    atoms = Atoms(sym, frac, gen_lattice_fcc(A_SI))
    ham = Hamiltonian(atoms, [PsPot_GTH(os.path.join(PSP, "Si-q4.lda.gth"))], 8.0)
    rho = guess_rhoe(ham)
    assert np.abs(calc_forces_Ps_loc(ham, rho)).max() < 1e-10
    assert np.abs(calc_forces_NN(ham)).max() < 1e-8
