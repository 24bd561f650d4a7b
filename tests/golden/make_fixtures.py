"""Writes the frozen fixtures of tests/golden/ from the oracle (oracle/, the CPU restatement
of the reference): small inputs and the oracle's outputs, so that a later change to the
oracle or to the device path is caught against fixed numbers.

  si2_ops.npz    Si fcc (2 atoms, a = 10.2631 bohr, GTH LDA, 15 Ha, 2x2x2 k-mesh, time reversal
                 only) with V_loc = V_Ps_loc + dV (closed form below): 2 seeded bands at k-point
                 1 and their op_H; 1 seeded band at every k-point with seeded occupations, and
                 64 samples of the density calc_rhoe builds from them, with its integral and
                 2-norm.
  si2_scf.json   the same cell with a 3x3x3 mesh: KS_solve_SCF (simple mixing 0.5) total
                 energy and components, converged to 1e-11 Ha.

Run from the repository root:  python tests/golden/make_fixtures.py"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
PSP = os.path.join(ROOT, "pwdft.jl_amd", "pwdft_hip", "pseudopotentials", "Si-q4.lda.gth")
A_SI = 10.2631


def si2(meshk):
    from oracle.energy import Hamiltonian
    from oracle.setup import Atoms, PsPot_GTH, gen_lattice_fcc
    atoms = Atoms(["Si", "Si"], [[0, 0, 0], [0.25, 0.25, 0.25]], gen_lattice_fcc(A_SI))
    return Hamiltonian(atoms, [PsPot_GTH(PSP)], 15.0, meshk=meshk, extra_states=4)


def dV(Ns):
    """The V_loc perturbation of si2_ops (a closed form, so it is not stored): column-major
    grid, 0.1 cos(2 pi i/n1) cos(4 pi j/n2) + 0.05 sin(2 pi k/n3)."""
    n1, n2, n3 = Ns
    i, j, k = np.meshgrid(np.arange(n1), np.arange(n2), np.arange(n3), indexing="ij")
    v = 0.1 * np.cos(2 * np.pi * i / n1) * np.cos(4 * np.pi * j / n2) + 0.05 * np.sin(2 * np.pi * k / n3)
    return v.reshape(-1, order="F")


def ops_fixture():
    from oracle import ops as O
    from oracle.density import calc_rhoe
    H = si2((2, 2, 2))
    rng = np.random.default_rng(2026)
    H.V_loc[:, 0] = H.V_Ps_loc + dV(H.pw.Ns)
    H.ik = 1
    n1 = H.pw.Ngw[1]
    # inputs stored in complex64 (their exact values are what the outputs were computed from)
    psi = (rng.standard_normal((n1, 2)) + 1j * rng.standard_normal((n1, 2))).astype(np.complex64)
    hpsi = O.op_H(H, psi.astype(np.complex128))
    psiks = [(rng.standard_normal((n, 1)) + 1j * rng.standard_normal((n, 1))).astype(np.complex64)
             for n in H.pw.Ngw]
    Focc = rng.uniform(0.5, 2.0, size=(1, H.pw.Nkpt))
    rho = calc_rhoe(H, [p.astype(np.complex128) for p in psiks], Focc)
    idx = rng.choice(H.pw.Npoints, size=64, replace=False)
    out = {"psi": psi, "hpsi": hpsi, "Focc": Focc, "rho_idx": idx, "rho_samples": rho[idx],
           "rho_integral": np.sum(rho) * H.pw.CellVolume / H.pw.Npoints, "rho_norm": np.linalg.norm(rho),
           "Ngw": np.asarray(H.pw.Ngw)}
    for ik, p in enumerate(psiks):
        out[f"psik{ik}"] = p
    np.savez_compressed(os.path.join(HERE, "si2_ops.npz"), **out)


def scf_fixture():
    from oracle.energy import KS_solve_SCF
    H = si2((3, 3, 3))
    e = KS_solve_SCF(H, betamix=0.5, etot_conv_thr=1e-11, ethr_evals_last=1e-13)[0]
    json.dump({k: float(v) for k, v in e.items()}, open(os.path.join(HERE, "si2_scf.json"), "w"), indent=1,
              sort_keys=True)


if __name__ == "__main__":
    ops_fixture()
    if "--ops-only" not in sys.argv:
        scf_fixture()
