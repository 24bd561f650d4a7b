"""Per-geometry set-up built on the device (pw_build_projectors_dev, pw_build_vps_loc_dev)
against the oracle's PsPotNL and calc_V_Ps_loc, including d and f projectors (the CeSi
test compound: Ce GTH-PADE-q12 has l = 0..3), and set_positions against a Hamiltonian
built from scratch at the new positions."""
import os

import numpy as np
import pytest

from _systems import A_SI, PSP, rand_psi, rel

pytestmark = pytest.mark.gpu


def _pair(symbols, frac, L, psps, ecut, meshk):
    import pwdft_hip as ph
    from oracle import setup as S
    H = ph.Hamiltonian(ph.Atoms(symbols, frac, L), psps, ecut, meshk=meshk, use_symmetry=False)
    oat = S.Atoms(symbols, frac, L)
    opw = S.PWGrid(ecut, L, kpoints=(H.pw.k, H.pw.wk))
    opsp = [S.PsPot_GTH(os.path.join(PSP, p)) for p in psps]
    return H, oat, opw, opsp


@pytest.mark.parametrize("name", ["si_fcc", "cesi_fcc", "si8_displaced"])
def test_device_projectors_and_vps_vs_oracle(name):
    import pwdft_hip as ph
    from oracle import setup as S
    print(f"case set-up {name}", flush=True)
    if name == "si_fcc":
        args = (["Si", "Si"], [[0, 0, 0], [0.25, 0.25, 0.25]], ph.gen_lattice_fcc(A_SI), ["Si-q4.lda.gth"], 15.0,
                (2, 2, 2))
    elif name == "cesi_fcc":
        args = (["Ce", "Si"], [[0, 0, 0], [0.25, 0.25, 0.25]], ph.gen_lattice_fcc(A_SI),
                ["Ce-q12.lda.gth", "Si-q4.lda.gth"], 15.0, (2, 2, 2))
    else:
        frac = ph.diamond_supercell(1.0, 1) + np.random.default_rng(3).uniform(-0.005, 0.005, (8, 3))
        args = (["Si"] * 8, frac, ph.gen_lattice_sc(A_SI), ["Si-q4.lda.gth"], 20.0, (1, 1, 1))
    H, oat, opw, opsp = _pair(*args)
    onl = S.PsPotNL(oat, opw, opsp)
    assert H.pspotNL.NbetaNL == onl.NbetaNL
    assert np.array_equal(H.pspotNL.D, onl.D)
    nb = onl.NbetaNL
    for ik in range(H.pw.Nkpt):
        out = ph.DeviceMatrix(H.ctx, nb, H.pw.Ngw[ik])
        # beta^H applied to the identity rows: download beta itself through a product with I
        eye = ph.DeviceMatrix.from_host(H.ctx, np.eye(H.pw.Ngw[ik], dtype=complex))
        H.ctx.call("pw_betaNL_psi_dev", ik, H.pw.Ngw[ik], eye.p, H.pw.Ngw[ik], out.p, nb, 0)
        beta = out.download().conj().T        # (beta^H I)^H = beta
        assert rel(beta, onl.betaNL[ik]) < 1e-13, (ik, rel(beta, onl.betaNL[ik]))
        out.free()
        eye.free()
    V = S.calc_V_Ps_loc(oat, opw, opsp)
    assert rel(H.potentials.Ps_loc, V) < 1e-12, rel(H.potentials.Ps_loc, V)


def test_set_positions_matches_fresh_hamiltonian():
    import pwdft_hip as ph
    frac = ph.diamond_supercell(1.0, 1)
    L = ph.gen_lattice_sc(A_SI)
    H = ph.Hamiltonian(ph.Atoms(["Si"] * 8, frac, L), ["Si-q4.lda.gth"], 15.0, use_symmetry=False)
    rho = H.guess_rhoe()
    H.update(rho)
    new = H.atoms.positions + np.random.default_rng(8).uniform(-0.05, 0.05, (3, 8))
    H.set_positions(new)
    H.update(rho)
    H2 = ph.Hamiltonian(ph.Atoms(["Si"] * 8, new.T, L, frac=False), ["Si-q4.lda.gth"], 15.0, use_symmetry=False)
    H2.update(rho)
    assert np.array_equal(H.potentials.Ps_loc, H2.potentials.Ps_loc)
    psi = rand_psi(np.random.default_rng(2), H.pw.Ngw[0], 16)
    assert np.array_equal(ph.op_H(H, psi), ph.op_H(H2, psi))
    assert H.E_NN == H2.E_NN
