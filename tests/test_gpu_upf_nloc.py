"""The nonlocal operator in the form of a norm-conserving UPF Hamiltonian on the device:
dense, coupled nh x nh Deeq blocks per atom installed through pw_set_dij /
pw_set_projectors (Hamiltonian.set_nonlocal_deeq), against the oracle's restatement of
op_V_Ps_nloc.jl:82-126, calc_energies.jl:61-128 and calc_forces_Ps_nloc.jl:31-160
(oracle/nloc_upf.py).  Tolerances as for the GTH operator: FP64 1e-12 relative, mixed
(FP32 MFMA projector products) 2e-5."""
import numpy as np
import pytest

from _systems import A_SI, oracle_ham, rand_psi, rel

pytestmark = pytest.mark.gpu

SHIFT = np.array([[0.041, -0.027, 0.033, -0.012, 0.018, -0.045, 0.006, 0.029],
                  [-0.036, 0.014, 0.022, -0.049, 0.031, 0.008, -0.019, 0.043],
                  [0.025, 0.047, -0.038, 0.011, -0.023, 0.035, -0.044, -0.007]])


@pytest.fixture(scope="module")
def si8_deeq():
    import pwdft_hip as ph
    L = ph.gen_lattice_sc(A_SI)
    frac = np.array(ph.diamond_supercell(1.0, 1), dtype=float)
    frac = np.linalg.solve(L, L @ frac.T + SHIFT).T
    H = ph.Hamiltonian(ph.Atoms(["Si"] * 8, frac, L), ["Si-q4.lda.gth"], 6.0, meshk=(2, 1, 1), Nstates=16,
                       use_symmetry=False)
    oH = oracle_ham(["Si"] * 8, frac, L, "Si-q4.lda.gth", 6.0, kpoints=(H.pw.k, H.pw.wk), Nstates=16)
    nat = 8
    nh = [H.pspotNL.NbetaNL // nat]
    indv = [a * nh[0] for a in range(nat)]
    A = np.random.default_rng(77).standard_normal((nh[0], nh[0], nat))
    Deeq = 0.5 * (A + A.transpose(1, 0, 2))
    H.set_nonlocal_deeq(Deeq, nh, indv)
    rng = np.random.default_rng(78)
    V = oH.V_Ps_loc + 0.2 * rng.standard_normal(oH.pw.Npoints)
    oH.V_loc[:, 0] = V
    H.set_Vloc(V)
    sp = [0] * nat
    return H, oH, Deeq, nh, indv, sp


@pytest.mark.parametrize("mixed", [False, True])
def test_op_V_Ps_nloc_deeq(si8_deeq, mixed):
    import pwdft_hip as ph
    from oracle.nloc_upf import op_V_Ps_nloc_upf
    H, oH, Deeq, nh, indv, sp = si8_deeq
    rng = np.random.default_rng(5)
    H.mixed = mixed
    try:
        for ik in range(H.pw.Nkpt):
            H.ik = oH.ik = ik
            psi = rand_psi(rng, H.pw.Ngw[ik], 16)
            print(f"case UPF-form op_V_Ps_nloc ik={ik} mixed={mixed}", flush=True)
            got = ph.op_V_Ps_nloc(H, psi)
            ref = op_V_Ps_nloc_upf(oH.pspotNL.betaNL[ik], Deeq, nh, indv, sp, psi)
            assert rel(got, ref) < (2e-5 if mixed else 1e-12), rel(got, ref)
    finally:
        H.mixed = False


def test_op_H_deeq_and_lobpcg(si8_deeq):
    """Full op_H with the Deeq nonlocal term, and diag_LOBPCG! on it against dense eigh of
    the oracle's explicit H."""
    import pwdft_hip as ph
    from oracle import ops as O
    from oracle.nloc_upf import op_V_Ps_nloc_upf
    H, oH, Deeq, nh, indv, sp = si8_deeq
    ik = 1
    H.ik = oH.ik = ik
    n = H.pw.Ngw[ik]
    beta = oH.pspotNL.betaNL[ik]
    opH = lambda Y: O.op_K(oH, Y) + O.op_V_loc(oH, Y) + op_V_Ps_nloc_upf(beta, Deeq, nh, indv, sp, Y)
    psi = rand_psi(np.random.default_rng(6), n, 16)
    print("case UPF-form op_H", flush=True)
    assert rel(ph.op_H(H, psi), opH(psi)) < 1e-12
    Hd = opH(np.eye(n, dtype=complex))
    exact = np.linalg.eigvalsh(0.5 * (Hd + Hd.conj().T))[:16]
    X = ph.DeviceMatrix.from_host(H.ctx, np.linalg.qr(rand_psi(np.random.default_rng(7), n, 16))[0])
    print("case UPF-form LOBPCG", flush=True)
    ev, _ = ph.diag_LOBPCG_(H, X, tol=1e-11, NiterMax=400)
    assert np.abs(ev - exact).max() < 1e-8, np.abs(ev - exact).max()


def test_energy_and_forces_deeq(si8_deeq):
    import pwdft_hip as ph
    from oracle.nloc_upf import calc_E_Ps_nloc_upf, calc_forces_Ps_nloc_upf
    H, oH, Deeq, nh, indv, sp = si8_deeq
    rng = np.random.default_rng(9)
    psiks = [np.asfortranarray(np.linalg.qr(rand_psi(rng, H.pw.Ngw[ik], 16))[0]) for ik in range(H.pw.Nkpt)]
    Focc = rng.uniform(0.2, 2.0, size=(16, H.pw.Nkpt))
    print("case UPF-form E_Ps_nloc and forces", flush=True)
    E = ph.scf.calc_E_Ps_nloc(H, psiks, Focc)
    Eo = calc_E_Ps_nloc_upf(oH.pspotNL.betaNL, Deeq, nh, indv, sp, psiks, Focc, oH.pw.wk)
    assert abs(E - Eo) < 1e-12 * max(1.0, abs(Eo)), (E, Eo)
    F = ph.calc_forces_Ps_nloc(H, psiks, Focc)
    Fo = calc_forces_Ps_nloc_upf(oH.pspotNL.betaNL, oH.pw.G, oH.pw.idx_gw2g, Deeq, nh, indv, sp, psiks, Focc,
                                 oH.pw.wk)
    assert np.abs(Fo).max() > 1e-3
    assert np.abs(F - Fo).max() < 1e-11 * np.abs(Fo).max(), np.abs(F - Fo).max()


def test_set_nonlocal_deeq_rejects_bad_blocks(si8_deeq):
    H, _, Deeq, nh, indv, _ = si8_deeq
    bad = Deeq.copy()
    bad[0, 1, 2] += 1.0
    with pytest.raises(ValueError):
        H.set_nonlocal_deeq(bad, nh, indv)
    with pytest.raises(ValueError):
        H.set_nonlocal_deeq(Deeq, nh, [0] * 8)
    with pytest.raises(ValueError):
        H.set_positions(H.atoms.positions)
