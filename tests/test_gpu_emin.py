"""Direct minimisation (KS_solve_Emin_PCG!) on the device and the H atom system of the
reference's validations (test/validations_old/H_atom.jl: H GTH-PADE-q1 in a 16 bohr cube,
15 Ha, one state with occupation 1): calc_grad against the oracle (with occupations that
switch on the occupation-matrix term), the H atom SCF and Emin PCG with each
conjugate-direction formula against the reference log, and Emin PCG on Si fcc against the
reference's test tolerance."""
import numpy as np
import pytest

from _systems import A_SI, oracle_ham, product_ham, rand_psi, rel, si_fcc_structure

pytestmark = pytest.mark.gpu

E_H_LOG = -0.4400340433      # test/LOG_runtest_v.1.1.1:463 (Emin PCG, Fletcher-Reeves)
E_H_ABINIT = -0.4400340349   # test/validations_old/H_atom.jl:15


def _h_atom():
    import pwdft_hip as ph
    return ph.Hamiltonian(ph.Atoms(["H"], [[0, 0, 0]], ph.gen_lattice_sc(16.0)), ["H-q1.lda.gth"], 15.0)


def test_calc_grad_vs_oracle():
    import pwdft_hip as ph
    from oracle import ops as O
    from oracle.solvers import calc_grad
    sym, frac = si_fcc_structure()
    H = product_ham(sym, frac, ph.gen_lattice_fcc(A_SI), "Si-q4.lda.gth", 15.0, meshk=(2, 2, 2), extra_states=4)
    H.update(H.guess_rhoe())
    oH = oracle_ham(sym, frac, ph.gen_lattice_fcc(A_SI), "Si-q4.lda.gth", 15.0, kpoints=(H.pw.k, H.pw.wk),
                    extra_states=4)
    oH.V_loc[:, 0] = H.potentials.Total[:, 0]
    ik = 1
    H.ik = oH.ik = ik
    psi = rand_psi(np.random.default_rng(3), H.pw.Ngw[ik], 8)
    psi = np.linalg.qr(psi)[0]
    for Focc in (np.full(8, 2.0), np.array([2.0, 2.0, 2.0, 1.5, 0.5, 0.0, 0.0, 0.0])):
        d = ph.DeviceMatrix.from_host(H.ctx, psi)
        G = ph.DeviceMatrix(H.ctx, d.rows, 8)
        S = ph.DeviceMatrix(H.ctx, 8, 8)
        ph.calc_grad_(H, d, G, S, Focc, H.pw.wk[ik])
        ref = calc_grad(lambda Y: O.op_H(oH, Y), psi, Focc, H.pw.wk[ik])
        assert rel(G.download(), ref) < 1e-12, Focc


def test_scf_h_atom():
    import pwdft_hip as ph
    H = _h_atom()
    assert (H.electrons.Nstates, H.pspotNL.NbetaNL) == (1, 0)
    out = ph.KS_solve_SCF_(H, betamix=0.7, etot_conv_thr=1e-8)
    assert abs(out["energies"]["Total"] - E_H_LOG) < 1e-6, out["energies"]["Total"]
    assert abs(out["energies"]["Total"] - E_H_ABINIT) < 1e-5


@pytest.mark.parametrize("cg_beta", [1, 2, 3, 4])
def test_emin_pcg_h_atom(cg_beta):
    import pwdft_hip as ph
    print(f"case Emin PCG H atom cg_beta={cg_beta}", flush=True)
    H = _h_atom()
    out = ph.KS_solve_Emin_PCG_(H, i_cg_beta=cg_beta)
    assert out["iterations"] < 200
    assert abs(out["energies"]["Total"] - E_H_LOG) < 2e-6, (out["energies"]["Total"], out["iterations"])


def test_emin_pcg_si_fcc():
    import pwdft_hip as ph
    print("case Emin PCG Si fcc", flush=True)
    sym, frac = si_fcc_structure()
    H = product_ham(sym, frac, ph.gen_lattice_fcc(A_SI), "Si-q4.lda.gth", 15.0, meshk=(3, 3, 3), use_symmetry=True)
    out = ph.KS_solve_Emin_PCG_(H, i_cg_beta=2)
    assert abs(out["energies"]["Total"] - (-7.9110087934)) < 1e-5, (out["energies"]["Total"], out["iterations"])
