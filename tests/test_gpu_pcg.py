"""diag_Emin_PCG! on the device (pwdft_hip/pcg.py) against the oracle's restatement
(oracle/solvers.py: diag_Emin_PCG) from the same orthonormal start with each
conjugate-direction formula, and the SCF with update_psi = "PCG" against the reference
log's Si fcc energy."""
import numpy as np
import pytest

from _systems import A_SI, oracle_ham, product_ham, rand_psi, si_fcc_structure

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pair():
    import pwdft_hip as ph
    sym, frac = si_fcc_structure()
    H = product_ham(sym, frac, ph.gen_lattice_fcc(A_SI), "Si-q4.lda.gth", 15.0, meshk=(3, 3, 3), extra_states=4)
    H.update(H.guess_rhoe())
    oH = oracle_ham(sym, frac, ph.gen_lattice_fcc(A_SI), "Si-q4.lda.gth", 15.0, kpoints=(H.pw.k, H.pw.wk),
                    extra_states=4)
    oH.V_loc[:, 0] = H.potentials.Total[:, 0]
    return H, oH


@pytest.mark.parametrize("cg_beta", [1, 2, 3, 4])
def test_pcg_vs_oracle(pair, cg_beta):
    import pwdft_hip as ph
    from oracle import ops as O
    from oracle.solvers import diag_Emin_PCG, ortho_sqrt
    H, oH = pair
    ik = 2
    H.ik = oH.ik = ik
    print(f"case PCG cg_beta={cg_beta}", flush=True)
    X0 = ortho_sqrt(rand_psi(np.random.default_rng(cg_beta), H.pw.Ngw[ik], 8))
    X = X0.copy()
    ev0, it0, _ = diag_Emin_PCG(lambda Y: O.op_H(oH, Y), lambda R: O.Kprec(oH, R), X, tol=0.0, NiterMax=12,
                                tol_ebands=0.0, I_CG_BETA=cg_beta)
    d = ph.DeviceMatrix.from_host(H.ctx, X0)
    ev, st = ph.diag_Emin_PCG_(H, d, tol=0.0, NiterMax=12, tol_ebands=0.0, I_CG_BETA=cg_beta)
    assert st["iterations"] == it0 == 12
    assert np.abs(ev - ev0).max() < 1e-9, (ev, ev0)
    Xd = d.download()
    assert np.abs(Xd.conj().T @ Xd - np.eye(8)).max() < 1e-12
    # same invariant subspace: the projectors onto the 8 vectors agree
    assert np.abs(Xd @ Xd.conj().T - X @ X.conj().T).max() < 1e-8


def test_scf_pcg_si_fcc():
    import pwdft_hip as ph
    print("case SCF Si fcc PCG", flush=True)
    sym, frac = si_fcc_structure()
    # init_Ham_Si_fcc (test/validations_old/Si_fcc.jl:1-13): the default band count (4,
    # all occupied); PCG's early exit on the band-energy change (tol_ebands) leaves empty
    # bands poorly converged, which the default count avoids
    H = product_ham(sym, frac, ph.gen_lattice_fcc(A_SI), "Si-q4.lda.gth", 15.0, meshk=(3, 3, 3), use_symmetry=True)
    out = ph.KS_solve_SCF_(H, betamix=0.5, update_psi="PCG", etot_conv_thr=1e-7, NiterMax=100)
    assert out["iterations"] < 100
    assert abs(out["energies"]["Total"] - (-7.9110091083)) < 5e-6, (out["energies"]["Total"], out["iterations"])
