"""diag_CheFSI! on the device (pwdft_hip/chefsi.py) against the oracle's restatement
(oracle/solvers.py: chebyfilt, get_ub_lb_lanczos, diag_CheFSI) on the same Hamiltonian,
block and Lanczos start vector, and the SCF with update_psi = "CheFSI" against the
reference's Si fcc validation (test/validations_old/Si_fcc.jl:82-91: rpulay, betamix 0.2,
|E - E_ABINIT| < 1e-5)."""
import numpy as np
import pytest

from _systems import A_SI, oracle_ham, product_ham, rand_psi, rel, si_fcc_structure

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


def _opH(oH, ik):
    from oracle import ops as O

    def f(Z):
        oH.ik = ik
        return O.op_H(oH, Z)
    return f


@pytest.mark.parametrize("degree", [1, 4, 8])
def test_chebyfilt_vs_oracle(pair, degree):
    import pwdft_hip as ph
    from oracle.solvers import chebyfilt
    H, oH = pair
    ik = 2
    H.ik = ik
    X = rand_psi(np.random.default_rng(degree), H.pw.Ngw[ik], 8)
    lb, ub = 0.3, 12.0
    d = ph.DeviceMatrix.from_host(H.ctx, X)
    ph.chebyfilt(H, d, degree, lb, ub)
    ref = chebyfilt(_opH(oH, ik), X, degree, lb, ub)
    assert rel(d.download(), ref) < 1e-12, rel(d.download(), ref)


def test_lanczos_bounds_vs_oracle(pair):
    import pwdft_hip as ph
    from oracle.solvers import get_ub_lb_lanczos
    H, oH = pair
    ik = 1
    H.ik = ik
    rng = np.random.default_rng(3)
    v0 = rng.standard_normal(H.pw.Ngw[ik]) + 1j * rng.standard_normal(H.pw.Ngw[ik])
    lb, ub = ph.get_ub_lb_lanczos(H, 16, v0)
    lb0, ub0 = get_ub_lb_lanczos(_opH(oH, ik), v0, 16)
    assert abs(lb - lb0) < 1e-10 * abs(lb0) and abs(ub - ub0) < 1e-10 * abs(ub0), (lb, lb0, ub, ub0)


def test_diag_chefsi_vs_oracle(pair):
    import pwdft_hip as ph
    from oracle.solvers import diag_CheFSI
    H, oH = pair
    rng = np.random.default_rng(11)
    host = [rand_psi(rng, n, 8) for n in H.pw.Ngw]
    v0s = [rng.standard_normal(n) + 1j * rng.standard_normal(n) for n in H.pw.Ngw]
    psiks = [ph.DeviceMatrix.from_host(H.ctx, a) for a in host]
    ev, st = ph.diag_CheFSI_(H, psiks, 8, v0s=v0s)
    for ik in range(H.pw.Nkpt):
        ev0, Y = diag_CheFSI(_opH(oH, ik), host[ik], 8, v0s[ik])
        assert np.abs(ev[:, ik] - ev0).max() < 1e-10, ik
        assert rel(psiks[ik].download(), Y) < 1e-10, ik
    assert st["ncols"] == H.pw.Nkpt * (16 + 8 * 8 + 8)


def test_scf_chefsi_si_fcc():
    import pwdft_hip as ph
    print("case SCF Si fcc CheFSI rpulay", flush=True)
    sym, frac = si_fcc_structure()
    # init_Ham_Si_fcc (Si_fcc.jl:1-13): default band count, Nstates = 4, all occupied -- the
    # density of a filtered block that is not rotated to Ritz vectors is then exact
    H = product_ham(sym, frac, ph.gen_lattice_fcc(A_SI), "Si-q4.lda.gth", 15.0, meshk=(3, 3, 3), use_symmetry=True)
    assert H.electrons.Nstates == 4
    out = ph.KS_solve_SCF_(H, betamix=0.2, mix_method="rpulay", update_psi="CheFSI", NiterMax=150)
    E = out["energies"]["Total"]
    assert abs(E - (-7.9110087934)) < 1e-5, (E, out["iterations"])
