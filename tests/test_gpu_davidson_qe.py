"""diag_davidson_qe! on the device (pwdft_hip/solvers.py: diag_davidson_qe_) against the
oracle's restatement (oracle/solvers.py: diag_davidson_qe) from the same orthonormal start:
the same iteration count and op_H column count (so the same convergence pattern, growth and
refreshes of the search space), the same Ritz values and the same Ritz subspace."""
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


@pytest.mark.parametrize("thr,niter", [(1e-30, 5), (1e-4, 12), (1e-9, 200)])
def test_davidson_qe_vs_oracle(pair, thr, niter):
    """thr 1e-30: no root converges, the basis is refreshed every iteration; 1e-4: roots
    converge one by one and the basis grows by the unconverged ones; 1e-9: to convergence."""
    import pwdft_hip as ph
    from oracle import ops as O
    from oracle.solvers import diag_davidson_qe, ortho_sqrt
    H, oH = pair
    ik = 3
    H.ik = oH.ik = ik
    print(f"case davidson_qe thr={thr} niter={niter}", flush=True)
    X0 = ortho_sqrt(rand_psi(np.random.default_rng(11), H.pw.Ngw[ik], 8))
    X = X0.copy()
    ev0, it0, nc0 = diag_davidson_qe(lambda Y: O.op_H(oH, Y), lambda R: O.Kprec(oH, R), X, EBANDS_THR=thr,
                                     NiterMax=niter)
    d = ph.DeviceMatrix.from_host(H.ctx, X0)
    ev, st = ph.diag_davidson_qe_(H, d, EBANDS_THR=thr, NiterMax=niter)
    assert (st["iter"], st["ncols"]) == (it0, nc0)
    if thr == 1e-30:
        assert it0 == niter and nc0 == 8 * (niter + 1)
    elif thr == 1e-9:
        assert it0 < niter
    assert np.abs(ev - ev0).max() < 1e-9, (ev, ev0)
    Xd = d.download()
    assert np.abs(Xd.conj().T @ Xd - np.eye(8)).max() < 1e-10
    assert np.abs(Xd @ Xd.conj().T - X @ X.conj().T).max() < 1e-7


def test_davidson_qe_blochwavefunc():
    """diag_davidson_qe!(Ham, psiks) over all k-points of Si fcc at the guess density agrees
    with diag_LOBPCG! on the same Hamiltonian."""
    import pwdft_hip as ph
    sym, frac = si_fcc_structure()
    H = product_ham(sym, frac, ph.gen_lattice_fcc(A_SI), "Si-q4.lda.gth", 15.0, meshk=(2, 2, 2), extra_states=4)
    H.update(H.guess_rhoe())
    print("case davidson_qe BlochWavefunc", flush=True)
    ev, stats = ph.diag_davidson_qe(H, ph.rand_BlochWavefunc(H, seed=3), tol=1e-10, NiterMax=100)
    ev_l, _ = ph.diag_LOBPCG(H, ph.rand_BlochWavefunc(H, seed=4), tol=1e-10, NiterMax=300)
    assert all(st["iter"] < 100 for st in stats)
    assert np.abs(ev - ev_l).max() < 1e-8
