"""Pins of the oracle's norm-conserving UPF nonlocal terms (oracle/nloc_upf.py):

* with each atom's Deeq block set to its GTH h-matrix blocks, the dense-block operator,
  energy and forces equal the oracle's GTH restatements, which are written as separate
  scalar loops over (l, m, i, j) (op_V_Ps_nloc.jl:39-78, calc_energies.jl:7-58,
  calc_forces_Ps_nloc.jl:291-343); the GTH force uses G + k and the UPF one G, so this also
  checks that the k part of the projector derivative cancels for a real symmetric Deeq;
* for random dense symmetric Deeq blocks, the forces are minus the central finite
  differences of the energy when an atom's projectors are translated
  (beta_a(tau + h e) = beta_a(tau) exp(-i (G + k).h e), PsPotNL.jl:63-64);
* the operator is Hermitian and equals beta D beta' with D block-diagonal."""
import numpy as np
import pytest

from _systems import A_SI, oracle_ham, rand_psi
from oracle import nloc_upf as U
from oracle import ops as O
from oracle.energy import calc_E_Ps_nloc
from oracle.forces import calc_forces_Ps_nloc
from oracle.setup import gen_lattice_sc

SHIFT = np.array([[0.041, -0.027, 0.033, -0.012, 0.018, -0.045, 0.006, 0.029],
                  [-0.036, 0.014, 0.022, -0.049, 0.031, 0.008, -0.019, 0.043],
                  [0.025, 0.047, -0.038, 0.011, -0.023, 0.035, -0.044, -0.007]])


@pytest.fixture(scope="module")
def si8():
    L = gen_lattice_sc(A_SI)
    base = np.array([[0, 0, 0], [0.25, 0.25, 0.25], [0, 0.5, 0.5], [0.25, 0.75, 0.75],
                     [0.5, 0, 0.5], [0.75, 0.25, 0.75], [0.5, 0.5, 0], [0.75, 0.75, 0.25]], dtype=float)
    frac = np.linalg.solve(L, L @ base.T + SHIFT).T
    oH = oracle_ham(["Si"] * 8, frac, L, "Si-q4.lda.gth", 6.0, meshk=(2, 1, 1), Nstates=16)
    rng = np.random.default_rng(41)
    psiks = [np.linalg.qr(rand_psi(rng, oH.pw.Ngw[ik], 16))[0] for ik in range(oH.pw.Nkpt)]
    Focc = rng.uniform(0.2, 2.0, size=(16, oH.pw.Nkpt))
    return oH, psiks, Focc


def _layout(oH):
    nl = oH.pspotNL
    nat = oH.atoms.Natoms
    nh = [nl.NbetaNL // nat]
    indv = [a * nh[0] for a in range(nat)]
    return nh, indv, list(oH.atoms.atm2species)


def _gth_deeq(oH):
    nh, indv, sp = _layout(oH)
    n = nh[0]
    return np.stack([oH.pspotNL.D[o:o + n, o:o + n] for o in indv], axis=2)


def _random_deeq(oH, seed):
    nh, indv, _ = _layout(oH)
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((nh[0], nh[0], len(indv)))
    return 0.5 * (A + A.transpose(1, 0, 2))


def test_gth_blocks_reduce_to_gth_terms(si8):
    oH, psiks, Focc = si8
    nh, indv, sp = _layout(oH)
    Deeq = _gth_deeq(oH)
    for ik in range(oH.pw.Nkpt):
        oH.ik = ik
        ref = O.op_V_Ps_nloc(oH, psiks[ik])
        got = U.op_V_Ps_nloc_upf(oH.pspotNL.betaNL[ik], Deeq, nh, indv, sp, psiks[ik])
        assert np.abs(got - ref).max() < 1e-13 * np.abs(ref).max()
    E = U.calc_E_Ps_nloc_upf(oH.pspotNL.betaNL, Deeq, nh, indv, sp, psiks, Focc, oH.pw.wk)
    assert abs(E - calc_E_Ps_nloc(oH, psiks, Focc)) < 1e-12 * abs(E)
    F = U.calc_forces_Ps_nloc_upf(oH.pspotNL.betaNL, oH.pw.G, oH.pw.idx_gw2g, Deeq, nh, indv, sp, psiks, Focc,
                                  oH.pw.wk)
    Fg = calc_forces_Ps_nloc(oH, psiks, Focc)
    assert np.abs(Fg).max() > 1e-3
    assert np.abs(F - Fg).max() < 1e-11 * np.abs(Fg).max()


def test_dense_blocks_operator(si8):
    oH, psiks, _ = si8
    nh, indv, sp = _layout(oH)
    Deeq = _random_deeq(oH, 5)
    D = U.assemble_D(Deeq, nh, indv, sp, oH.pspotNL.NbetaNL)
    rng = np.random.default_rng(8)
    for ik in range(oH.pw.Nkpt):
        beta = oH.pspotNL.betaNL[ik]
        x = rand_psi(rng, oH.pw.Ngw[ik], 3)
        y = rand_psi(rng, oH.pw.Ngw[ik], 3)
        Vx = U.op_V_Ps_nloc_upf(beta, Deeq, nh, indv, sp, x)
        Vy = U.op_V_Ps_nloc_upf(beta, Deeq, nh, indv, sp, y)
        assert np.abs(y.conj().T @ Vx - (Vy.conj().T @ x)).max() < 1e-12 * np.abs(Vx).max() * x.shape[0]
        assert np.abs(Vx - beta @ (D @ (beta.conj().T @ x))).max() < 1e-13 * np.abs(Vx).max()


def test_dense_blocks_forces_fd(si8):
    oH, psiks, Focc = si8
    nh, indv, sp = _layout(oH)
    Deeq = _random_deeq(oH, 6)
    pw = oH.pw
    F = U.calc_forces_Ps_nloc_upf(oH.pspotNL.betaNL, pw.G, pw.idx_gw2g, Deeq, nh, indv, sp, psiks, Focc, pw.wk)

    def energy(ia, ipol, h):
        betas = []
        for ik in range(pw.Nkpt):
            b = oH.pspotNL.betaNL[ik].copy()
            r = slice(indv[ia], indv[ia] + nh[0])
            b[:, r] *= np.exp(-1j * h * pw.gk(ik)[ipol])[:, None]
            betas.append(b)
        return U.calc_E_Ps_nloc_upf(betas, Deeq, nh, indv, sp, psiks, Focc, pw.wk)

    h = 1e-4
    for ia, ipol in [(0, 0), (3, 1), (6, 2)]:
        fd = -(energy(ia, ipol, h) - energy(ia, ipol, -h)) / (2 * h)
        assert abs(fd - F[ipol, ia]) < 1e-7 * max(1.0, np.abs(F).max()), (ia, ipol, fd, F[ipol, ia])
