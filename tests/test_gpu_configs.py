"""GPU parity at the sizes the benchmarks run, against the oracle, through the C ABI.

- Si64 H-psi microbenchmark (2x2x2 conventional diamond cells, a = 20.5262 bohr, GTH LDA,
  15 Ha, 72^3 grid, Gamma, Ngw 23 871, Nbeta 320, 160 bands): op_H and calc_rhoe on 8 bands
  against the oracle (FP64 <= 1e-12, mixed <= 2e-5 relative); the mixed op_H over all 160
  bands against the FP64 op_H (band energies within 1e-6 Ha/atom).
- Si8 displaced at its stated size (20 Ha, 45^3 grid, 2x2x2 k-mesh, 16 bands): the local
  and nonlocal forces against the oracle and against central differences (h = 1e-4 bohr)
  of the device energies.
- Si216 PBE (3x3x3 cells, a = 30.789 bohr, atoms moved by up to 0.02 bohr, 108^3 grid,
  Ngw 80 797, Nbeta 1 080): op_H on 4 bands against the oracle; one 480-band LOBPCG
  iteration: orthonormal Ritz vectors, Ritz values equal to the Rayleigh quotients, the
  residual orthogonal to the Ritz vectors, and no Ritz value above its starting Rayleigh
  quotient.
- SCF end points: the FP64 GPU SCF of Si fcc 3x3x3 against the oracle's SCF (1e-10 Ha), and
  the Pt fcc smearing run (d projectors, Fermi level, entropy) against the reference's log.
"""
import os

import numpy as np
import pytest

from _systems import PSP, rand_psi, rel

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOLMIX = 2e-5

SHIFT8 = np.array([[0.041, -0.027, 0.033, -0.012, 0.018, -0.045, 0.006, 0.029],
                   [-0.036, 0.014, 0.022, -0.049, 0.031, 0.008, -0.019, 0.043],
                   [0.025, 0.047, -0.038, 0.011, -0.023, 0.035, -0.044, -0.007]])


def _oracle(symbols, cart, L, psp, ecut, **kw):
    """Oracle Hamiltonian from Cartesian positions given 3 x Natoms (the reference's layout);
    the oracle's Atoms takes one row per atom."""
    from oracle.energy import Hamiltonian
    from oracle.setup import Atoms, PsPot_GTH
    pos = np.asarray(cart, dtype=float)
    assert pos.shape == (3, len(symbols))
    return Hamiltonian(Atoms(symbols, pos.T, L, frac=False), [PsPot_GTH(os.path.join(PSP, psp))], ecut, **kw)


# ----------------------------------------------------------------------------- Si64


@pytest.fixture(scope="module")
def si64():
    import pwdft_hip as ph
    frac = np.array(ph.diamond_supercell(1.0, 2), dtype=float)
    L = ph.gen_lattice_sc(20.5262)
    H = ph.Hamiltonian(ph.Atoms(["Si"] * 64, frac, L), ["Si-q4.lda.gth"], 15.0, Ns_=(72, 72, 72), Nstates=160,
                       band_batch=160, use_symmetry=False)
    oH = _oracle(["Si"] * 64, L @ frac.T, L, "Si-q4.lda.gth", 15.0, Ns=(72, 72, 72), Nstates=160)
    assert H.pw.Ngw[0] == oH.pw.Ngw[0] == 23871 and H.pspotNL.NbetaNL == 320
    rho = H.guess_rhoe()
    H.update(rho)
    oH.update(rho)
    assert rel(H.potentials.Total[:, 0], oH.V_loc[:, 0]) < 1e-12
    return H, oH


@pytest.mark.parametrize("mixed", [False, True])
def test_si64_op_H_vs_oracle(si64, mixed):
    import pwdft_hip as ph
    from oracle import ops as O
    H, oH = si64
    H.ik = oH.ik = 0
    psi = rand_psi(np.random.default_rng(64), H.pw.Ngw[0], 8)
    print(f"case Si64 op_H 8 bands mixed={mixed}", flush=True)
    H.mixed = mixed
    try:
        got = ph.op_H(H, psi)
    finally:
        H.mixed = False
    ref = O.op_H(oH, psi)
    err = rel(got, ref)
    assert err < (TOLMIX if mixed else TOL64), err


@pytest.mark.parametrize("mixed", [False, True])
def test_si64_calc_rhoe_vs_oracle(si64, mixed):
    import pwdft_hip as ph
    from oracle.density import calc_rhoe
    H, oH = si64
    rng = np.random.default_rng(65)
    psi = rand_psi(rng, H.pw.Ngw[0], 8)
    Focc = rng.uniform(0.5, 2.0, size=(8, 1))
    print(f"case Si64 calc_rhoe 8 bands mixed={mixed}", flush=True)
    H.mixed = mixed
    try:
        got = ph.calc_rhoe(H, [psi], Focc=Focc)[:, 0]
    finally:
        H.mixed = False
    ref = calc_rhoe(oH, [psi], Focc)
    err = rel(got, ref)
    assert err < (TOLMIX if mixed else TOL64), err


def test_si64_mixed_vs_fp64_all_bands(si64):
    """All 160 Loewdin-orthonormal bands (the benchmark's input): mixed op_H against FP64
    op_H, and the band-energy sum sum_b <psi_b|H|psi_b> within 1e-6 Ha/atom."""
    import pwdft_hip as ph
    from pwdft_hip._lib import PW_MIXED, PW_TERM_ALL
    H, _ = si64
    H.ik = 0
    n = H.pw.Ngw[0]
    rng = np.random.default_rng(20260101)
    X = ph.DeviceMatrix.from_host(H.ctx, rng.standard_normal((n, 160)) + 1j * rng.standard_normal((n, 160)))
    ph.ortho_sqrt_(H, X)
    psi = X.download()
    assert np.abs(psi.conj().T @ psi - np.eye(160)).max() < 1e-12
    out = {}
    for mixed in (False, True):
        Y = ph.DeviceMatrix(H.ctx, n, 160)
        print(f"case Si64 op_H 160 bands mixed={mixed}", flush=True)
        H.ctx.call("pw_op_H_dev", 0, 0, 160, X.p, n, Y.p, n, PW_TERM_ALL | (PW_MIXED if mixed else 0))
        out[mixed] = Y.download()
    assert rel(out[True], out[False]) < TOLMIX
    eb = {m: float(np.real(np.sum(psi.conj() * out[m]))) for m in out}
    assert abs(eb[True] - eb[False]) / 64 < 1e-6, eb


# ----------------------------------------------------------------------------- Si8, 20 Ha


@pytest.fixture(scope="module")
def si8_20():
    import pwdft_hip as ph
    L = ph.gen_lattice_sc(10.2631)
    frac = np.array(ph.diamond_supercell(1.0, 1), dtype=float)
    cart = L @ frac.T + SHIFT8
    frac = np.linalg.solve(L, cart).T
    H = ph.Hamiltonian(ph.Atoms(["Si"] * 8, frac, L), ["Si-q4.lda.gth"], 20.0, meshk=(2, 2, 2), Nstates=16,
                       use_symmetry=False)
    assert H.pw.Ns == (45, 45, 45) and H.pw.Nkpt == 8
    oH = _oracle(["Si"] * 8, cart, L, "Si-q4.lda.gth", 20.0, kpoints=(H.pw.k, H.pw.wk), Nstates=16)
    rng = np.random.default_rng(45)
    psiks = [np.asfortranarray(np.linalg.qr(rand_psi(rng, n, 16))[0]) for n in H.pw.Ngw]
    Focc = rng.uniform(0.2, 2.0, size=(16, H.pw.Nkpt))
    rho = H.guess_rhoe()
    rho = rho * (1.0 + 0.05 * np.cos(0.37 * np.arange(rho.size)))
    return H, oH, psiks, Focc, rho


def test_si8_20Ha_forces_vs_oracle(si8_20):
    import pwdft_hip as ph
    from oracle.forces import calc_forces_Ps_loc, calc_forces_Ps_nloc
    H, oH, psiks, Focc, rho = si8_20
    print("case Si8 20 Ha 45^3 forces vs oracle", flush=True)
    Fl, Flo = ph.calc_forces_Ps_loc(H, rho), calc_forces_Ps_loc(oH, rho)
    Fn, Fno = ph.calc_forces_Ps_nloc(H, psiks, Focc), calc_forces_Ps_nloc(oH, psiks, Focc)
    assert np.abs(Flo).max() > 1e-3 and np.abs(Fno).max() > 1e-3
    assert np.abs(Fl - Flo).max() < 1e-11 * max(1.0, np.abs(Flo).max()), np.abs(Fl - Flo).max()
    assert np.abs(Fn - Fno).max() < 1e-11 * max(1.0, np.abs(Fno).max()), np.abs(Fn - Fno).max()


def test_si8_20Ha_forces_fd(si8_20):
    """Central differences (h = 1e-4 bohr) of the device E_Ps_loc (rho fixed) and E_Ps_nloc
    (psi fixed, projectors rebuilt on the device at each displaced geometry)."""
    import pwdft_hip as ph
    H, _, psiks, Focc, rho = si8_20
    print("case Si8 20 Ha 45^3 forces finite differences", flush=True)
    Fl = ph.calc_forces_Ps_loc(H, rho)
    Fn = ph.calc_forces_Ps_nloc(H, psiks, Focc)
    pos0 = H.atoms.positions.copy()
    h = 1e-4
    try:
        for ia, d in ((1, 0), (4, 1), (7, 2)):
            el, en = [], []
            for s in (h, -h):
                pos = pos0.copy()
                pos[d, ia] += s
                H.set_positions(pos)
                el.append(ph.E_Ps_loc_grad(H, rho)[0])
                en.append(ph.rrule_E_Ps_nloc(H, psiks, Focc)[0])
            fl = -(el[0] - el[1]) / (2 * h)
            fn = -(en[0] - en[1]) / (2 * h)
            assert abs(fl - Fl[d, ia]) < 2e-8 * max(1.0, abs(fl)), (ia, d, fl, Fl[d, ia])
            assert abs(fn - Fn[d, ia]) < 2e-8 * max(1.0, abs(fn)), (ia, d, fn, Fn[d, ia])
    finally:
        H.set_positions(pos0)


# ----------------------------------------------------------------------------- Si216 PBE


@pytest.fixture(scope="module")
def si216():
    import pwdft_hip as ph
    frac = np.array(ph.diamond_supercell(1.0, 3), dtype=float)
    L = ph.gen_lattice_sc(30.789)
    cart = L @ frac.T + np.random.default_rng(216).uniform(-0.02, 0.02, size=(3, len(frac)))
    print("case Si216 PBE set-up", flush=True)
    H = ph.Hamiltonian(ph.Atoms(["Si"] * 216, np.linalg.solve(L, cart).T, L), ["Si-q4.pbe.gth"], 15.0,
                       Ns_=(108, 108, 108), Nstates=480, xcfunc="PBE", use_symmetry=False, band_batch=96)
    assert H.pw.Ngw[0] == 80797 and H.pspotNL.NbetaNL == 1080
    H.update(H.guess_rhoe())
    return H, cart, L


@pytest.mark.parametrize("mixed", [False, True])
def test_si216_op_H_vs_oracle(si216, mixed):
    import pwdft_hip as ph
    from oracle import ops as O
    H, cart, L = si216
    oH = _oracle(["Si"] * 216, cart, L, "Si-q4.pbe.gth", 15.0, Ns=(108, 108, 108), Nstates=480, xcfunc="PBE")
    oH.V_loc[:, 0] = H.potentials.Total[:, 0]
    H.ik = oH.ik = 0
    psi = rand_psi(np.random.default_rng(217), H.pw.Ngw[0], 4)
    print(f"case Si216 op_H 4 bands mixed={mixed}", flush=True)
    H.mixed = mixed
    try:
        got = ph.op_H(H, psi)
    finally:
        H.mixed = False
    ref = O.op_H(oH, psi)
    err = rel(got, ref)
    assert err < (TOLMIX if mixed else TOL64), err


def test_si216_lobpcg_iteration(si216):
    """One diag_LOBPCG! iteration on all 480 bands (Rayleigh-Ritz on [X W] of size 960)."""
    import pwdft_hip as ph
    from pwdft_hip._lib import PW_TERM_ALL
    H, _, _ = si216
    H.ik = 0
    n = H.pw.Ngw[0]
    rng = np.random.default_rng(480)
    X = ph.DeviceMatrix.from_host(H.ctx, rng.standard_normal((n, 480)) + 1j * rng.standard_normal((n, 480)))
    ph.ortho_sqrt_(H, X)
    x0 = X.download()
    HX = ph.DeviceMatrix(H.ctx, n, 480)
    H.ctx.call("pw_op_H_dev", 0, 0, 480, X.p, n, HX.p, n, PW_TERM_ALL)
    rq0 = np.sort(np.real(np.einsum("ij,ij->j", x0.conj(), HX.download())))
    print("case Si216 one LOBPCG iteration, 480 bands", flush=True)
    ev, st = ph.diag_LOBPCG_(H, X, tol=0.0, NiterMax=1)
    x = X.download()
    H.ctx.call("pw_op_H_dev", 0, 0, 480, X.p, n, HX.p, n, PW_TERM_ALL)
    hx = HX.download()
    assert np.abs(x.conj().T @ x - np.eye(480)).max() < 1e-10
    xhx = x.conj().T @ hx
    assert np.abs(np.real(np.diag(xhx)) - ev).max() < 1e-9 * max(1.0, np.abs(ev).max())
    assert np.abs(xhx - np.diag(np.diag(xhx))).max() < 1e-9 * max(1.0, np.abs(ev).max())
    assert np.all(np.diff(ev) >= -1e-12)
    assert np.all(ev <= rq0 + 1e-9), (ev - rq0).max()
    assert ev.sum() < rq0.sum() - 1.0


# ----------------------------------------------------------------------------- SCF end points


def test_scf_si_fcc_fp64_matches_oracle_scf():
    """Drop-in FP64 parity of the converged SCF: Si fcc, GTH LDA, 15 Ha, 3x3x3 (14 k-points
    by time reversal), LOBPCG, simple mixing beta 0.5.  The GPU SCF and the oracle's
    KS_solve_SCF, each run until two successive energies differ by < 1e-12 with the
    eigensolver tolerance floor at 1e-13, end on total energies within 1e-10 Ha.  (Both
    also land on ABINIT's -7.9110087934 of test/validations_old/Si_fcc.jl:15.)"""
    import pwdft_hip as ph
    from oracle.energy import KS_solve_SCF
    from oracle.setup import gen_lattice_fcc
    from _systems import A_SI, oracle_ham, product_ham, si_fcc_structure
    sym, frac = si_fcc_structure()
    oH = oracle_ham(sym, frac, gen_lattice_fcc(A_SI), "Si-q4.lda.gth", 15.0, meshk=(3, 3, 3), extra_states=4)
    H = product_ham(sym, frac, ph.gen_lattice_fcc(A_SI), "Si-q4.lda.gth", 15.0, meshk=(3, 3, 3), extra_states=4)
    assert H.pw.Nkpt == oH.pw.Nkpt == 14
    print("case Si fcc SCF FP64 vs oracle SCF (oracle)", flush=True)
    eo = KS_solve_SCF(oH, betamix=0.5, etot_conv_thr=1e-12, ethr_evals_last=1e-13)[0]
    print(f"  oracle {eo['Total']:.12f}; GPU", flush=True)
    out = ph.KS_solve_SCF_(H, betamix=0.5, etot_conv_thr=1e-12, ethr_evals_last=1e-13, NiterMax=150)
    eg = out["energies"]
    print(f"  GPU {eg['Total']:.12f} after {out['iterations']} iterations", flush=True)
    assert abs(eg["Total"] - eo["Total"]) < 1e-10, (eg["Total"], eo["Total"])
    assert abs(eg["Total"] - (-7.9110087934)) < 1e-8


def test_scf_pt_fcc_smearing_matches_reference_log():
    """The reference's Pt fcc smearing run (test/LOG_runtest_v.1.1.1:2630-2663,
    test/validations_old/Pt_fcc.jl:1-22): a = 3.9231 angstrom, GTH-PADE-q10 (s, p, d
    channels with two projectors each), 30 Ha, 3x3x3 reduced by symmetry, Davidson, simple
    mixing beta 0.5, Fermi-Dirac kT = 1e-3.  The log stopped at |dE| 2.6e-7: free energy
    -26.2945052148 (ABINIT -26.2945054360), -TS -0.0005158875, E_F -0.0148842322 in the
    log's convention, which left the G = 0 local-potential constant mean(V_Ps_loc) out of
    the eigenvalues (its separate PspCore term)."""
    import pwdft_hip as ph
    a = 3.9231 * 1.8897261254578281
    H = ph.Hamiltonian(ph.Atoms(["Pt"], [[0, 0, 0]], ph.gen_lattice_fcc(a)), ["Pt-q10.lda.gth"], 30.0,
                       meshk=(3, 3, 3), extra_states=4)
    assert H.pspotNL.NbetaNL == 18 and H.electrons.Nstates == 9 and H.pw.Nkpt == 4
    print("case Pt fcc SCF with smearing", flush=True)
    out = ph.KS_solve_SCF_(H, betamix=0.5, etot_conv_thr=1e-8, ethr_evals_last=1e-10, use_smearing=True,
                           kT=1e-3, update_psi="davidson", NiterMax=150)
    e = out["energies"]
    shift = float(np.mean(H.potentials.Ps_loc))
    print(f"  free energy {e['Total']:.10f}  -TS {e['mTS']:.10f}  E_F {H.electrons.E_fermi - shift:.10f}",
          flush=True)
    assert abs(e["Total"] - (-26.2945052148)) < 1e-6, e
    assert abs(e["Total"] - (-26.2945054360)) < 1e-6, e
    assert abs(e["mTS"] - (-0.0005158875)) < 1e-7, e
    assert abs(shift * 10 - 6.3258423493) < 1e-8, shift
    assert abs(H.electrons.E_fermi - shift - (-0.0148842322)) < 1e-5, H.electrons.E_fermi
