"""GPU forces and energy gradients through the C ABI: parity with the oracle's
calc_forces_Ps_loc / calc_forces_Ps_nloc on the same rho and psi, central finite
differences of the device energies, and the self-consistent total force against central
differences of the self-consistent total energy (FP64 and mixed precision)."""
import numpy as np
import pytest

from _systems import A_SI, oracle_ham, rand_psi

pytestmark = pytest.mark.gpu


SHIFT = np.array([[0.041, -0.027, 0.033, -0.012, 0.018, -0.045, 0.006, 0.029],
                  [-0.036, 0.014, 0.022, -0.049, 0.031, 0.008, -0.019, 0.043],
                  [0.025, 0.047, -0.038, 0.011, -0.023, 0.035, -0.044, -0.007]])


def _si8(ecut=10.0, meshk=(2, 2, 2), Nstates=16, **kw):
    """Si8 cubic cell (a = 10.2631 bohr), every atom moved by a fixed offset <= 0.05 bohr."""
    import pwdft_hip as ph
    L = ph.gen_lattice_sc(A_SI)
    frac = np.array(ph.diamond_supercell(1.0, 1), dtype=float)
    cart = L @ frac.T + SHIFT
    frac = np.linalg.solve(L, cart).T
    atoms = ph.Atoms(["Si"] * 8, frac, L)
    H = ph.Hamiltonian(atoms, ["Si-q4.lda.gth"], ecut, meshk=meshk, Nstates=Nstates, use_symmetry=False, **kw)
    return H, frac, L


def _oracle_for(H, frac, L):
    return oracle_ham(["Si"] * 8, frac, L, "Si-q4.lda.gth", H.pw.ecutwfc, kpoints=(H.pw.k, H.pw.wk),
                      Nstates=16)


def _orthonormal_psiks(H, seed):
    rng = np.random.default_rng(seed)
    out = []
    for ik in range(H.pw.Nkpt):
        q, _ = np.linalg.qr(rand_psi(rng, H.pw.Ngw[ik], H.electrons.Nstates))
        out.append(np.asfortranarray(q))
    return out


@pytest.fixture(scope="module")
def si8():
    H, frac, L = _si8()
    oH = _oracle_for(H, frac, L)
    rho = H.guess_rhoe()
    rho = rho * (1.0 + 0.05 * np.cos(0.37 * np.arange(rho.size)))
    psiks = _orthonormal_psiks(H, 11)
    Focc = np.random.default_rng(12).uniform(0.2, 2.0, size=(16, H.pw.Nkpt))
    return H, oH, rho, psiks, Focc


def test_forces_Ps_loc_vs_oracle(si8):
    import pwdft_hip as ph
    from oracle.forces import calc_forces_Ps_loc
    H, oH, rho, _, _ = si8
    print("case Si8 forces_Ps_loc", flush=True)
    F = ph.calc_forces_Ps_loc(H, rho)
    Fo = calc_forces_Ps_loc(oH, rho)
    assert np.abs(Fo).max() > 1e-3
    assert np.abs(F - Fo).max() < 1e-11 * max(1.0, np.abs(Fo).max()), np.abs(F - Fo).max()


def test_E_Ps_loc_grad_vs_oracle(si8):
    """E_Ps_loc on the rho sphere equals the reference's real-space <V_Ps_loc, rho> dV, and
    dE/dV_sp(G) summed on a G shell equals the change of that real-space energy per unit
    change of the shell's form factor (exact: the energy is linear in V_sp)."""
    import pwdft_hip as ph
    from oracle.energy import calc_E_local
    H, oH, rho, _, _ = si8
    print("case Si8 E_Ps_loc grad", flush=True)
    E, dvg, dtau = ph.E_Ps_loc_grad(H, rho)
    oH.update(rho)
    Eo = calc_E_local(oH, rho)[0]
    assert abs(E - Eo) < 1e-10 * max(1.0, abs(Eo)), (E, Eo)
    assert np.abs(dtau + ph.calc_forces_Ps_loc(H, rho)).max() < 1e-14
    pw = H.pw
    Sf = ph.basis.strfact(H.atoms.positions, H.atoms.atm2species, 1, pw.G)[:, 0]
    for ish in (1, 2, 5, 17):
        sel = pw.idx_g2shells == ish
        Vg = np.zeros(pw.Npoints, dtype=complex)
        Vg[pw.idx_g2r[sel]] = Sf[sel] / pw.CellVolume
        dV = np.real(np.fft.ifftn(Vg.reshape(pw.Ns, order="F")).reshape(-1, order="F")) * pw.Npoints
        dE = float(np.dot(dV, rho) * pw.CellVolume / pw.Npoints)
        assert abs(dvg[sel, 0].sum() - dE) < 1e-10 * max(1.0, abs(dE)), (ish, dvg[sel, 0].sum(), dE)
    _, pull = ph.rrule_E_Ps_loc(H, rho)
    gsh, gt = pull(2.0)
    assert np.allclose(gt, 2.0 * dtau, rtol=0, atol=1e-14)
    assert abs(gsh[5, 0] - 2.0 * dvg[pw.idx_g2shells == 5, 0].sum()) < 1e-12


def test_forces_Ps_nloc_vs_oracle(si8):
    import pwdft_hip as ph
    from oracle.energy import calc_E_Ps_nloc
    from oracle.forces import calc_forces_Ps_nloc
    H, oH, _, psiks, Focc = si8
    print("case Si8 forces_Ps_nloc", flush=True)
    F = ph.calc_forces_Ps_nloc(H, psiks, Focc)
    Fo = calc_forces_Ps_nloc(oH, psiks, Focc)
    assert np.abs(Fo).max() > 1e-3
    assert np.abs(F - Fo).max() < 1e-11 * max(1.0, np.abs(Fo).max()), np.abs(F - Fo).max()
    E, pull = ph.rrule_E_Ps_nloc(H, psiks, Focc)
    Eo = calc_E_Ps_nloc(oH, psiks, Focc)
    assert abs(E - Eo) < 1e-11 * max(1.0, abs(Eo)), (E, Eo)
    assert np.allclose(pull(1.0), -F, rtol=0, atol=1e-15)
    # device-resident psi gives the same
    dps = [ph.DeviceMatrix.from_host(H.ctx, p) for p in psiks]
    assert np.abs(ph.calc_forces_Ps_nloc(H, dps, Focc) - F).max() < 1e-15


def test_energy_gradients_fd():
    """Central differences (h = 1e-4 bohr) of the device energies E_Ps_loc (rho fixed) and
    E_Ps_nloc (psi fixed) against the adjoint kernels' dE/dtau."""
    import pwdft_hip as ph
    H, _, _ = _si8(meshk=(1, 1, 1))
    rho = H.guess_rhoe()
    psiks = _orthonormal_psiks(H, 21)
    Focc = np.full((16, 1), 2.0)
    print("case Si8 energy-gradient finite differences", flush=True)
    _, _, dtau = ph.E_Ps_loc_grad(H, rho)
    _, pull = ph.rrule_E_Ps_nloc(H, psiks, Focc)
    dnl = pull(1.0)
    pos0 = H.atoms.positions.copy()
    h = 1e-4
    for ia, d in ((0, 0), (3, 1), (6, 2)):
        el, en = [], []
        for s in (h, -h):
            pos = pos0.copy()
            pos[d, ia] += s
            H.set_positions(pos)
            el.append(ph.E_Ps_loc_grad(H, rho)[0])
            en.append(ph.rrule_E_Ps_nloc(H, psiks, Focc)[0])
            H.set_positions(pos0)
        fl = (el[0] - el[1]) / (2 * h)
        fn = (en[0] - en[1]) / (2 * h)
        assert abs(fl - dtau[d, ia]) < 2e-8 * max(1.0, abs(fl)), (ia, d, fl, dtau[d, ia])
        assert abs(fn - dnl[d, ia]) < 2e-8 * max(1.0, abs(fn)), (ia, d, fn, dnl[d, ia])


def _si2_displaced(**kw):
    import pwdft_hip as ph
    L = ph.gen_lattice_fcc(A_SI)
    atoms = ph.Atoms(["Si", "Si"], [[0, 0, 0], [0.25, 0.25, 0.25]], L)
    atoms.positions[:, 0] += [0.06, -0.035, 0.02]
    return ph.Hamiltonian(atoms, ["Si-q4.lda.gth"], 10.0, meshk=(2, 2, 2), extra_states=2, use_symmetry=False,
                          **kw)


def _scf(H, mixed_until=None):
    import pwdft_hip as ph
    return ph.KS_solve_SCF_(H, betamix=0.5, etot_conv_thr=1e-11, ethr_evals_last=1e-11, NiterMax=80,
                            mixed_until=mixed_until)


def test_scf_forces_fd():
    """Self-consistent Si2 (one atom displaced): total force NN + Ps_loc + Ps_nloc against
    central differences (h = 2e-3 bohr) of the converged total energy."""
    import pwdft_hip as ph
    H = _si2_displaced()
    print("case Si2 SCF forces", flush=True)
    out = _scf(H)
    rho = ph.calc_rhoe(H, out["psiks"])
    F = ph.calc_forces(H, out["psiks"], rho)
    pos0 = H.atoms.positions.copy()
    h = 2e-3
    for ia, d in ((0, 0), (0, 1), (1, 2)):
        E = []
        for s in (h, -h):
            pos = pos0.copy()
            pos[d, ia] += s
            H.set_positions(pos)
            E.append(_scf(H)["energies"]["Total"])
        H.set_positions(pos0)
        fd = -(E[0] - E[1]) / (2 * h)
        print(f"  F[{d},{ia}] = {F[d, ia]:+.8f}  fd = {fd:+.8f}", flush=True)
        assert abs(F[d, ia] - fd) < 2e-5, (ia, d, F[d, ia], fd)


def test_scf_forces_mixed_vs_fp64():
    """Si8 (displaced, 2x2x2 k-points, 20 bands): the mixed precision SCF (FP32 FFT/V and
    projectors, FP64 refinement at the end) gives total energy within 1e-6 Ha/atom and
    forces within 1e-4 Ha/bohr of the FP64 SCF."""
    import pwdft_hip as ph
    print("case Si8 SCF mixed vs FP64 forces", flush=True)
    res = {}
    for mixed in (False, True):
        H, _, _ = _si8(meshk=(2, 2, 2), Nstates=20, mixed=mixed)
        out = ph.KS_solve_SCF_(H, betamix=0.3, etot_conv_thr=1e-9, ethr_evals_last=1e-10, NiterMax=80,
                               mixed_until=1e-6 if mixed else None)
        rho = ph.calc_rhoe(H, out["psiks"])
        res[mixed] = (out["energies"]["Total"], ph.calc_forces(H, out["psiks"], rho))
    dE = abs(res[True][0] - res[False][0]) / 8
    dF = np.abs(res[True][1] - res[False][1]).max()
    print(f"  dE/atom = {dE:.2e}  max|dF| = {dF:.2e}", flush=True)
    assert dE < 1e-6 and dF < 1e-4
