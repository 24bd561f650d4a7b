"""The density side of the SCF iteration on the device, through the C ABI: update!
(Poisson_solve!, Slater + VWN and PBE exchange-correlation, V_Total and the local energy
terms), symmetrize_rhoe!, simple mixing and calc_rhoe into device memory, against the
oracle's restatements (oracle/density.py, oracle/symmetry.py).  FP64 throughout; the
tolerances are relative to the largest entry."""
import numpy as np
import pytest

from _systems import A_SI, PSP, oracle_ham, product_ham, rand_psi, rel, si_fcc_structure

pytestmark = pytest.mark.gpu


def _oracle_pw(H):
    from oracle.setup import PWGrid
    return PWGrid(H.pw.ecutwfc, H.atoms.LatVecs, Ns=H.pw.Ns)


def _test_density(H, seed=0):
    """Atomic-superposition density scaled by a ramp (low-density corners where the
    functionals' thresholds switch terms off), clamped at eps() so that it also carries
    components outside the rho sphere."""
    rho = H.guess_rhoe()
    ramp = np.linspace(0.02, 1.0, rho.size)
    rng = np.random.default_rng(seed)
    return np.maximum(rho * ramp * (1.0 + 0.01 * rng.standard_normal(rho.size)), np.finfo(float).eps)


def _systems():
    import pwdft_hip as ph
    sym, frac = si_fcc_structure()
    yield "si2_27", lambda xc: product_ham(sym, frac, ph.gen_lattice_fcc(A_SI), "Si-q4.lda.gth", 15.0, xcfunc=xc)
    yield "si8_45", lambda xc: product_ham(["Si"] * 8, ph.diamond_supercell(1.0, 1), ph.gen_lattice_sc(A_SI),
                                           "Si-q4.lda.gth", 20.0, xcfunc=xc)
    yield "si8_108", lambda xc: product_ham(["Si"] * 8, ph.diamond_supercell(1.0, 1), ph.gen_lattice_sc(A_SI),
                                            "Si-q4.lda.gth", 20.0, xcfunc=xc, Ns_=(108, 108, 108))


@pytest.mark.parametrize("xc", ["VWN", "PBE"])
@pytest.mark.parametrize("name,make", list(_systems()))
def test_update_potential_vs_oracle(name, make, xc):
    from oracle.density import Poisson_solve, calc_epsxc_Vxc_PBE, calc_epsxc_Vxc_VWN
    print(f"case update_potential {name} {xc}", flush=True)
    H = make(xc)
    opw = _oracle_pw(H)
    assert np.array_equal(opw.idx_g2r, H.pw.idx_g2r)
    rho = _test_density(H)
    E = H.update(rho)
    VH = Poisson_solve(opw, rho)
    eps, Vxc = calc_epsxc_Vxc_PBE(opw, rho) if xc == "PBE" else calc_epsxc_Vxc_VWN(rho)
    p = H.potentials
    assert rel(p.Hartree, VH) < 1e-12, rel(p.Hartree, VH)
    assert rel(p.XC[:, 0], Vxc) < (1e-11 if xc == "PBE" else 1e-13), rel(p.XC[:, 0], Vxc)
    assert rel(p.Total[:, 0], p.Ps_loc + VH + Vxc) < 1e-12
    dV = H.pw.CellVolume / H.pw.Npoints
    ref = (np.dot(p.Ps_loc, rho) * dV, 0.5 * np.dot(VH, rho) * dV, np.sum(eps * rho) * dV)
    for got, want in zip(E, ref):
        assert abs(got - want) < 1e-11 * max(abs(want), 1.0), (E, ref)


@pytest.mark.parametrize("mode", [1, 2])
def test_update_potential_compressed_planes(mode):
    """The same update (PBE: R->G and G->R of several fields) and V_Ps_loc build with the
    compressed plane kernels (mode 1) or the global-plane pass (mode 2, compressed planes
    in HBM, three launches per plane pass) forced on a grid whose planes fit in LDS give the
    result of the in-LDS kernels to rounding."""
    import pwdft_hip as ph
    from pwdft_hip._lib import PW_TUNE_PLANE_COMPRESSED
    sym, frac = si_fcc_structure()
    H = product_ham(sym, frac, ph.gen_lattice_fcc(A_SI), "Si-q4.lda.gth", 15.0, xcfunc="PBE")
    rho = _test_density(H, 3)
    E0 = H.update(rho)
    V0 = H.potentials.Total[:, 0].copy()
    P0 = H.potentials.Ps_loc.copy()
    print(f"case update plane mode {mode}", flush=True)
    H.ctx.call("pw_tune_set", PW_TUNE_PLANE_COMPRESSED, mode)
    try:
        H._build_geometry()
        E1 = H.update(rho)
    finally:
        H.ctx.call("pw_tune_set", PW_TUNE_PLANE_COMPRESSED, 0)
    assert rel(H.potentials.Ps_loc, P0) < 1e-14
    assert rel(H.potentials.Total[:, 0], V0) < 1e-14
    assert np.allclose(E0, E1, rtol=1e-13, atol=0)


def test_si216_pbe_setup_and_update_vs_oracle():
    """The Si216 PBE configuration (simple cubic a = 30.789 bohr, atoms moved by up to 0.02
    bohr, 15 Ha, 108^3 grid): its rho-sphere planes exceed LDS even compressed (107 active x
    lines of 108 in FP64), so V_Ps_loc, R->G and G->R run the global-plane pass.  V_Ps_loc and
    the potentials of update(guess density) against the oracle at the full size."""
    import pwdft_hip as ph
    from oracle.density import Poisson_solve, calc_epsxc_Vxc_PBE
    from oracle.setup import Atoms, PsPot_GTH, calc_V_Ps_loc
    import os
    frac = np.array(ph.diamond_supercell(1.0, 3), dtype=float)
    L = ph.gen_lattice_sc(30.789)
    cart = L @ frac.T + np.random.default_rng(216).uniform(-0.02, 0.02, size=(3, len(frac)))
    print("case Si216 PBE 108^3 set-up", flush=True)
    H = ph.Hamiltonian(ph.Atoms(["Si"] * len(frac), np.linalg.solve(L, cart).T, L), ["Si-q4.pbe.gth"], 15.0,
                       Ns_=(108, 108, 108), Nstates=480, xcfunc="PBE", use_symmetry=False, band_batch=96)
    assert H.pw.Ns == (108, 108, 108)
    opw = _oracle_pw(H)
    oat = Atoms(["Si"] * len(frac), cart.T, L, frac=False)
    Vps = calc_V_Ps_loc(oat, opw, [PsPot_GTH(os.path.join(PSP, "Si-q4.pbe.gth"))])
    assert rel(H.potentials.Ps_loc, Vps) < 1e-12, rel(H.potentials.Ps_loc, Vps)
    rho = H.guess_rhoe()
    print("case Si216 PBE update", flush=True)
    H.update(rho)
    VH = Poisson_solve(opw, rho)
    _, Vxc = calc_epsxc_Vxc_PBE(opw, rho)
    assert rel(H.potentials.Hartree, VH) < 1e-12, rel(H.potentials.Hartree, VH)
    assert rel(H.potentials.XC[:, 0], Vxc) < 1e-11, rel(H.potentials.XC[:, 0], Vxc)


@pytest.mark.parametrize("mixed", [False, True])
def test_installed_potential_matches_set_vloc(mixed):
    """op_H after update uses exactly V_Total: the same bits as uploading V_Total from the
    host with pw_set_vloc, in FP64 and in the mixed path."""
    import pwdft_hip as ph
    sym, frac = si_fcc_structure()
    H = product_ham(sym, frac, ph.gen_lattice_fcc(A_SI), "Si-q4.lda.gth", 15.0)
    H.update(_test_density(H, 1))
    psi = rand_psi(np.random.default_rng(5), H.pw.Ngw[0], 8)
    H.mixed = mixed
    a = ph.op_H(H, psi)
    H.set_Vloc(H.potentials.Total[:, 0].copy())
    b = ph.op_H(H, psi)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("name,symbols,frac,lat,psp", [
    ("si_fcc", *si_fcc_structure(), "fcc", "Si-q4.lda.gth"),
    ("al_fcc", ["Al"], [[0, 0, 0]], "fcc_al", "Al-q3.lda.gth"),
])
def test_symmetrize_vs_oracle(name, symbols, frac, lat, psp):
    """symmetrize_rhoe! on the device against the oracle's orbit algorithm, for a density
    that is not band-limited (its components outside the sphere must be kept), then
    idempotence on the device."""
    import pwdft_hip as ph
    from oracle.symmetry import symmetrize_rhoe
    print(f"case symmetrize {name}", flush=True)
    L = ph.gen_lattice_fcc(A_SI) if lat == "fcc" else ph.gen_lattice_fcc(7.6)
    H = product_ham(symbols, frac, L, psp, 15.0, meshk=(3, 3, 3), use_symmetry=True)
    ops = H.sym_ops
    assert len(ops) == 48
    inv_s = [np.rint(np.linalg.inv(R).T).astype(np.int64) for R, _ in ops]
    ft = [-t for _, t in ops]
    rho = np.random.default_rng(2).uniform(size=H.pw.Npoints)
    ref = symmetrize_rhoe(_oracle_pw(H), inv_s, ft, rho)
    d = ph.DeviceVector.from_host(H.ctx, rho)
    H.ctx.call("pw_symmetrize_rhoe_dev", d.p)
    got = d.download()
    assert rel(got, ref) < 1e-13, rel(got, ref)
    H.ctx.call("pw_symmetrize_rhoe_dev", d.p)
    assert rel(d.download(), got) < 1e-13


def test_mix_and_clamp():
    import pwdft_hip as ph
    sym, frac = si_fcc_structure()
    H = product_ham(sym, frac, ph.gen_lattice_fcc(A_SI), "Si-q4.lda.gth", 15.0)
    rng = np.random.default_rng(9)
    a = rng.standard_normal(H.pw.Npoints)
    b = rng.standard_normal(H.pw.Npoints)
    da, db = ph.DeviceVector.from_host(H.ctx, a), ph.DeviceVector.from_host(H.ctx, b)
    H.ctx.call("pw_mix_rhoe_dev", 0.3, da.p, db.p, db.p)
    want = np.maximum(0.3 * a + 0.7 * b, np.finfo(float).eps)
    got = db.download()
    assert np.abs(got - want).max() < 4e-16 * np.abs(want).max()
    assert got.min() >= np.finfo(float).eps


def test_calc_rhoe_into_device_vector():
    """calc_rhoe_ into a DeviceVector (symmetrised in place) equals calc_rhoe_ into a host
    array, and the unsymmetrised sum agrees with the oracle's calc_rhoe symmetrised by the
    oracle's symmetrize_rhoe."""
    import pwdft_hip as ph
    from oracle.density import calc_rhoe as o_rhoe
    from oracle.symmetry import symmetrize_rhoe
    sym, frac = si_fcc_structure()
    H = product_ham(sym, frac, ph.gen_lattice_fcc(A_SI), "Si-q4.lda.gth", 15.0, meshk=(3, 3, 3),
                    use_symmetry=True, extra_states=4)
    rng = np.random.default_rng(4)
    host = [rand_psi(rng, n, H.electrons.Nstates) for n in H.pw.Ngw]
    psiks = [ph.DeviceMatrix.from_host(H.ctx, a) for a in host]
    d = ph.DeviceVector(H.ctx, H.pw.Npoints)
    ph.calc_rhoe_(H, psiks, d)
    r = ph.calc_rhoe(H, psiks)[:, 0]
    assert rel(d.download(), r) < 1e-15
    oH = oracle_ham(sym, frac, ph.gen_lattice_fcc(A_SI), "Si-q4.lda.gth", 15.0, kpoints=(H.pw.k, H.pw.wk),
                    extra_states=4)
    ops = H.sym_ops
    inv_s = [np.rint(np.linalg.inv(R).T).astype(np.int64) for R, _ in ops]
    ref = symmetrize_rhoe(oH.pw, inv_s, [-t for _, t in ops], o_rhoe(oH, host, Focc=H.electrons.Focc))
    assert rel(r, ref) < 1e-12, rel(r, ref)


def test_psi_energies_vs_oracle():
    """E_kin and E_Ps_nloc reduced on the device (pw_energies_psi_dev) against the oracle's
    calc_E_kin / calc_E_Ps_nloc, Si fcc on a 2x2x2 mesh, random occupations."""
    import pwdft_hip as ph
    from pwdft_hip.scf import _psi_energies
    from oracle.energy import calc_E_Ps_nloc, calc_E_kin
    sym, frac = si_fcc_structure()
    H = product_ham(sym, frac, ph.gen_lattice_fcc(A_SI), "Si-q4.lda.gth", 15.0, meshk=(2, 2, 2), extra_states=4)
    oH = oracle_ham(sym, frac, ph.gen_lattice_fcc(A_SI), "Si-q4.lda.gth", 15.0, kpoints=(H.pw.k, H.pw.wk),
                    extra_states=4)
    rng = np.random.default_rng(12)
    host = [rand_psi(rng, n, H.electrons.Nstates) for n in H.pw.Ngw]
    Focc = rng.uniform(0.0, 2.0, size=(H.electrons.Nstates, H.pw.Nkpt))
    psiks = [ph.DeviceMatrix.from_host(H.ctx, a) for a in host]
    ek, en = _psi_energies(H, psiks, Focc)
    ek0, en0 = calc_E_kin(oH, host, Focc), calc_E_Ps_nloc(oH, host, Focc)
    assert abs(ek - ek0) < 1e-12 * abs(ek0), (ek, ek0)
    assert abs(en - en0) < 1e-12 * abs(en0), (en, en0)
