"""The op_H autotuner and the launch parameters it chooses among, through the C ABI: every
FFT radix mode (PW_TUNE_FFT_ORDER) gives the default op_H to rounding on grids of 27^3
(27 = 9 x 3), 20^3 (5 x 4) and 72^3 (9 x 8), FP64 <= 1e-13 and mixed <= 1e-5 relative; a tuned
Hamiltonian's op_H equals the untuned one's; the tuning cache is written on the first
Hamiltonian of a shape and read, without timing, by the next one."""
import numpy as np
import pytest

from _systems import A_SI, product_ham, rand_psi, rel, si_fcc_structure

pytestmark = pytest.mark.gpu

TUNE_FFT_ORDER = 15


def _si_fcc(**kw):
    import pwdft_hip as ph
    sym, frac = si_fcc_structure()
    return product_ham(sym, frac, ph.gen_lattice_fcc(A_SI), "Si-q4.lda.gth", 15.0, meshk=(2, 2, 2),
                       extra_states=4, **kw)


def _al(**kw):
    import pwdft_hip as ph
    return ph.Hamiltonian(ph.Atoms(["Al"], [[0, 0, 0]], ph.gen_lattice_fcc(7.6)), ["Al-q3.lda.gth"], 15.0,
                          meshk=(2, 2, 2), extra_states=4, use_symmetry=False, **kw)


def _si64():
    import pwdft_hip as ph
    frac = np.array(ph.diamond_supercell(1.0, 2), dtype=float)
    return ph.Hamiltonian(ph.Atoms(["Si"] * 64, frac, ph.gen_lattice_sc(20.5262)), ["Si-q4.lda.gth"], 15.0,
                          Ns_=(72, 72, 72), Nstates=160, use_symmetry=False)


@pytest.mark.parametrize("system", ["si_fcc_27", "al_20", "si64_72"])
def test_fft_radix_modes_match_default(system):
    import pwdft_hip as ph
    H = {"si_fcc_27": _si_fcc, "al_20": _al, "si64_72": _si64}[system]()
    rng = np.random.default_rng(9)
    V = H.potentials.Ps_loc + 0.2 * rng.standard_normal(H.pw.Npoints)
    H.set_Vloc(V)
    H.ik = 0
    psi = rand_psi(rng, H.pw.Ngw[0], 12)
    try:
        for mixed in (False, True):
            H.mixed = mixed
            H.ctx.call("pw_tune_set", TUNE_FFT_ORDER, 0)
            ref = ph.op_H(H, psi)
            for mode in (1, 2):
                print(f"case {system} radix mode {mode} mixed={mixed}", flush=True)
                H.ctx.call("pw_tune_set", TUNE_FFT_ORDER, mode)
                err = rel(ph.op_H(H, psi), ref)
                assert err < (1e-5 if mixed else 1e-13), (mode, mixed, err)
    finally:
        H.mixed = False
        H.ctx.call("pw_tune_set", TUNE_FFT_ORDER, 0)


def test_autotune_cache_and_tuned_op_H(tmp_path):
    import pwdft_hip as ph
    from pwdft_hip import autotune as at
    path = str(tmp_path / "tune.json")
    H0 = _si_fcc()
    rng = np.random.default_rng(4)
    V = H0.potentials.Ps_loc + 0.2 * rng.standard_normal(H0.pw.Npoints)
    H0.set_Vloc(V)
    H0.ik = 1
    psi = rand_psi(rng, H0.pw.Ngw[1], H0.electrons.Nstates)
    ref = ph.op_H(H0, psi)
    print("case autotune (timed)", flush=True)
    settings, source = at.autotune(H0, path=path, quick=True, steps=2)
    assert source == "tuned" and settings["op_H_ms"] > 0
    assert set(at.DEFAULTS) <= set(settings)
    assert rel(ph.op_H(H0, psi), ref) < 1e-13
    key = at.shape_key(H0.pw.Ns, max(H0.pw.Ngw), H0.pspotNL.NbetaNL, H0.electrons.Nstates, False)
    assert at.load_cache(path)[key]["fft_order"] == settings["fft_order"]
    print("case autotune (cache)", flush=True)
    import os
    os.environ["PWDFT_HIP_TUNE_CACHE"] = path
    try:
        H1 = _si_fcc(autotune=True)
    finally:
        del os.environ["PWDFT_HIP_TUNE_CACHE"]
    assert H1.tuned[1] == "cache" and H1.tuned[0]["fft_order"] == settings["fft_order"]
    H1.set_Vloc(V)
    H1.ik = 1
    assert rel(ph.op_H(H1, psi), ref) < 1e-13
