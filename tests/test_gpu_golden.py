"""The device path against the frozen fixtures of tests/golden/ (the oracle's outputs, written
by tests/golden/make_fixtures.py), through the C ABI: op_H (FP64 <= 1e-12, mixed <= 2e-5
relative) and calc_rhoe (1e-12) on Si fcc 2x2x2; the FP64 GPU SCF of Si fcc 3x3x3 against the
frozen total energy and components (1e-9 Ha)."""
import json
import os
import sys

import numpy as np
import pytest

from _systems import A_SI, product_ham, si_fcc_structure

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
sys.path.insert(0, GOLD)


def _si2(meshk):
    import pwdft_hip as ph
    sym, frac = si_fcc_structure()
    return product_ham(sym, frac, ph.gen_lattice_fcc(A_SI), "Si-q4.lda.gth", 15.0, meshk=meshk, extra_states=4)


@pytest.mark.parametrize("mixed", [False, True])
def test_device_ops_match_fixture(mixed):
    import pwdft_hip as ph
    from make_fixtures import dV
    d = np.load(os.path.join(GOLD, "si2_ops.npz"))
    H = _si2((2, 2, 2))
    assert np.array_equal(np.asarray(H.pw.Ngw), d["Ngw"])
    H.set_Vloc(H.potentials.Ps_loc + dV(H.pw.Ns))
    H.ik = 1
    H.mixed = mixed
    print(f"case golden op_H / calc_rhoe mixed={mixed}", flush=True)
    try:
        hp = ph.op_H(H, d["psi"].astype(np.complex128))
        psiks = [np.asfortranarray(d[f"psik{ik}"].astype(np.complex128)) for ik in range(H.pw.Nkpt)]
        rho = ph.calc_rhoe(H, psiks, Focc=d["Focc"])[:, 0]
    finally:
        H.mixed = False
    tol = 2e-5 if mixed else 1e-12
    assert np.abs(hp - d["hpsi"]).max() < tol * np.abs(d["hpsi"]).max()
    assert np.abs(rho[d["rho_idx"]] - d["rho_samples"]).max() < tol * np.abs(d["rho_samples"]).max()
    assert abs(np.linalg.norm(rho) - d["rho_norm"]) < tol * d["rho_norm"]


def test_device_scf_matches_fixture():
    import pwdft_hip as ph
    ref = json.load(open(os.path.join(GOLD, "si2_scf.json")))
    H = _si2((3, 3, 3))
    print("case golden SCF Si fcc 3x3x3", flush=True)
    out = ph.KS_solve_SCF_(H, betamix=0.5, etot_conv_thr=1e-11, ethr_evals_last=1e-13, NiterMax=150)
    e = out["energies"]
    for k, v in ref.items():
        assert abs(e[k] - v) < 1e-9, (k, e[k], v)
