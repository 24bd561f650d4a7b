"""The oracle against the frozen fixtures of tests/golden/ (written by
tests/golden/make_fixtures.py): op_H and calc_rhoe on Si fcc 2x2x2 (1e-13), and the SCF total
energy and components of Si fcc 3x3x3 (1e-9 Ha).  A change to the oracle that moves any of
these numbers fails here."""
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
sys.path.insert(0, GOLD)


def test_oracle_ops_match_fixture():
    from make_fixtures import dV, si2
    from oracle import ops as O
    from oracle.density import calc_rhoe
    d = np.load(os.path.join(GOLD, "si2_ops.npz"))
    H = si2((2, 2, 2))
    assert np.array_equal(np.asarray(H.pw.Ngw), d["Ngw"])
    H.V_loc[:, 0] = H.V_Ps_loc + dV(H.pw.Ns)
    H.ik = 1
    hp = O.op_H(H, d["psi"].astype(np.complex128))
    assert np.abs(hp - d["hpsi"]).max() < 1e-13 * np.abs(d["hpsi"]).max()
    psiks = [d[f"psik{ik}"].astype(np.complex128) for ik in range(H.pw.Nkpt)]
    rho = calc_rhoe(H, psiks, d["Focc"])
    assert np.abs(rho[d["rho_idx"]] - d["rho_samples"]).max() < 1e-13 * np.abs(d["rho_samples"]).max()
    assert abs(np.sum(rho) * H.pw.CellVolume / H.pw.Npoints - d["rho_integral"]) < 1e-12
    assert abs(np.linalg.norm(rho) - d["rho_norm"]) < 1e-12 * d["rho_norm"]


@pytest.mark.slow
def test_oracle_scf_matches_fixture():
    from make_fixtures import si2
    from oracle.energy import KS_solve_SCF
    ref = json.load(open(os.path.join(GOLD, "si2_scf.json")))
    e = KS_solve_SCF(si2((3, 3, 3)), betamix=0.5, etot_conv_thr=1e-11, ethr_evals_last=1e-13)[0]
    for k, v in ref.items():
        assert abs(e[k] - v) < 1e-9, (k, e[k], v)
