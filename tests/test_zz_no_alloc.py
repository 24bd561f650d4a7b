"""pw_op_H_dev_noalloc (op_H with every workspace in the library's fixed device array)
against pw_op_H_dev, through the C ABI, on the Si64 H-psi system (72^3 grid, Ngw 23 871,
Nbeta 320): two calls in a row, 160 bands then 40, FP64 and mixed.  160 FP64 bands need
four launches of the FFT passes through the fixed work array (one band's work array is
1.16 MB) and a shorter split of the projector products' K, so the results match to
rounding: FP64 <= 1e-13 and mixed <= 1e-6 relative to max|Hpsi| (the mixed path's own
tolerance against the oracle is 2e-5).  Sizes beyond the fixed array are rejected."""
import numpy as np
import pytest

from _systems import rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def si64():
    import pwdft_hip as ph
    frac = np.array(ph.diamond_supercell(1.0, 2), dtype=float)
    L = ph.gen_lattice_sc(20.5262)
    H = ph.Hamiltonian(ph.Atoms(["Si"] * 64, frac, L), ["Si-q4.lda.gth"], 15.0, Ns_=(72, 72, 72), Nstates=160,
                       band_batch=160, use_symmetry=False)
    assert H.pw.Ngw[0] == 23871 and H.pspotNL.NbetaNL == 320
    H.update(H.guess_rhoe())
    return H


def _apply(H, entry, X, Y, nb, flags):
    H.ctx.call(entry, 0, 0, nb, X.p, X.rows, Y.p, Y.rows, flags)
    return Y.download()[:, :nb]


@pytest.mark.parametrize("mixed", [False, True])
def test_noalloc_matches_op_H_dev_twice(si64, mixed):
    import pwdft_hip as ph
    from pwdft_hip._lib import PW_MIXED, PW_TERM_ALL
    H = si64
    flags = PW_TERM_ALL | (PW_MIXED if mixed else 0)
    tol = 1e-6 if mixed else 1e-13
    rng = np.random.default_rng(355)
    n = H.pw.Ngw[0]
    a = rng.standard_normal((n, 160)) + 1j * rng.standard_normal((n, 160))
    X = ph.DeviceMatrix.from_host(H.ctx, a)
    Yr = ph.DeviceMatrix(H.ctx, n, 160)
    Yn = ph.DeviceMatrix(H.ctx, n, 160)
    for nb in (160, 40):
        print(f"case noalloc Si64 nb={nb} mixed={mixed}", flush=True)
        ref = _apply(H, "pw_op_H_dev", X, Yr, nb, flags)
        # the fixed array keeps the previous call's contents: fill the output with garbage
        # first so that a missed write shows
        Yn.upload(np.full((n, 160), 7.0 + 3.0j))
        got = _apply(H, "pw_op_H_dev_noalloc", X, Yn, nb, flags)
        err = rel(got, ref)
        print(f"  rel err {err:.3e}", flush=True)
        assert err < tol, (nb, err)


def test_noalloc_rejects_oversized_beta_block(si64):
    """beta^H psi of Nbeta x nb beyond the fixed 8 MiB share is an argument error (no
    kernel runs): 320 projectors x 1700 bands x 16 B = 8.7 MB."""
    import pwdft_hip as ph
    from pwdft_hip._lib import PW_TERM_ALL, PWError
    H = si64
    n = H.pw.Ngw[0]
    X = ph.DeviceMatrix(H.ctx, n, 1700)
    Y = ph.DeviceMatrix(H.ctx, n, 1700)
    with pytest.raises(PWError) as ei:
        H.ctx.call("pw_op_H_dev_noalloc", 0, 0, 1700, X.p, X.rows, Y.p, Y.rows, PW_TERM_ALL)
    assert ei.value.code == -1 and "fixed workspace" in str(ei.value)
