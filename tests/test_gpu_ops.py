"""GPU parity of the Hamiltonian application, density and dense products against the CPU
oracle, through the C ABI.  Tolerances: FP64 path <= 1e-12 relative to max|ref| (FFT
order and FMA contraction differ from numpy's; the arithmetic is the same); mixed path
(FP32 FFT, FP32 MFMA projectors) <= 2e-5 relative."""
import numpy as np
import pytest

from _systems import A_SI, oracle_ham, product_ham, rand_psi, rel, si_fcc_structure

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOLMIX = 2e-5


@pytest.fixture(scope="module")
def si_pair():
    import pwdft_hip as ph
    from oracle.setup import gen_lattice_fcc
    sym, frac = si_fcc_structure()
    kw = dict(meshk=(2, 2, 2), extra_states=4)
    oH = oracle_ham(sym, frac, gen_lattice_fcc(A_SI), "Si-q4.lda.gth", 15.0, **kw)
    H = product_ham(sym, frac, ph.gen_lattice_fcc(A_SI), "Si-q4.lda.gth", 15.0, **kw)
    rng = np.random.default_rng(11)
    V = oH.V_Ps_loc + 0.3 * rng.standard_normal(oH.pw.Npoints)
    oH.V_loc[:, 0] = V
    H.set_Vloc(V)
    return H, oH


def test_setup_matches_oracle(si_pair):
    H, oH = si_pair
    assert H.pw.Ns == oH.pw.Ns == (27, 27, 27)
    for ik in range(H.pw.Nkpt):
        assert np.array_equal(H.pw.idx_gw2r[ik], oH.pw.idx_gw2r[ik])
        assert rel(H.pspotNL.betaNL[ik], oH.pspotNL.betaNL[ik]) < 1e-13


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("term", ["K", "V_loc", "V_Ps_nloc", "H"])
def test_op_terms(si_pair, term, mixed):
    import pwdft_hip as ph
    from oracle import ops as O
    H, oH = si_pair
    H.mixed = mixed
    rng = np.random.default_rng(3)
    try:
        for ik in range(H.pw.Nkpt):
            H.ik = oH.ik = ik
            psi = rand_psi(rng, H.pw.Ngw[ik], 8)
            print(f"case op_{term} ik={ik} mixed={mixed}", flush=True)
            got = getattr(ph, "op_" + term)(H, psi)
            ref = getattr(O, "op_" + term)(oH, psi)
            err = rel(got, ref)
            assert err < (TOLMIX if mixed else TOL64), (term, ik, err)
    finally:
        H.mixed = False


def test_accumulate_forms(si_pair):
    import pwdft_hip as ph
    from oracle import ops as O
    H, oH = si_pair
    H.ik = oH.ik = 2
    rng = np.random.default_rng(5)
    psi = rand_psi(rng, H.pw.Ngw[2], 5)
    base = rand_psi(rng, H.pw.Ngw[2], 5)
    out = base.copy(order="F")
    ph.op_K_(H, psi, out)
    ph.op_V_loc_(H, psi, out)
    ph.op_V_Ps_nloc_(H, psi, out)
    assert rel(out, base + O.op_H(oH, psi)) < TOL64
    out2 = base.copy(order="F")
    ph.op_H_(H, psi, out2)  # overwrites
    assert rel(out2, O.op_H(oH, psi)) < TOL64


@pytest.mark.parametrize("mixed", [False, True])
def test_calc_rhoe(si_pair, mixed):
    import pwdft_hip as ph
    from oracle.density import calc_rhoe
    H, oH = si_pair
    H.mixed = mixed
    rng = np.random.default_rng(9)
    try:
        psiks = [rand_psi(rng, n, 8) for n in H.pw.Ngw]
        Focc = np.abs(rng.standard_normal((8, H.pw.Nkpt)))
        print(f"case calc_rhoe mixed={mixed}", flush=True)
        got = ph.calc_rhoe(H, psiks, Focc=Focc)[:, 0]
        ref = calc_rhoe(oH, psiks, Focc)
        assert rel(got, ref) < (TOLMIX if mixed else TOL64)
    finally:
        H.mixed = False


@pytest.mark.parametrize("transa", [0, 1])
@pytest.mark.parametrize("shape", [(1, 1, 1), (37, 19, 100), (64, 64, 64), (100, 130, 3001), (700, 8, 20)])
@pytest.mark.parametrize("mixed", [False, True])
def test_zgemm(si_pair, transa, shape, mixed):
    from pwdft_hip._lib import DeviceMatrix, PW_MIXED, ptr
    H, _ = si_pair
    c = H.ctx
    M, N, K = shape
    rng = np.random.default_rng(M * 7 + N * 3 + K)
    A = rand_psi(rng, K, M) if transa else rand_psi(rng, M, K)
    B = rand_psi(rng, K, N)
    C0 = rand_psi(rng, M, N)
    alpha = np.array([0.7, -0.2])
    beta = np.array([0.5, 0.25])
    dA, dB, dC = DeviceMatrix.from_host(c, A), DeviceMatrix.from_host(c, B), DeviceMatrix.from_host(c, C0)
    print(f"case zgemm transa={transa} shape={shape} mixed={mixed}", flush=True)
    c.call("pw_zgemm_dev", transa, M, N, K, ptr(alpha), dA.p, dA.rows, dB.p, dB.rows, ptr(beta), dC.p, M,
           PW_MIXED if mixed else 0)
    got = dC.download()
    opA = A.conj().T if transa else A
    ref = (0.7 - 0.2j) * (opA @ B) + (0.5 + 0.25j) * C0
    scale = np.abs(opA).max() * np.abs(B).max() * np.sqrt(K) + np.abs(C0).max()
    err = np.abs(got - ref).max() / scale
    assert err < (1e-5 if mixed else 1e-14), err


@pytest.mark.parametrize("split", [0, 1])
@pytest.mark.parametrize("tile", [-1, 0, 1, 2, 3, 4, 5])
def test_zgemm_every_tile(si_pair, tile, split):
    """Each workgroup tile shape (pw_tune_set PW_TUNE_GEMM_*_SHAPE) on ragged sizes, with
    split-K (K = 3001) and without, both precisions and both op(A); split = 1 runs the
    mixed products on BF16 MFMA with three-part operands (PW_TUNE_GEMM_SPLIT_BF16), held to
    the FP32 path's tolerance."""
    from pwdft_hip._lib import DeviceMatrix, PW_MIXED, ptr
    H, _ = si_pair
    c = H.ctx
    try:
        c.call("pw_tune_set", 0, tile)
        c.call("pw_tune_set", 1, tile)
        c.call("pw_tune_set", 8, split)
        for transa in (0, 1):
            for M, N, K in ((37, 19, 100), (100, 130, 3001), (300, 160, 70), (1, 1, 1)):
                rng = np.random.default_rng(M + 5 * N + 11 * K + transa)
                A = rand_psi(rng, K, M) if transa else rand_psi(rng, M, K)
                B = rand_psi(rng, K, N)
                C0 = rand_psi(rng, M, N)
                alpha = np.array([0.3, 0.9])
                beta = np.array([-0.5, 0.1])
                opA = A.conj().T if transa else A
                ref = (0.3 + 0.9j) * (opA @ B) + (-0.5 + 0.1j) * C0
                scale = np.abs(opA).max() * np.abs(B).max() * np.sqrt(K) + np.abs(C0).max()
                for mixed in (False, True):
                    dA, dB, dC = (DeviceMatrix.from_host(c, X) for X in (A, B, C0))
                    print(f"case zgemm tile={tile} transa={transa} shape={(M, N, K)} mixed={mixed}", flush=True)
                    c.call("pw_zgemm_dev", transa, M, N, K, ptr(alpha), dA.p, dA.rows, dB.p, dB.rows, ptr(beta),
                           dC.p, M, PW_MIXED if mixed else 0)
                    err = np.abs(dC.download() - ref).max() / scale
                    assert err < (1e-5 if mixed else 1e-14), (tile, transa, (M, N, K), mixed, err)
    finally:
        c.call("pw_tune_set", 0, -1)
        c.call("pw_tune_set", 1, -1)
        c.call("pw_tune_set", 8, -1)


def test_split_bf16_op_H_and_error_level(si_pair):
    """The mixed op_H with its projector products on three-part BF16 MFMA against the
    oracle (2e-5, the mixed tolerance), and the GEMM error of the split form at the level
    of the FP32 MFMA path (within 4x of it) on a long inner product."""
    import pwdft_hip as ph
    from oracle import ops as O
    from pwdft_hip._lib import DeviceMatrix, PW_MIXED, ptr
    H, oH = si_pair
    c = H.ctx
    rng = np.random.default_rng(23)
    M, N, K = 96, 40, 20000
    A = rand_psi(rng, K, M)
    B = rand_psi(rng, K, N)
    ref = A.conj().T @ B
    one, zero = np.array([1.0, 0.0]), np.array([0.0, 0.0])
    errs = {}
    for split in (0, 1):
        c.call("pw_tune_set", 8, split)
        dA, dB, dC = DeviceMatrix.from_host(c, A), DeviceMatrix.from_host(c, B), DeviceMatrix(c, M, N)
        c.call("pw_zgemm_dev", 1, M, N, K, ptr(one), dA.p, K, dB.p, K, ptr(zero), dC.p, M, PW_MIXED)
        errs[split] = np.abs(dC.download() - ref).max() / np.abs(ref).max()
    print(f"case split bf16 GEMM error {errs[1]:.2e} vs FP32 MFMA {errs[0]:.2e}", flush=True)
    assert errs[1] < 4 * errs[0] + 1e-7, errs
    try:
        c.call("pw_tune_set", 8, 1)
        H.mixed = True
        for ik in range(H.pw.Nkpt):
            H.ik = oH.ik = ik
            psi = rand_psi(rng, H.pw.Ngw[ik], 8)
            assert rel(ph.op_H(H, psi), O.op_H(oH, psi)) < TOLMIX
    finally:
        H.mixed = False
        c.call("pw_tune_set", 8, -1)


def test_kprec(si_pair):
    import pwdft_hip as ph
    from oracle import ops as O
    from pwdft_hip._lib import DeviceMatrix
    H, oH = si_pair
    H.ik = oH.ik = 1
    psi = rand_psi(np.random.default_rng(1), H.pw.Ngw[1], 4)
    got = ph.Kprec(H, DeviceMatrix.from_host(H.ctx, psi)).download()
    assert rel(got, O.Kprec(oH, psi)) < 1e-15


def test_betaNL_psi(si_pair):
    import pwdft_hip as ph
    from oracle import ops as O
    H, oH = si_pair
    H.ik = oH.ik = 3
    psi = rand_psi(np.random.default_rng(2), H.pw.Ngw[3], 6)
    got = ph.calc_betaNL_psi(H, psi)
    assert rel(got, O.calc_betaNL_psi(oH.pspotNL.betaNL[3], psi)) < TOL64


@pytest.mark.parametrize("case", ["al_fcc_20", "si8_45", "si8_72", "si8_72_compressed", "si8_72_global",
                                  "si8_108"])
def test_other_grids(case):
    """Grids with other radix mixes: 20 = 5*4, 45 = 9*5, 72 = 9*8, 108 = 9*4*3 (Ns_ override).
    At 108 the padded FP64 z-plane (188 KB) exceeds LDS and the compressed plane kernels
    run; "si8_72_compressed" forces them where the padded plane would fit, and
    "si8_72_global" forces the global-plane pass (compressed planes in HBM).  calc_rhoe is
    checked on the same grids."""
    import pwdft_hip as ph
    from oracle import ops as O
    from oracle import setup as S
    if case == "al_fcc_20":
        sym, frac, Lo, Lp, psp, ecut, kw = ["Al"], [[0, 0, 0]], S.gen_lattice_fcc(7.6), ph.gen_lattice_fcc(7.6), \
            "Al-q3.lda.gth", 15.0, dict(meshk=(2, 2, 2), shiftk=(1, 1, 1), Nstates=8)
    else:
        sym = ["Si"] * 8
        frac = ph.diamond_supercell(1.0, 1)
        Lo, Lp = S.gen_lattice_sc(A_SI), ph.gen_lattice_sc(A_SI)
        psp, ecut = "Si-q4.lda.gth", 20.0
        kw = dict(meshk=(1, 1, 1), extra_states=4)
        if case.startswith("si8_72"):
            kw["Ns"] = (72, 72, 72)
        if case == "si8_108":
            kw["Ns"] = (108, 108, 108)
    kwp = dict(kw)
    if "Ns" in kwp:
        kwp["Ns_"] = kwp.pop("Ns")
    oH = oracle_ham(sym, frac, Lo, psp, ecut, **kw)
    H = product_ham(sym, frac, Lp, psp, ecut, **kwp)
    if case == "si8_72_compressed":
        H.ctx.call("pw_tune_set", 5, 1)
    if case == "si8_72_global":
        H.ctx.call("pw_tune_set", 5, 2)
    rng = np.random.default_rng(4)
    V = oH.V_Ps_loc + 0.2 * rng.standard_normal(oH.pw.Npoints)
    oH.V_loc[:, 0] = V
    H.set_Vloc(V)
    print(f"case {case} Ns={H.pw.Ns} Ngw={H.pw.Ngw}", flush=True)
    for ik in range(H.pw.Nkpt):
        H.ik = oH.ik = ik
        psi = rand_psi(rng, H.pw.Ngw[ik], 6)
        assert rel(ph.op_H(H, psi), O.op_H(oH, psi)) < TOL64
        H.mixed = True
        assert rel(ph.op_H(H, psi), O.op_H(oH, psi)) < TOLMIX
        H.mixed = False
    from oracle.density import calc_rhoe as o_rhoe
    psiks = [rand_psi(rng, n, 6) for n in H.pw.Ngw]
    Focc = np.full((6, H.pw.Nkpt), 2.0)
    ref = o_rhoe(oH, psiks, Focc)
    assert rel(ph.calc_rhoe(H, psiks, Focc=Focc)[:, 0], ref) < TOL64
    H.mixed = True
    assert rel(ph.calc_rhoe(H, psiks, Focc=Focc)[:, 0], ref) < TOLMIX
    H.mixed = False


@pytest.mark.parametrize("mixed", [False, True])
def test_overlapped_nonlocal_stream(mixed):
    """op_H with the nonlocal products on the second stream (default) against the
    single-stream order, with and without accumulation, over several band batches; and a
    captured op_H (fork and join inside the graph) replayed against the eager result."""
    import pwdft_hip as ph
    from pwdft_hip._lib import PW_ACCUMULATE, PW_MIXED, PW_TERM_ALL
    atoms = ph.Atoms(["Si"] * 8, ph.diamond_supercell(1.0, 1), ph.gen_lattice_sc(10.2631))
    H = ph.Hamiltonian(atoms, ["Si-q4.lda.gth"], 20.0, use_symmetry=False, band_batch=7)
    H.update(H.guess_rhoe())
    rng = np.random.default_rng(17)
    n = H.pw.Ngw[0]
    psi = ph.DeviceMatrix.from_host(H.ctx, rand_psi(rng, n, 20))
    h0 = rand_psi(rng, n, 20)
    flags = PW_TERM_ALL | (PW_MIXED if mixed else 0)
    res = {}
    for ov in (0, 1):
        H.ctx.call("pw_tune_set", 6, ov)
        for acc in (0, PW_ACCUMULATE):
            out = ph.DeviceMatrix.from_host(H.ctx, h0)
            H.ctx.call("pw_op_H_dev", 0, 0, 20, psi.p, n, out.p, n, flags | acc)
            res[(ov, acc)] = out.download()
    for acc in (0, PW_ACCUMULATE):
        assert rel(res[(1, acc)], res[(0, acc)]) < 1e-13, acc
    out = ph.DeviceMatrix(H.ctx, n, 20)
    H.ctx.call("pw_op_H_dev", 0, 0, 20, psi.p, n, out.p, n, flags)   # warm workspaces
    H.ctx.call("pw_synchronize")
    with ph.Graph(H.ctx) as g:
        H.ctx.call("pw_op_H_dev", 0, 0, 20, psi.p, n, out.p, n, flags)
    for rep in range(3):
        out.upload(np.full((n, 20), np.nan + 0j))
        g.launch()
        H.ctx.call("pw_synchronize")
        assert np.array_equal(out.download(), res[(1, 0)]), rep
    g.free()
