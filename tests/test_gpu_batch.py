"""Batched k-point launches through the C ABI: op_H over all k-points of the Al fcc
12x12x12 shifted mesh (182 symmetry-reduced k-points, uneven plane-wave counts) in one
launch per stage against the per-k op_H and the oracle, and the batched complex GEMM
against numpy."""
import ctypes

import numpy as np
import pytest

from _systems import rand_psi

pytestmark = pytest.mark.gpu


def _al(**kw):
    import pwdft_hip as ph
    atoms = ph.Atoms(["Al"], [[0, 0, 0]], ph.gen_lattice_fcc(7.6))
    return ph.Hamiltonian(atoms, ["Al-q3.lda.gth"], 15.0, meshk=(12, 12, 12), shiftk=(1, 1, 1), Nstates=8,
                          use_smearing=True, smearing_kT=0.01, **kw)


@pytest.fixture(scope="module")
def al():
    H = _al()
    H.update(H.guess_rhoe())
    rng = np.random.default_rng(31)
    host = [rand_psi(rng, n, 8) for n in H.pw.Ngw]
    return H, host


def test_al_mesh_shape(al):
    H, _ = al
    assert H.pw.Nkpt == 182
    assert min(H.pw.Ngw) < max(H.pw.Ngw)
    assert abs(H.pw.wk.sum() - 1.0) < 1e-14


@pytest.mark.parametrize("mixed", [False, True])
def test_op_H_batch_vs_per_k(al, mixed):
    import pwdft_hip as ph
    H, host = al
    print(f"case Al 182k op_H batch mixed={mixed}", flush=True)
    dev = [ph.DeviceMatrix.from_host(H.ctx, a) for a in host]
    H.mixed = mixed
    try:
        out = ph.op_H_batch(H, dev)
        out2 = ph.op_H_batch(H, dev)      # cached descriptors
    finally:
        H.mixed = False
    worst = 0.0
    for ik in range(H.pw.Nkpt):
        H.ik = ik
        ref = ph.op_H(H, host[ik])        # per-k FP64 path
        got = out[ik].download()
        assert np.array_equal(got, out2[ik].download())
        worst = max(worst, np.abs(got - ref).max() / np.abs(ref).max())
    assert worst < (2e-5 if mixed else 1e-13), worst


def test_op_H_batch_vs_oracle(al):
    """A few k-points of the batch against the oracle's op_H (same k list and V)."""
    import pwdft_hip as ph
    from _systems import oracle_ham
    H, host = al
    oH = oracle_ham(["Al"], [[0, 0, 0]], ph.gen_lattice_fcc(7.6), "Al-q3.lda.gth", 15.0,
                    kpoints=(H.pw.k, H.pw.wk), Nstates=8)
    oH.V_loc[:, 0] = H.potentials.Total[:, 0]
    from oracle import ops as O
    dev = [ph.DeviceMatrix.from_host(H.ctx, a) for a in host]
    out = ph.op_H_batch(H, dev)
    for ik in (0, 57, 181):
        oH.ik = ik
        ref = O.op_H(oH, host[ik])
        got = out[ik].download()
        assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max(), ik


def test_op_H_batch_terms_and_accumulate(al):
    import pwdft_hip as ph
    from pwdft_hip._lib import PW_TERM_K, PW_TERM_NLOC, PW_TERM_VLOC
    H, host = al
    dev = [ph.DeviceMatrix.from_host(H.ctx, a) for a in host[:20]]
    iks = list(range(20))
    full = [m.download() for m in ph.op_H_batch(H, dev, iks=iks)]
    acc = ph.op_H_batch(H, dev, terms=PW_TERM_K, iks=iks)
    ph.op_H_batch(H, dev, out=acc, terms=PW_TERM_VLOC, accumulate=True, iks=iks)
    ph.op_H_batch(H, dev, out=acc, terms=PW_TERM_NLOC, accumulate=True, iks=iks)
    for a, f in zip(acc, full):
        assert np.abs(a.download() - f).max() < 1e-13 * np.abs(f).max()


# batches of (m, n, k): mixed sizes (the 16 x 16-tile kernel), and batches whose every
# product fits the small-product kernels (A B: k <= 64, n <= 32; A^H B: m, n <= 32)
ZGEMM_BATCHES = {
    0: [[(5, 8, 300), (24, 24, 311), (1, 1, 1), (17, 33, 5), (300, 8, 24), (40, 16, 4634)],
        [(300, 8, 24), (5, 8, 16), (1, 1, 1), (517, 32, 64), (17, 3, 5), (256, 9, 33)]],
    1: [[(5, 8, 300), (24, 24, 311), (1, 1, 1), (17, 33, 5), (300, 8, 24), (40, 16, 4634)],
        [(24, 24, 311), (5, 8, 300), (1, 1, 1), (32, 32, 1000), (17, 3, 5), (8, 24, 33)]],
}


@pytest.mark.parametrize("small", [1, 0])
def test_zgemm_batch_vs_numpy(small):
    """pw_zgemm_batch_dev against numpy (1e-12), with the small-product kernels allowed (1)
    and not (0), on batches that fit them and batches that do not."""
    import pwdft_hip as ph
    from pwdft_hip._lib import PW_TUNE_BATCH_GEMM_SMALL
    ctx = ph.Context(0)
    ctx.call("pw_tune_set", PW_TUNE_BATCH_GEMM_SMALL, small)
    rng = np.random.default_rng(4)
    for transa, shapes in [(t, sh) for t in (0, 1) for sh in ZGEMM_BATCHES[t]]:
        print(f"case small={small} transa={transa} shapes={shapes}", flush=True)
        As, Bs, Cs, refs = [], [], [], []
        alpha = np.array([0.7, -0.2])
        beta = np.array([0.3, 0.1])
        for m, n, k in shapes:
            A = rand_psi(rng, k, m) if transa else rand_psi(rng, m, k)
            B = rand_psi(rng, k, n)
            C = rand_psi(rng, m, n)
            opA = A.conj().T if transa else A
            refs.append(complex(*alpha) * opA @ B + complex(*beta) * C)
            As.append(ph.DeviceMatrix.from_host(ctx, A))
            Bs.append(ph.DeviceMatrix.from_host(ctx, B))
            Cs.append(ph.DeviceMatrix.from_host(ctx, C))
        arr = lambda xs: (ctypes.c_void_p * len(xs))(*[x.p.value for x in xs])
        i32 = lambda xs: np.ascontiguousarray(xs, dtype=np.int32)
        i64 = lambda xs: np.ascontiguousarray(xs, dtype=np.int64)
        ms, ns, ks = (i32([s[i] for s in shapes]) for i in range(3))
        la, lb, lc = i64([a.rows for a in As]), i64([b.rows for b in Bs]), i64([c.rows for c in Cs])
        from pwdft_hip._lib import ptr
        ctx.call("pw_zgemm_batch_dev", transa, len(shapes), ptr(ms), ptr(ns), ptr(ks), ptr(alpha), arr(As),
                 ptr(la), arr(Bs), ptr(lb), ptr(beta), arr(Cs), ptr(lc))
        for Cd, ref, s in zip(Cs, refs, shapes):
            assert np.abs(Cd.download() - ref).max() < 1e-12 * max(1.0, np.abs(ref).max()), (transa, s)


@pytest.mark.parametrize("n,nbatch", [(1, 3), (8, 182), (24, 182), (40, 5)])
def test_heevd_batch_vs_numpy(n, nbatch):
    """pw_heevd_batch_dev: the eigenvalues of every stacked Hermitian matrix equal numpy's,
    the eigenvectors are orthonormal and satisfy A v = l v (only the upper triangle is read:
    the strict lower triangle holds garbage)."""
    import pwdft_hip as ph
    from pwdft_hip._lib import Context, DeviceMatrix, PWError, ptr
    c = Context(0)
    rng = np.random.default_rng(n * 100 + nbatch)
    A = rng.standard_normal((nbatch, n, n)) + 1j * rng.standard_normal((nbatch, n, n))
    A = 0.5 * (A + np.conj(np.swapaxes(A, 1, 2)))
    dev = np.array(A)
    low = np.tril_indices(n, -1)
    for k in range(nbatch):
        dev[k][low] = 7.0 + 3.0j             # must be ignored
    buf = np.ascontiguousarray(np.transpose(dev, (0, 2, 1)))   # column-major blocks
    d = DeviceMatrix.from_host(c, buf.reshape(-1, 1))
    w = np.empty((nbatch, n))
    print(f"case heevd_batch n={n} nbatch={nbatch}", flush=True)
    c.call("pw_heevd_batch_dev", n, nbatch, d.p, n, n * n, ptr(w))
    V = np.transpose(d.download().reshape(nbatch, n, n), (0, 2, 1))
    ref = np.linalg.eigvalsh(A)
    assert np.abs(w - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())
    eye = np.eye(n)
    for k in range(nbatch):
        assert np.abs(V[k].conj().T @ V[k] - eye).max() < 1e-12
        assert np.abs(A[k] @ V[k] - V[k] * w[k][None, :]).max() < 1e-11 * max(1.0, np.abs(ref).max())
    with pytest.raises(PWError):
        c.call("pw_heevd_batch_dev", n, nbatch, d.p, n, n * n - 1, ptr(w))
    d.free()
    c.close()


def test_heevd_batch_graded_spectrum():
    """Matrices like the Rayleigh-Ritz matrices of a batched LOBPCG step: eigenvalues near
    zero beside ones of order 1-10 (so the diagonal products that scale the Jacobi threshold
    are tiny for some pairs).  The batched eigensolver converges (a non-converged matrix
    raises) and matches numpy to the same tolerance as random spectra."""
    from pwdft_hip._lib import Context, DeviceMatrix, ptr
    c = Context(0)
    rng = np.random.default_rng(77)
    n, nbatch = 24, 64
    lam = np.concatenate([[-3e-7, 1e-9, 2e-6, -1e-4], np.linspace(-0.3, 0.8, 12), np.geomspace(1.0, 9.0, 8)])
    A = np.empty((nbatch, n, n), dtype=complex)
    for k in range(nbatch):
        Q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
        A[k] = (Q * rng.permutation(lam)[None, :]) @ Q.conj().T
        A[k] = 0.5 * (A[k] + A[k].conj().T)
    buf = np.ascontiguousarray(np.transpose(A, (0, 2, 1)))
    d = DeviceMatrix.from_host(c, buf.reshape(-1, 1))
    w = np.empty((nbatch, n))
    print("case heevd_batch graded spectrum", flush=True)
    c.call("pw_heevd_batch_dev", n, nbatch, d.p, n, n * n, ptr(w))
    V = np.transpose(d.download().reshape(nbatch, n, n), (0, 2, 1))
    ref = np.linalg.eigvalsh(A)
    assert np.abs(w - ref).max() < 1e-12 * np.abs(ref).max()
    for k in range(nbatch):
        assert np.abs(V[k].conj().T @ V[k] - np.eye(n)).max() < 1e-12
        assert np.abs(A[k] @ V[k] - V[k] * w[k][None, :]).max() < 1e-11 * np.abs(ref).max()
    d.free()
    c.close()


@pytest.mark.parametrize("n,nbatch", [(1, 3), (8, 182), (24, 182), (40, 5), (100, 1), (160, 1), (480, 1), (1000, 1)])
def test_chol_inv_batch_vs_numpy(n, nbatch):
    """pw_chol_inv_batch_dev: every block becomes inv(cholesky(Hermitian(A)).U) of numpy's
    factor (upper triangle read, strict lower triangle zeroed; one matrix above order 64 by
    the recursive 2 x 2 block form); a batch holding an indefinite matrix raises
    PosDefException naming it.  pw_herm_fix_batch_dev modes 0 and 2 against
    numpy's Hermitian() and (A + A^H)/2."""
    from pwdft_hip._lib import Context, DeviceMatrix, PosDefException
    c = Context(0)
    rng = np.random.default_rng(n * 7 + nbatch)
    Z = rng.standard_normal((nbatch, n, 2 * n)) + 1j * rng.standard_normal((nbatch, n, 2 * n))
    A = Z @ np.conj(np.swapaxes(Z, 1, 2)) / (2 * n) + 0.1 * np.eye(n)
    dev = np.array(A)
    low = np.tril_indices(n, -1)
    for k in range(nbatch):
        dev[k][low] = 7.0 + 3.0j             # must be ignored
    col = lambda M: np.ascontiguousarray(np.transpose(M, (0, 2, 1))).reshape(-1, 1)
    back = lambda d: np.transpose(d.download().reshape(nbatch, n, n), (0, 2, 1))
    d = DeviceMatrix.from_host(c, col(dev))
    print(f"case chol_inv_batch n={n} nbatch={nbatch}", flush=True)
    c.call("pw_chol_inv_batch_dev", n, nbatch, d.p, n, n * n)
    Ui = back(d)
    U = np.conj(np.swapaxes(np.linalg.cholesky(A), 1, 2))
    ref = np.linalg.inv(U)
    scale = np.abs(ref).max()
    assert np.abs(Ui - ref).max() < 1e-11 * scale
    assert np.all(Ui[:, low[0], low[1]] == 0)
    assert np.abs(U @ Ui - np.eye(n)).max() < 1e-11 * scale
    if n > 1:
        bad = np.array(dev)
        bad[min(2, nbatch - 1)][0, 0] = -1.0
        d.upload(col(bad).ravel())
        with pytest.raises(PosDefException, match=f"matrix {min(2, nbatch - 1)}"):
            c.call("pw_chol_inv_batch_dev", n, nbatch, d.p, n, n * n)
    G = rng.standard_normal((nbatch, n, n)) + 1j * rng.standard_normal((nbatch, n, n))
    up = np.triu(G, 1)
    herm = up + np.conj(np.swapaxes(up, 1, 2))
    idx = np.arange(n)
    herm[:, idx, idx] = G[:, idx, idx].real
    for mode, want in ((0, herm), (2, 0.5 * (G + np.conj(np.swapaxes(G, 1, 2))))):
        d.upload(col(G).ravel())
        c.call("pw_herm_fix_batch_dev", n, nbatch, d.p, n, n * n, mode)
        assert np.abs(back(d) - want).max() < 1e-15 * max(1.0, np.abs(G).max()), mode
    d.free()
    c.close()


@pytest.mark.parametrize("n,nbatch", [(8, 182), (24, 182)])
def test_async_dense_batch_matches_sync(n, nbatch):
    """pw_chol_inv_batch_async_dev and pw_heevd_batch_async_dev give the synchronous calls'
    results bit for bit and leave their per-matrix status in device memory: zero for
    positive definite input, nonzero for the matrix made indefinite."""
    from pwdft_hip._lib import Context, DeviceMatrix, ptr
    c = Context(0)
    rng = np.random.default_rng(n + 3 * nbatch)
    Z = rng.standard_normal((nbatch, n, 2 * n)) + 1j * rng.standard_normal((nbatch, n, 2 * n))
    A = Z @ np.conj(np.swapaxes(Z, 1, 2)) / (2 * n) + 0.1 * np.eye(n)
    col = lambda M: np.ascontiguousarray(np.transpose(M, (0, 2, 1))).reshape(-1, 1)
    d1, d2 = DeviceMatrix.from_host(c, col(A)), DeviceMatrix.from_host(c, col(A))
    info = DeviceMatrix(c, nbatch, 1)                    # 4 * nbatch ints fit in 16 * nbatch bytes
    wdev = DeviceMatrix(c, (n * nbatch + 1) // 2, 1)
    get_info = lambda m: (lambda h: (c.call("pw_memcpy_d2h", ptr(h), info.p, h.nbytes), h)[1])(np.empty(m, np.int32))
    print(f"case async dense n={n} nbatch={nbatch}", flush=True)
    c.call("pw_chol_inv_batch_dev", n, nbatch, d1.p, n, n * n)
    c.call("pw_chol_inv_batch_async_dev", n, nbatch, d2.p, n, n * n, info.p)
    assert np.array_equal(d1.download(), d2.download())
    assert not get_info(2 * nbatch).any()
    w = np.empty((nbatch, n))
    d1.upload(col(A).ravel())
    d2.upload(col(A).ravel())
    c.call("pw_heevd_batch_dev", n, nbatch, d1.p, n, n * n, ptr(w))
    c.call("pw_heevd_batch_async_dev", n, nbatch, d2.p, n, n * n, wdev.p, info.p)
    w2 = np.empty((nbatch, n))
    c.call("pw_memcpy_d2h", ptr(w2), wdev.p, w2.nbytes)
    assert np.array_equal(w, w2) and np.array_equal(d1.download(), d2.download())
    assert not get_info(nbatch).any()
    bad = np.array(A)
    bad[5][0, 0] = -1.0
    d2.upload(col(bad).ravel())
    c.call("pw_chol_inv_batch_async_dev", n, nbatch, d2.p, n, n * n, info.p)
    st = get_info(2 * nbatch)
    assert st[5] != 0 and np.count_nonzero(st[:nbatch]) == 1
    for m in (d1, d2, info, wdev):
        m.free()
    c.close()


def test_lobpcg_batch_vs_per_k(al):
    """diag_LOBPCG! over all 182 k-points in lockstep gives the per-k solver's eigenvalues."""
    import pwdft_hip as ph
    from pwdft_hip.solvers import diag_LOBPCG_
    H, _ = al
    print("case Al 182k LOBPCG batch", flush=True)
    ps = ph.rand_BlochWavefunc(H, seed=8)
    ev, st = ph.diag_LOBPCG_batch(H, ps, tol=1e-10, NiterMax=200)
    assert ev.shape == (8, 182)
    ps2 = ph.rand_BlochWavefunc(H, seed=8)
    for ik in (0, 1, 90, 181):
        H.ik = ik
        ev2, _ = diag_LOBPCG_(H, ps2[ik], tol=1e-10, NiterMax=200)
        assert np.abs(ev[:, ik] - ev2).max() < 1e-8, (ik, ev[:, ik], ev2)
        X = ps[ik].download()
        assert np.abs(X.conj().T @ X - np.eye(8)).max() < 1e-10


def test_lobpcg_batch_lock_branch(al):
    """The batched solver's lock path (a k-point that reaches the lock branch finishes the
    iteration through the per-k lock step, with device blocks reused across k-points): with
    the lock threshold raised from 10 eps to 1e-8 the k-points lock on the way to
    tol = 1e-11, and the eigenvalues still equal the per-k solver's (no lock) to the accuracy
    a locked band keeps (1e-6); the Ritz vectors stay orthonormal."""
    import pwdft_hip as ph
    import pwdft_hip.batch as BT
    import pwdft_hip.solvers as PS
    H, _ = al
    iks = list(range(0, 182, 9))
    print(f"case Al LOBPCG batch lock branch, {len(iks)} k-points", flush=True)
    ps = ph.rand_BlochWavefunc(H, seed=9)
    sub = [ps[k] for k in iks]
    saved = (PS.EPS, BT.EPS)
    PS.EPS = BT.EPS = 1e-9
    try:
        ev, st = ph.diag_LOBPCG_batch(H, sub, tol=1e-11, NiterMax=200, iks=iks)
    finally:
        PS.EPS, BT.EPS = saved
    assert sum(x["locks"] for x in st) >= 1, st
    ps2 = ph.rand_BlochWavefunc(H, seed=9)
    for j, ik in enumerate(iks[::5]):
        H.ik = ik
        ev2, _ = PS.diag_LOBPCG_(H, ps2[ik], tol=1e-11, NiterMax=200)
        assert np.abs(ev[:, iks.index(ik)] - ev2).max() < 1e-6, (ik, ev[:, iks.index(ik)], ev2)
    for m in sub:
        X = m.download()
        assert np.abs(X.conj().T @ X - np.eye(8)).max() < 1e-10


@pytest.mark.parametrize("mixed", [False, True])
def test_lobpcg_batch_graph_replay_vs_eager(al, mixed):
    """The batched LOBPCG's phases replayed from captured hipGraphs (within a call, and from
    graphs kept by the context across calls) give the eager run's Ritz values and vectors
    bit for bit."""
    import pwdft_hip as ph
    from pwdft_hip._lib import PW_TUNE_LOBPCG_GRAPH
    H, _ = al
    iks = list(range(0, 182, 3))
    runs = []
    for graph, reps in ((0, 1), (1, 3)):
        H.ctx.call("pw_tune_set", PW_TUNE_LOBPCG_GRAPH, graph)
        for rep in range(reps):
            print(f"case Al LOBPCG batch graph={graph} rep={rep} mixed={mixed}", flush=True)
            ps = ph.rand_BlochWavefunc(H, seed=12)
            sub = [ps[k] for k in iks]
            # a bounded number of iterations: run to full convergence, the mixed path's P
            # block can turn rank-deficient (Cholesky of P^H P fails, as the reference's
            # cholesky() would)
            ev, st = ph.diag_LOBPCG_batch(H, sub, tol=1e-9, NiterMax=12, iks=iks, mixed=mixed)
            runs.append((ev, [m.download() for m in sub], st))
    H.ctx.call("pw_tune_set", PW_TUNE_LOBPCG_GRAPH, 1)
    ev0, X0, st0 = runs[0]
    assert max(x["iter"] for x in st0) > 3
    for ev, X, st in runs[1:]:
        assert np.array_equal(ev, ev0)
        assert st == st0
        for a, b in zip(X, X0):
            assert np.array_equal(a, b)


def test_descriptor_cache_beyond_its_slots():
    """More distinct batched-GEMM descriptor arrays than the cache has slots (256), first
    with no captured graph (slots are reused after a wait) and then while a captured graph
    lives (slots are added, none the graph reads is rewritten): every product equals numpy's,
    and the graph's replay still computes its own product."""
    import pwdft_hip as ph
    from pwdft_hip._lib import ptr
    ctx = ph.Context(0)
    rng = np.random.default_rng(17)
    m, n, k = 6, 5, 40
    A = rand_psi(rng, k, 320 * m)       # 320 distinct column offsets -> 320 distinct arrays
    B = rand_psi(rng, k, n)
    dA, dB = ph.DeviceMatrix.from_host(ctx, A), ph.DeviceMatrix.from_host(ctx, B)
    dC = ph.DeviceMatrix(ctx, m, n)
    one, zero = np.array([1.0, 0.0]), np.array([0.0, 0.0])
    # the size arrays stay referenced for the whole call (a temporary's buffer may be
    # freed, and reused, before the library reads it)
    mm, nn, kk = (np.array([v], dtype=np.int32) for v in (m, n, k))
    lda, ldc = np.array([k], dtype=np.int64), np.array([m], dtype=np.int64)

    def product(j):
        pa = np.ascontiguousarray([dA.p.value + 16 * k * m * j], dtype=np.uint64)
        pb = np.ascontiguousarray([dB.p.value], dtype=np.uint64)
        pc = np.ascontiguousarray([dC.p.value], dtype=np.uint64)
        ctx.call("pw_zgemm_batch_dev", 1, 1, ptr(mm), ptr(nn), ptr(kk), ptr(one), ptr(pa), ptr(lda),
                 ptr(pb), ptr(lda), ptr(zero), ptr(pc), ptr(ldc))

    def ref(j):
        return A[:, j * m:(j + 1) * m].conj().T @ B

    print("case descriptor cache without a graph", flush=True)
    for j in range(300):
        product(j)
        if j % 37 == 0 or j >= 290:
            assert np.abs(dC.download() - ref(j)).max() < 1e-12 * np.abs(ref(j)).max(), j
    print("case descriptor cache with a live graph", flush=True)
    product(7)                          # its descriptor is cached before the capture
    with ph.Graph(ctx) as g:
        product(7)
    for j in range(300, 320):
        product(j)
        assert np.abs(dC.download() - ref(j)).max() < 1e-12 * np.abs(ref(j)).max(), j
    for j in range(0, 300, 7):
        product(j)
    ctx.call("pw_synchronize")
    dC.upload(np.zeros(m * n, dtype=complex))
    g.launch()
    ctx.call("pw_synchronize")
    assert np.abs(dC.download() - ref(7)).max() < 1e-12 * np.abs(ref(7)).max()
    g.free()
    ctx.close()


def test_scf_batched_vs_per_k():
    """Al fcc 4x4x4 shifted (10 k-points, Fermi-Dirac smearing 0.01 Ha): the SCF with the
    batched eigensolver converges to the per-k SCF's free energy."""
    import pwdft_hip as ph
    print("case Al 10k SCF batched vs per-k", flush=True)
    E = []
    for batched in (False, True):
        atoms = ph.Atoms(["Al"], [[0, 0, 0]], ph.gen_lattice_fcc(7.6))
        H = ph.Hamiltonian(atoms, ["Al-q3.lda.gth"], 15.0, meshk=(4, 4, 4), shiftk=(1, 1, 1), Nstates=8,
                           use_smearing=True, smearing_kT=0.01)
        out = ph.KS_solve_SCF_(H, betamix=0.5, use_smearing=True, kT=0.01, etot_conv_thr=1e-9,
                               ethr_evals_last=1e-10, batched=batched)
        E.append(out["energies"]["Total"])
    assert abs(E[0] - E[1]) < 1e-8, E


@pytest.mark.parametrize("mixed", [False, True])
def test_op_H_batch_graph_replay(al, mixed):
    """The batched op_H captured into a hipGraph and replayed: every replay, with eager
    launches and synchronisations in between, reproduces the eager result bit for bit."""
    import pwdft_hip as ph
    H, host = al
    print(f"case Al 182k op_H batch graph mixed={mixed}", flush=True)
    dev = [ph.DeviceMatrix.from_host(H.ctx, a) for a in host]
    H.mixed = mixed
    try:
        eager = [m.download() for m in ph.op_H_batch(H, dev)]
        out = [ph.DeviceMatrix(H.ctx, a.shape[0], a.shape[1]) for a in host]
        ph.op_H_batch(H, dev, out=out)        # warm: workspaces and descriptors exist
        H.ctx.call("pw_synchronize")
        for o in out:                         # poison the outputs before the replays
            o.upload(np.full(o.rows * o.cols, np.nan + 0j))
        with ph.Graph(H.ctx) as g:
            ph.op_H_batch(H, dev, out=out)
        for rep in range(3):
            g.launch()
            H.ctx.call("pw_synchronize")
            for ik in (0, 91, 181):
                assert np.array_equal(out[ik].download(), eager[ik]), (rep, ik)
            ph.op_H_batch(H, dev[:5], iks=range(5))   # eager work between replays
            H.ctx.call("pw_synchronize")
            for o in out:
                o.upload(np.zeros(o.rows * o.cols, dtype=complex))
        g.free()
    finally:
        H.mixed = False


@pytest.mark.parametrize("mixed", [False, True])
def test_op_H_batch_fused_vs_three_stage(al, mixed):
    """The fused launches (one workgroup per (k-point, band): the register-line kernel on the
    20^3 box compressed to the active x, at 256 and 512 threads, and the kernel with the whole
    box in LDS) and the three-stage launches (column / plane / column kernels with the
    intermediate in HBM, batched GEMMs for the nonlocal term) give the same op_H, per term
    and accumulated."""
    import pwdft_hip as ph
    from pwdft_hip._lib import (PW_ACCUMULATE, PW_MIXED, PW_TERM_ALL, PW_TERM_K, PW_TERM_NLOC, PW_TERM_VLOC,
                                PW_TUNE_FUSED_BATCH, PW_TUNE_FUSED_NT)
    H, host = al
    dev = [ph.DeviceMatrix.from_host(H.ctx, a) for a in host]
    base = [ph.DeviceMatrix.from_host(H.ctx, 0.5 * a) for a in host]
    tol = 2e-5 if mixed else 1e-13
    modes = [("reg", 1, 0), ("reg512", 1, 512), ("box", 2, 0), ("three-stage", 0, 0)]
    try:
        for terms in (PW_TERM_ALL, PW_TERM_VLOC, PW_TERM_K | PW_TERM_VLOC, PW_TERM_NLOC):
            for acc in (0, PW_ACCUMULATE):
                res = {}
                for name, fused, nt in modes:
                    H.ctx.call("pw_tune_set", PW_TUNE_FUSED_BATCH, fused)
                    H.ctx.call("pw_tune_set", PW_TUNE_FUSED_NT, nt)
                    out = [ph.DeviceMatrix.from_host(H.ctx, b.download()) for b in base]
                    print(f"case {name} terms={terms} acc={acc} mixed={mixed}", flush=True)
                    ph.op_H_batch(H, dev, out=out, terms=terms | (PW_MIXED if mixed else 0), accumulate=bool(acc))
                    res[name] = [o.download() for o in out]
                for name in ("reg", "reg512", "box"):
                    worst = max(np.abs(a - b).max() / np.abs(b).max() for a, b in zip(res[name], res["three-stage"]))
                    assert worst < tol, (name, terms, acc, worst)
    finally:
        H.ctx.call("pw_tune_set", PW_TUNE_FUSED_BATCH, 1)
        H.ctx.call("pw_tune_set", PW_TUNE_FUSED_NT, 0)


@pytest.mark.parametrize("n,bad", [(480, 300), (160, 70), (100, 0)])
def test_chol_inv_recursive_failure_index_and_rocsolver(n, bad):
    """The recursive Cholesky inverse reports the failing leading minor counted over the
    whole matrix (a leaf deep in the recursion), as rocSOLVER's ZPOTRF does
    (PW_TUNE_CHOL_REC = 0), and both give the same inverse factor of a definite matrix."""
    from pwdft_hip._lib import Context, DeviceMatrix, PosDefException, PW_TUNE_CHOL_REC
    c = Context(0)
    rng = np.random.default_rng(n + bad)
    Z = rng.standard_normal((n, 2 * n)) + 1j * rng.standard_normal((n, 2 * n))
    A = Z @ Z.conj().T / (2 * n) + 0.1 * np.eye(n)
    d = DeviceMatrix.from_host(c, A)
    out = []
    for mode in (1, 0):
        c.call("pw_tune_set", PW_TUNE_CHOL_REC, mode)
        try:
            d.upload(A)
            c.call("pw_chol_inv_batch_dev", n, 1, d.p, n, n * n)
            out.append(d.download())
            B = np.array(A)
            B[bad, bad] = -1.0
            d.upload(B)
            print(f"case chol_inv recursive failure n={n} bad={bad} mode={mode}", flush=True)
            with pytest.raises(PosDefException, match=f"leading minor {bad + 1} "):
                c.call("pw_chol_inv_batch_dev", n, 1, d.p, n, n * n)
        finally:
            c.call("pw_tune_set", PW_TUNE_CHOL_REC, 1)
    ref = np.linalg.inv(np.linalg.cholesky(A).conj().T)
    scale = np.abs(ref).max()
    assert np.abs(out[0] - ref).max() < 1e-11 * scale
    assert np.abs(out[0] - out[1]).max() < 1e-11 * scale
    d.free()
    c.close()
