"""The device Hermitian eigensolver (tridiagonal reduction with the matrix in the LDS of a
persistent cooperative launch; the tridiagonal eigenpairs by rocSOLVER's divide and conquer
or, for the lowest ones, by the device bisection / inverse-iteration kernels; reflectors
applied to the eigenvectors) through pw_heevd_dev and pw_heev_lowest_dev, against numpy's
eigh on random, degenerate, clustered, diagonal and graded Hermitian matrices of every order
class the eigensolvers meet; and against rocSOLVER's ZHEEVD (PW_TUNE_EIG_PATH = 1)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PW_TUNE_EIG_PATH = 9
PW_TUNE_EIG_SYNC = 14
PW_TUNE_EIG_WG = 17


@pytest.fixture(scope="module")
def ctx():
    import pwdft_hip as ph
    return ph.Context(0)


def _herm(rng, n, kind):
    if kind == "random":
        a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        return a + a.conj().T
    q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    if kind == "degenerate":
        lam = np.repeat(rng.standard_normal((n + 2) // 3), 3)[:n]
    elif kind == "graded":
        lam = np.sign(rng.standard_normal(n)) * np.logspace(-8, 2, n)
    else:
        lam = rng.standard_normal(n)
    if kind == "diagonal":
        return np.diag(lam).astype(complex)
    return (q * lam) @ q.conj().T


def _eig(ctx, A, path=0):
    import ctypes
    from pwdft_hip._lib import DeviceMatrix, ptr
    n = A.shape[0]
    # only the upper triangle is read: poison the strict lower one
    Au = np.triu(A) + np.tril(np.full((n, n), np.nan + 1j * np.nan), -1)
    dA = DeviceMatrix.from_host(ctx, Au)
    w = np.zeros(n)
    ctx.call("pw_tune_set", PW_TUNE_EIG_PATH, path)
    try:
        ctx.call("pw_heevd_dev", n, dA.p, n, w.ctypes.data_as(ctypes.c_void_p))
    finally:
        ctx.call("pw_tune_set", PW_TUNE_EIG_PATH, 0)
    return w, dA.download()


@pytest.mark.parametrize("n", [2, 3, 8, 33, 100, 320, 480, 777, 960])
@pytest.mark.parametrize("kind", ["random", "degenerate", "graded", "diagonal"])
def test_heevd_device_vs_numpy(ctx, n, kind):
    rng = np.random.default_rng(n * 31 + len(kind))
    A = _herm(rng, n, kind)
    print(f"case heevd n={n} kind={kind}", flush=True)
    w, X = _eig(ctx, A)
    ref = np.linalg.eigvalsh(A)
    scale = max(1.0, np.abs(ref).max())
    assert np.abs(w - ref).max() < 1e-12 * scale * max(1.0, np.sqrt(n) / 4), np.abs(w - ref).max()
    assert np.abs(X.conj().T @ X - np.eye(n)).max() < 1e-12 * max(1.0, np.sqrt(n) / 4)
    assert np.abs(A @ X - X * w).max() < 1e-12 * scale * max(1.0, np.sqrt(n) / 4)


def test_heevd_values_only_and_rocsolver_path(ctx):
    import ctypes
    from pwdft_hip._lib import DeviceMatrix
    rng = np.random.default_rng(5)
    A = _herm(rng, 480, "random")
    w0, X0 = _eig(ctx, A, path=0)
    w1, X1 = _eig(ctx, A, path=1)
    assert np.abs(w0 - w1).max() < 1e-11 * np.abs(w1).max()
    # same eigenvectors up to a phase per column (the spectrum is simple)
    ov = np.abs(np.einsum("ij,ij->j", X0.conj(), X1))
    assert np.abs(ov - 1).max() < 1e-9
    dA = DeviceMatrix.from_host(ctx, A)
    wv = np.zeros(480)
    ctx.call("pw_heev_values_dev", 480, dA.p, 480, wv.ctypes.data_as(ctypes.c_void_p))
    assert np.abs(wv - w1).max() < 1e-11 * np.abs(w1).max()


@pytest.mark.parametrize("n", [33, 160, 480, 960])
def test_heevd_fenced_exchanges(ctx, n):
    """The reduction with two exchanges per column, ordered by agent-scope fences
    (PW_TUNE_EIG_SYNC = 0) or as device-coherent atomics (1), and with one exchange per
    column (2), against numpy and against the default choice by order (3)."""
    rng = np.random.default_rng(n + 7)
    A = _herm(rng, n, "random")
    w0, _ = _eig(ctx, A)
    ref = np.linalg.eigvalsh(A)
    scale = max(1.0, np.abs(ref).max())
    for mode in (0, 1, 2):
        print(f"case heevd n={n} exchange mode {mode}", flush=True)
        ctx.call("pw_tune_set", PW_TUNE_EIG_SYNC, mode)
        try:
            w, X = _eig(ctx, A)
        finally:
            ctx.call("pw_tune_set", PW_TUNE_EIG_SYNC, 3)
        assert np.abs(w - ref).max() < 1e-12 * scale * max(1.0, np.sqrt(n) / 4)
        assert np.abs(w - w0).max() < 1e-12 * scale * max(1.0, np.sqrt(n) / 4)
        assert np.abs(X.conj().T @ X - np.eye(n)).max() < 1e-12 * max(1.0, np.sqrt(n) / 4)
        assert np.abs(A @ X - X * w).max() < 1e-12 * scale * max(1.0, np.sqrt(n) / 4)


@pytest.mark.parametrize("n", [100, 480])
def test_heevd_workgroup_counts(ctx, n):
    """The reduction shared by the fewest workgroups that fit (PW_TUNE_EIG_WG = 0), by 32 and
    by 128, against numpy and against the default (64)."""
    rng = np.random.default_rng(n + 11)
    A = _herm(rng, n, "degenerate")
    w0, _ = _eig(ctx, A)
    ref = np.linalg.eigvalsh(A)
    scale = max(1.0, np.abs(ref).max())
    for wg in (0, 32, 128):
        print(f"case heevd n={n} workgroups >= {wg}", flush=True)
        ctx.call("pw_tune_set", PW_TUNE_EIG_WG, wg)
        try:
            w, X = _eig(ctx, A)
        finally:
            ctx.call("pw_tune_set", PW_TUNE_EIG_WG, 64)
        assert np.abs(w - ref).max() < 1e-12 * scale * max(1.0, np.sqrt(n) / 4)
        assert np.abs(w - w0).max() < 1e-12 * scale * max(1.0, np.sqrt(n) / 4)
        assert np.abs(X.conj().T @ X - np.eye(n)).max() < 1e-12 * max(1.0, np.sqrt(n) / 4)
        assert np.abs(A @ X - X * w).max() < 1e-12 * scale * max(1.0, np.sqrt(n) / 4)


PW_TUNE_EIG_TRI = 18


def _herm_clustered(rng, n, kind):
    """Spectra with runs of close eigenvalues: 'cluster5' five-fold runs 1e-10 apart,
    'cluster40' one run of 40 exactly equal values (wider than one Rayleigh-Ritz chunk) among
    simple ones, 'near_identity' I + 1e-9 H (a single run of all n)."""
    q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    if kind == "cluster5":
        base = np.sort(rng.standard_normal((n + 4) // 5))
        lam = (np.repeat(base, 5) + 1e-10 * np.tile(np.arange(5), len(base)))[:n]
    elif kind == "cluster40":
        lam = np.sort(rng.standard_normal(n))
        lam[10:50] = lam[10]
    else:
        a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        return np.eye(n) + 1e-9 * (a + a.conj().T)
    return (q * lam) @ q.conj().T


def _lowest(ctx, A, m):
    import ctypes
    from pwdft_hip._lib import DeviceMatrix
    n = A.shape[0]
    Au = np.triu(A) + np.tril(np.full((n, n), np.nan + 1j * np.nan), -1)
    dA = DeviceMatrix.from_host(ctx, Au)
    w = np.zeros(m)
    ctx.call("pw_heev_lowest_dev", n, m, dA.p, n, w.ctypes.data_as(ctypes.c_void_p))
    return w, dA.download()[:, :m]


def _check_pairs(A, w, X, ref, n):
    m = len(w)
    scale = max(1.0, np.abs(ref).max())
    tol = 1e-12 * max(1.0, np.sqrt(n) / 4)
    assert np.abs(w - ref[:m]).max() < tol * scale, np.abs(w - ref[:m]).max()
    assert np.abs(X.conj().T @ X - np.eye(m)).max() < tol, np.abs(X.conj().T @ X - np.eye(m)).max()
    assert np.abs(A @ X - X * w).max() < tol * scale, np.abs(A @ X - X * w).max()


@pytest.mark.parametrize("n", [65, 160, 320, 480, 960])
@pytest.mark.parametrize("kind", ["random", "degenerate", "graded", "diagonal", "cluster5"])
def test_heev_lowest_vs_numpy(ctx, n, kind):
    """pw_heev_lowest_dev (the Rayleigh-Ritz eigensolve of the block solvers: reduction, then
    the lowest third of the tridiagonal eigenpairs on the device by multisection bisection,
    inverse iteration and orthogonalisation within runs of close eigenvalues) against numpy."""
    rng = np.random.default_rng(n * 17 + len(kind))
    A = _herm_clustered(rng, n, kind) if kind.startswith("cluster") else _herm(rng, n, kind)
    ref = np.linalg.eigvalsh(A)
    for m in (n // 3, n):
        print(f"case heev_lowest n={n} m={m} kind={kind}", flush=True)
        w, X = _lowest(ctx, A, m)
        _check_pairs(A, w, X, ref, n)


@pytest.mark.parametrize("n,kind", [(100, "near_identity"), (160, "cluster40"), (480, "cluster40"),
                                    (480, "degenerate"), (320, "graded")])
def test_heevd_device_tridiagonal_all_pairs(ctx, n, kind):
    """pw_heevd_dev with every tridiagonal eigenpair from the device kernels (PW_TUNE_EIG_TRI =
    2), including runs of close eigenvalues longer than one Rayleigh-Ritz chunk, against
    numpy and against ZSTEDC (the default for pw_heevd_dev)."""
    rng = np.random.default_rng(n + len(kind))
    A = _herm_clustered(rng, n, kind) if kind in ("near_identity", "cluster40") else _herm(rng, n, kind)
    ref = np.linalg.eigvalsh(A)
    w0, _ = _eig(ctx, A)
    print(f"case heevd device tridiagonal n={n} kind={kind}", flush=True)
    ctx.call("pw_tune_set", PW_TUNE_EIG_TRI, 2)
    try:
        w, X = _eig(ctx, A)
    finally:
        ctx.call("pw_tune_set", PW_TUNE_EIG_TRI, 1)
    _check_pairs(A, w, X, ref, n)
    assert np.abs(w - w0).max() < 1e-12 * max(1.0, np.abs(ref).max()) * max(1.0, np.sqrt(n) / 4)


def test_heev_lowest_zstedc_mode(ctx):
    """PW_TUNE_EIG_TRI = 0 routes pw_heev_lowest_dev through ZSTEDC: same eigenpairs."""
    rng = np.random.default_rng(3)
    A = _herm(rng, 320, "random")
    ref = np.linalg.eigvalsh(A)
    w1, X1 = _lowest(ctx, A, 107)
    ctx.call("pw_tune_set", PW_TUNE_EIG_TRI, 0)
    try:
        w0, X0 = _lowest(ctx, A, 107)
    finally:
        ctx.call("pw_tune_set", PW_TUNE_EIG_TRI, 1)
    _check_pairs(A, w0, X0, ref, 320)
    _check_pairs(A, w1, X1, ref, 320)
    ov = np.abs(np.einsum("ij,ij->j", X0.conj(), X1))
    assert np.abs(ov - 1).max() < 1e-9
