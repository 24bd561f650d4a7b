"""The device Hermitian eigensolver (tridiagonal reduction with the matrix in the LDS of a
persistent cooperative launch, rocSOLVER's divide and conquer on the tridiagonal form,
reflectors applied to the eigenvectors) through pw_heevd_dev, against numpy's eigh on
random, degenerate, diagonal and graded Hermitian matrices of every order class the
eigensolvers meet; and against rocSOLVER's ZHEEVD (PW_TUNE_EIG_PATH = 1)."""
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
