"""Pins of the oracle's CheFSI (oracle/solvers.py: chebyfilt, get_ub_lb_lanczos,
diag_CheFSI; CheFSI.jl:1-130) on a dense random Hermitian H (n = 60):

- chebyfilt equals the closed form T_m((H - c)/e) X / T_m(-2) evaluated on the eigenpairs of
  H, with T_m(x) = cos(m acos x) inside [-1, 1] and sign(x)^m cosh(m acosh |x|) outside;
- the Lanczos bounds bracket what the filter needs: lb inside the spectrum, ub above its
  top;
- repeated diag_CheFSI (Chebyshev-filtered subspace iteration) converges to the lowest
  eigenvalues of H from numpy's eigh."""
import numpy as np

from oracle.solvers import chebyfilt, diag_CheFSI, get_ub_lb_lanczos, ortho_sqrt


def _H(n=60, seed=0):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    H = (A + A.conj().T) / 2
    return H, np.linalg.eigh(H), rng


def _T(m, x):
    x = np.asarray(x, dtype=float)
    inside = np.cos(m * np.arccos(np.clip(x, -1.0, 1.0)))
    outside = np.sign(x) ** m * np.cosh(m * np.arccosh(np.maximum(np.abs(x), 1.0)))
    return np.where(np.abs(x) <= 1.0, inside, outside)


def test_chebyfilt_closed_form():
    H, (w, U), rng = _H()
    X = rng.standard_normal((60, 5)) + 1j * rng.standard_normal((60, 5))
    for m, lb, ub in ((1, w[10], w[-1]), (6, w[20], w[-1] + 0.1), (9, w[30], w[-1] + 2.0)):
        Y = chebyfilt(lambda Z: H @ Z, X, m, lb, ub)
        e, c = (ub - lb) / 2, (ub + lb) / 2
        ref = U @ np.diag(_T(m, (w - c) / e) / _T(m, -2.0)) @ U.conj().T @ X
        assert np.abs(Y - ref).max() < 1e-12 * np.abs(ref).max(), m


def test_lanczos_bounds():
    H, (w, U), rng = _H(seed=1)
    for nl in (10, 20, 40):
        lb, ub = get_ub_lb_lanczos(lambda Z: H @ Z, rng.standard_normal(60) + 1j * rng.standard_normal(60), nl)
        assert w[0] < lb < w[-1]
        assert ub >= w[-1]


def test_chefsi_converges_to_lowest_eigenvalues():
    H, (w, U), rng = _H(seed=3)          # gap between the 5th and 6th eigenvalue: 1.16
    X = ortho_sqrt(rng.standard_normal((60, 5)) + 1j * rng.standard_normal((60, 5)))
    for _ in range(40):
        ev, X = diag_CheFSI(lambda Z: H @ Z, X, 8, rng.standard_normal(60) + 1j * rng.standard_normal(60))
    assert np.abs(ev - w[:5]).max() < 1e-9, (ev, w[:5])
    assert np.abs(X.conj().T @ X - np.eye(5)).max() < 1e-12
