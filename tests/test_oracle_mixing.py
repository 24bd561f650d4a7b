"""Pins of the oracle's density mixers (oracle/mixing.py) by what the mathematics fixes,
on fixed-point problems x = g(x) with g affine, g(x) = A x + b, A symmetric with spectrum
in (-0.6, 0.95) (simple mixing converges slowly):

- the first iteration of every method is simple mixing, x + beta (g(x) - x), except the
  adaptive mixer, whose per-point beta starts at beta0 and grows by beta0 at once (the
  stored residual starts at zero);
- Pulay (mix_pulay) with MIXDIM = dim extrapolates exactly once its history is full: the
  M difference columns span the space, so x - X F^-1 f is the fixed point;
- restarted Pulay (mix_rpulay) is exact at iteration MIXDIM + 1 for the same reason;
- Anderson (mix_anderson) with a history longer than dim is exact by iteration dim + 2
  (it is GMRES on the affine residual);
- Broyden, periodic Pulay and the adaptive linear mixer converge to the fixed point, far
  below what simple mixing reaches in as many iterations.
The exactness checks use a small, well-conditioned problem (dim 4, spread spectrum) so that
the Krylov-like history columns are far from parallel.
A wrong sign, index or dropped term in any of them breaks one of these."""
import numpy as np
import pytest

from oracle import mixing as M

DIM = 6
SMALL = 4


def _problem(seed=0, dim=DIM, spread=False):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((dim, dim)))
    if spread:
        lam = np.linspace(-0.8, 0.9, dim)
    else:
        lam = np.concatenate([[0.95, -0.6], rng.uniform(-0.5, 0.9, dim - 2)])
    A = Q @ np.diag(lam) @ Q.T
    b = rng.uniform(1.0, 2.0, dim)
    xs = np.linalg.solve(np.eye(dim) - A, b)
    return (lambda x: A @ x + b), xs, rng.uniform(1.0, 2.0, dim)


def _run(method, g, x0, iters, beta=0.3, mixdim=DIM):
    n = x0.size
    x = x0.copy()
    st = dict(X=np.zeros((n, mixdim)), F=np.zeros((n, mixdim)), xo=np.zeros(n), fo=np.zeros(n),
              df=np.zeros((n, mixdim)), dv=np.zeros((n, mixdim)), bv=np.full(n, beta), f=np.zeros(n))
    hist = []
    for it in range(1, iters + 1):
        gx = g(x)
        if hist and np.abs(gx - x).max() < 1e-12 * np.abs(x).max():
            break   # converged: stop as the SCF loop does (the histories degenerate beyond)
        if method == "simple":
            x = beta * gx + (1 - beta) * x
        elif method == "pulay":
            M.mix_pulay(x, gx, beta, st["X"], st["F"], it, mixdim, st["xo"], st["fo"])
        elif method == "rpulay":
            M.mix_rpulay(x, gx, beta, st["X"], st["F"], it, mixdim, st["xo"], st["fo"])
        elif method == "ppulay":
            M.mix_ppulay(x, gx, beta, st["X"], st["F"], it, mixdim, 3, st["xo"], st["fo"])
        elif method == "anderson":
            M.mix_anderson(x, gx, beta, st["df"], st["dv"], it, mixdim)
        elif method == "broyden":
            M.mix_broyden(x, gx, beta, it, mixdim, st["df"], st["dv"])
        elif method == "linear_adaptive":
            M.mix_adaptive(x, gx, beta, st["bv"], st["f"])
        hist.append(x.copy())
    return hist


@pytest.mark.parametrize("method", ["pulay", "rpulay", "ppulay", "anderson", "broyden", "linear_adaptive"])
def test_first_iteration_is_simple_mixing(method):
    g, xs, x0 = _problem(1)
    b = min(0.6, 0.8) if method == "linear_adaptive" else 0.3
    assert np.allclose(_run(method, g, x0, 1)[0], x0 + b * (g(x0) - x0), rtol=1e-15, atol=1e-15)


def test_pulay_exact_when_history_spans():
    g, xs, x0 = _problem(2, SMALL, spread=True)
    h = _run("pulay", g, x0, SMALL + 1, mixdim=SMALL)
    assert np.abs(h[SMALL] - xs).max() < 1e-9 * np.abs(xs).max()      # output of iteration dim + 1
    assert np.abs(h[SMALL - 1] - xs).max() > 1e-3                        # not before


def test_rpulay_exact_at_full_active_history():
    g, xs, x0 = _problem(3, SMALL, spread=True)
    h = _run("rpulay", g, x0, SMALL + 1, mixdim=SMALL)
    assert np.abs(h[SMALL] - xs).max() < 1e-9 * np.abs(xs).max()
    assert np.abs(h[SMALL - 1] - xs).max() > 1e-3


def test_anderson_exact_like_gmres():
    g, xs, x0 = _problem(4, SMALL, spread=True)
    h = _run("anderson", g, x0, SMALL + 2, mixdim=SMALL + 2)
    assert np.abs(h[-1] - xs).max() < 1e-9 * np.abs(xs).max()
    assert len(h) <= SMALL + 2
    assert np.abs(h[SMALL - 2] - xs).max() > 1e-3


@pytest.mark.parametrize("method,iters,tol,gain", [("broyden", 40, 1e-10, 1e6), ("ppulay", 40, 1e-10, 1e6),
                                                    ("linear_adaptive", 300, 1e-3, 1e3)])
def test_converges_beyond_simple_mixing(method, iters, tol, gain):
    g, xs, x0 = _problem(5)
    h = _run(method, g, x0, iters)
    err = np.abs(h[-1] - xs).max()
    err_simple = np.abs(_run("simple", g, x0, len(h))[-1] - xs).max()
    assert err < tol and err_simple > gain * err, (err, err_simple)
