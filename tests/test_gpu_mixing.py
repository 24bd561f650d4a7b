"""The density mixers on device vectors (pwdft_hip/mixing.py over pw_vec_dots_dev,
pw_vec_lincomb_dev, pw_mix_adaptive_dev) against the oracle's restatements of mix_*.jl on
the same inputs, and the SCF with each mix_method against the reference log's Si fcc
total energy."""
import types

import numpy as np
import pytest

from _systems import A_SI, product_ham, si_fcc_structure

pytestmark = pytest.mark.gpu

EPS = np.finfo(float).eps


def _g(x):
    """A smooth contraction with a nonlinear part, positive on positive input."""
    return 1.0 + 0.45 * (np.roll(x, 1) + np.roll(x, -7)) * (1.0 + 0.05 * np.sin(x))


def _load(mixer, st, x, xd):
    """Copy the oracle's state (density and history, logical column order) to the device
    mixer, so that each step is compared from identical inputs."""
    xd.upload(x)
    cols = lambda a: [a[:, j] for j in range(a.shape[1])]
    if mixer.method in ("broyden", "anderson"):
        for v, c in zip(mixer.df + mixer.dv, cols(st["df"]) + cols(st["dv"])):
            v.upload(c)
    elif mixer.method == "linear_adaptive":
        mixer.betav.upload(st["bv"])
        mixer.f.upload(st["f"])
    else:
        for v, c in zip(mixer.X + mixer.F, cols(st["X"]) + cols(st["F"])):
            v.upload(c)
        mixer.x_old.upload(st["xo"])
        mixer.f_old.upload(st["fo"])


@pytest.mark.parametrize("method,tol", [("linear_adaptive", 1e-15), ("broyden", 1e-12), ("pulay", 1e-12),
                                        ("rpulay", 1e-12), ("ppulay", 1e-12), ("anderson", 1e-11)])
def test_mixer_vs_oracle(method, tol):
    """Twelve mixing steps; before each, the device mixer is given the oracle's density and
    history, and its output is compared with the oracle's step (rounding differences are
    not compounded through the history matrices)."""
    import pwdft_hip as ph
    from pwdft_hip.mixing import DensityMixer
    from oracle import mixing as M
    print(f"case mixer {method}", flush=True)
    n, mixdim, beta = 19683, 5, 0.4
    ham = types.SimpleNamespace(ctx=ph.Context(0), pw=types.SimpleNamespace(Npoints=n))
    mixer = DensityMixer(ham, method, beta, mixdim)
    rng = np.random.default_rng(7)
    x = rng.uniform(1.0, 2.0, n)
    xd = ph.DeviceVector(ham.ctx, n)
    gd = ph.DeviceVector(ham.ctx, n)
    st = dict(X=np.zeros((n, mixdim)), F=np.zeros((n, mixdim)), xo=np.zeros(n), fo=np.zeros(n),
              df=np.zeros((n, mixdim)), dv=np.zeros((n, mixdim)), bv=np.full(n, beta), f=np.zeros(n))
    for it in range(1, 13):
        gx = _g(x)
        _load(mixer, st, x, xd)
        gd.upload(gx)
        mixer.mix(xd, gd, it)
        if method == "linear_adaptive":
            M.mix_adaptive(x, gx, beta, st["bv"], st["f"])
        elif method == "broyden":
            M.mix_broyden(x, gx, beta, it, mixdim, st["df"], st["dv"])
        elif method == "pulay":
            M.mix_pulay(x, gx, beta, st["X"], st["F"], it, mixdim, st["xo"], st["fo"])
        elif method == "rpulay":
            M.mix_rpulay(x, gx, beta, st["X"], st["F"], it, mixdim, st["xo"], st["fo"])
        elif method == "ppulay":
            M.mix_ppulay(x, gx, beta, st["X"], st["F"], it, mixdim, 3, st["xo"], st["fo"])
        else:
            M.mix_anderson(x, gx, beta, st["df"], st["dv"], it, mixdim)
        np.maximum(x, EPS, out=x)       # the clamp of KS_solve_SCF.jl:277-281
        err = np.abs(xd.download() - x).max() / np.abs(x).max()
        assert err < tol, (it, err)


@pytest.mark.parametrize("method,betamix", [("linear_adaptive", 0.1), ("anderson", 0.5), ("pulay", 0.5),
                                            ("ppulay", 0.5), ("rpulay", 0.5), ("broyden", 0.5)])
def test_scf_mix_methods_si_fcc(method, betamix):
    """KS_solve_SCF! on Si fcc (3x3x3, 15 Ha) with each mixer and the log's betamix reaches
    the self-consistent energy of the reference log (-7.9110091083 with simple mixing at
    etot_conv_thr = 1e-6; the log's runs with these mixers agree with it to 5e-7)."""
    import pwdft_hip as ph
    print(f"case SCF Si fcc mix {method}", flush=True)
    sym, frac = si_fcc_structure()
    H = product_ham(sym, frac, ph.gen_lattice_fcc(A_SI), "Si-q4.lda.gth", 15.0, meshk=(3, 3, 3), extra_states=4,
                    use_symmetry=True)
    out = ph.KS_solve_SCF_(H, betamix=betamix, mix_method=method, etot_conv_thr=1e-8, ethr_evals_last=1e-9,
                           NiterMax=100)
    E = out["energies"]["Total"]
    assert out["iterations"] < 100
    assert abs(E - (-7.9110091083)) < 2e-6, (method, E, out["iterations"])


@pytest.mark.parametrize("method,betamix,ref", [("simple", 0.5, -7.9110091083), ("linear_adaptive", 0.1, -7.9110091083),
                                                ("broyden", 0.5, -7.9110091083)])
def test_scf_potmix_si_fcc(method, betamix, ref):
    """KS_solve_SCF_potmix! on Si fcc with the reference's validation settings
    (test/validations_old/Si_fcc.jl:114-140) reaches the self-consistent energy."""
    import pwdft_hip as ph
    print(f"case SCF potmix {method}", flush=True)
    sym, frac = si_fcc_structure()
    H = product_ham(sym, frac, ph.gen_lattice_fcc(A_SI), "Si-q4.lda.gth", 15.0, meshk=(3, 3, 3), use_symmetry=True)
    out = ph.KS_solve_SCF_potmix_(H, betamix=betamix, mix_method=method, etot_conv_thr=1e-8,
                                  ethr_evals_last=1e-9, NiterMax=150)
    assert out["iterations"] < 150
    assert abs(out["energies"]["Total"] - ref) < 2e-6, (method, out["energies"]["Total"], out["iterations"])
