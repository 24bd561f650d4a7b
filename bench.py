"""Benchmark: H psi band-applications/s on the Si64 H-psi microbenchmark.

Workload (one GPU): Si64 simple-cubic supercell (2x2x2 conventional diamond cells,
a = 20.5262 bohr), GTH LDA Si, ecutwfc 15 Ha, 72^3 dense FFT grid (Ns_ override),
Gamma point, Ngw = 23871 plane waves, 160 bands, Nbeta = 320 projectors.  psi is complex
i.i.d. normal from a seed, Loewdin-orthonormalised; V_loc(r) = V_Ps_loc + V_H + V_xc of
the Gaussian atomic-superposition density.  Inputs are resident in HBM before timing.

A step is one op_H over all 160 bands (a band-application is one column through op_H).
`value` is the mixed-precision path (FP32 FFT + FP32-MFMA projectors); the FP64 path is
timed beside it, and dE/atom between the two (sum_b <psi_b|H|psi_b> / Natoms) is reported.
Beside the op_H line: diag_LOBPCG! band-applications/s over a fixed number of iterations
(op_H plus the Rayleigh-Ritz products) and KS_solve_SCF! seconds per iteration, each in
FP64 and in mixed precision.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "pwdft.jl_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

A_SI64 = 20.5262
A_SI = 10.2631
NBANDS = 160
PEAK_HBM_GBS = 8000.0


def build_system(mixed=False):
    import pwdft_hip as ph
    frac = ph.diamond_supercell(1.0, 2)
    atoms = ph.Atoms(["Si"] * len(frac), frac, ph.gen_lattice_sc(A_SI64))
    H = ph.Hamiltonian(atoms, ["Si-q4.lda.gth"], 15.0, Ns_=(72, 72, 72), Nstates=NBANDS, mixed=mixed,
                       band_batch=NBANDS)
    H.update(H.guess_rhoe())
    return H


def make_psi(H, seed=20260101):
    import pwdft_hip as ph
    rng = np.random.default_rng(seed)
    n = H.pw.Ngw[0]
    a = rng.standard_normal((n, NBANDS)) + 1j * rng.standard_normal((n, NBANDS))
    X = ph.DeviceMatrix.from_host(H.ctx, a)
    H.ik = 0
    ph.ortho_sqrt_(H, X)
    return X


def cpu_threads():
    """Host cores this process may use, capped at 16 (one GPU's share of a host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(seconds=12.0, threads=1):
    """Time the oracle (CPU restatement, reference loop structure) on a bounded sample of
    the same workload in a child process: `threads` BLAS threads and FFT workers (numpy's
    FFT for one thread, scipy.fft with workers otherwise)."""
    code = r"""
import os, sys, time, json
sys.path.insert(0, %r); sys.path.insert(0, %r)
import numpy as np
import oracle.setup
oracle.setup.FFT_WORKERS = %d
from oracle.setup import Atoms, PsPot_GTH, gen_lattice_sc
from oracle.energy import Hamiltonian
from oracle.density import guess_rhoe
from oracle import ops
import pwdft_hip.basis as B
frac = B.diamond_supercell(1.0, 2)
atoms = Atoms(["Si"]*len(frac), frac, gen_lattice_sc(%r))
H = Hamiltonian(atoms, [PsPot_GTH(os.path.join(%r, "Si-q4.lda.gth"))], 15.0, Ns=(72,72,72), Nstates=%d)
H.update(guess_rhoe(H))
rng = np.random.default_rng(1)
n = H.pw.Ngw[0]
done = 0; t0 = time.perf_counter()
while time.perf_counter() - t0 < %r:
    psi = rng.standard_normal((n, 4)) + 1j*rng.standard_normal((n, 4))
    ops.op_H(H, psi); done += 4
dt = time.perf_counter() - t0
print(json.dumps({"bands": done, "seconds": dt}))
""" % (ROOT, os.path.join(ROOT, "pwdft.jl_amd"), threads, A_SI64,
       os.path.join(ROOT, "pwdft.jl_amd", "pwdft_hip", "pseudopotentials"), NBANDS, seconds)
    t = str(threads)
    env = dict(os.environ, OPENBLAS_NUM_THREADS=t, OMP_NUM_THREADS=t, MKL_NUM_THREADS=t)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, timeout=600)
    if r.returncode != 0:
        return {"value": None, "unit": "band-applications/s", "cores": threads, "kind": "port",
                "sample": "failed: " + r.stderr[-300:]}
    d = json.loads(r.stdout.strip().splitlines()[-1])
    fft = "numpy FFT" if threads == 1 else f"scipy.fft workers={threads}"
    return {"value": d["bands"] / d["seconds"], "unit": "band-applications/s", "cores": threads, "kind": "port",
            "sample": f"oracle op_H ({fft}, {threads} BLAS thread(s), GTH rank-1 nonlocal loop) on "
                      f"{d['bands']} bands of the same Si64 system in {d['seconds']:.1f} s"}


def time_op_H(H, X, Y, steps, warmup, mixed, profile=False):
    """Time `steps` op_H applications (all bands) between two events on the library's
    stream; with profile=True the library also brackets every kernel launch with events.
    Returns (GPU seconds, wall seconds, per-kernel profile or None)."""
    from pwdft_hip._lib import PW_MIXED, PW_TERM_ALL
    c = H.ctx
    flags = PW_TERM_ALL | (PW_MIXED if mixed else 0)
    for _ in range(warmup):
        c.call("pw_op_H_dev", 0, 0, NBANDS, X.p, X.rows, Y.p, Y.rows, flags)
    c.call("pw_synchronize")
    if profile:
        c.profile(True)
    c.call("pw_event_record", 0)
    t0 = time.perf_counter()
    for _ in range(steps):
        c.call("pw_op_H_dev", 0, 0, NBANDS, X.p, X.rows, Y.p, Y.rows, flags)
    c.call("pw_event_record", 1)
    c.call("pw_synchronize")
    wall = time.perf_counter() - t0
    import ctypes
    ms = ctypes.c_float()
    c.call("pw_event_elapsed_ms", 0, 1, ctypes.byref(ms))
    prof = None
    if profile:
        prof = c.profile_read()
        c.profile(False)
    return ms.value / 1e3, wall, prof


def time_rhoe(H, X, steps, warmup, mixed):
    """calc_rhoe! over the 160 bands (pw_rhoe_zero + pw_rhoe_accum_dev, weights 2/Nbands):
    seconds per step and the per-kernel profile of the density stage."""
    from pwdft_hip._lib import PW_MIXED, ptr
    c = H.ctx
    w = np.full(NBANDS, 2.0 / NBANDS)
    fl = PW_MIXED if mixed else 0
    for _ in range(warmup):
        c.call("pw_rhoe_zero")
        c.call("pw_rhoe_accum_dev", 0, NBANDS, X.p, X.rows, ptr(w), fl)
    c.call("pw_synchronize")
    t0 = time.perf_counter()
    for _ in range(steps):
        c.call("pw_rhoe_zero")
        c.call("pw_rhoe_accum_dev", 0, NBANDS, X.p, X.rows, ptr(w), fl)
    c.call("pw_synchronize")
    sec = (time.perf_counter() - t0) / steps
    c.profile(True)
    for _ in range(steps):
        c.call("pw_rhoe_zero")
        c.call("pw_rhoe_accum_dev", 0, NBANDS, X.p, X.rows, ptr(w), fl)
    prof = c.profile_read()
    c.profile(False)
    out = {"ms_per_step": sec * 1e3, "band_apps_per_s": NBANDS / sec}
    for k in ("col_ifft", "plane_rhoe"):
        ms, n, b = prof[k]
        if n:
            us = ms / n * 1e3
            out[k] = {"us": us, "achieved": b / n / (us * 1e-6) / 1e9, "unit": "GB/s", "peak": PEAK_HBM_GBS,
                      "frac": b / n / (us * 1e-6) / 1e9 / PEAK_HBM_GBS}
    if "plane_rhoe" in out:
        out["plane_rhoe"]["traffic"] = _traffic("plane_rhoe", mixed)
    return out


def time_lobpcg(H, X0, iters, mixed):
    """diag_LOBPCG! on the Si64 bands for a fixed number of iterations (tol = 0, so no band
    converges or locks) from the same orthonormal start; returns band-applications/s
    (op_H columns: the initial HX plus Nst per iteration) and seconds per iteration."""
    import pwdft_hip as ph
    from pwdft_hip.solvers import diag_LOBPCG_
    X = ph.DeviceMatrix(H.ctx, X0.rows, X0.cols)
    H.mixed = mixed
    H.ik = 0
    # untimed warm-up: two iterations, so the Rayleigh-Ritz orders 2 Nst and 3 Nst have each
    # run once (workspace growth and the dense solver library's first-call set-up)
    H.ctx.call("pw_copy_cols_dev", X0.rows, X0.cols, X0.p, X0.rows, X.p, X.rows)
    diag_LOBPCG_(H, X, tol=0.0, NiterMax=2)
    H.ctx.call("pw_copy_cols_dev", X0.rows, X0.cols, X0.p, X0.rows, X.p, X.rows)
    H.ctx.call("pw_synchronize")
    t0 = time.perf_counter()
    _, st = diag_LOBPCG_(H, X, tol=0.0, NiterMax=iters)
    H.ctx.call("pw_synchronize")
    dt = time.perf_counter() - t0
    H.mixed = False
    X.free()
    return {"band_apps_per_s": st["ncols"] / dt, "s_per_iter": dt / st["iter"], "iterations": st["iter"],
            "band_apps": st["ncols"]}


def time_scf(H, iters, mixed):
    """KS_solve_SCF! iterations on Si64 (LOBPCG + calc_rhoe + potential update + energies),
    seconds per SCF iteration after the initial guess."""
    import pwdft_hip as ph
    H.mixed = mixed
    out = ph.KS_solve_SCF_(H, betamix=0.5, NiterMax=iters, etot_conv_thr=0.0)
    H.mixed = False
    return {"s_per_iter": out["seconds"] / out["iterations"], "iterations": out["iterations"],
            "band_apps_per_s": out["ncols"] / out["seconds"], "Etot": out["energies"]["Total"]}


def time_batched_k(steps, warmup):
    """Al fcc (a = 7.6 bohr, GTH LDA, 15 Ha), 12x12x12 shifted mesh reduced by symmetry to
    182 k-points, 8 bands each: op_H over all k-points in one launch per stage
    (pw_op_H_batch_dev) against the per-k loop, FP64 and mixed, plus one batched LOBPCG."""
    import pwdft_hip as ph
    atoms = ph.Atoms(["Al"], [[0, 0, 0]], ph.gen_lattice_fcc(7.6))
    H = ph.Hamiltonian(atoms, ["Al-q3.lda.gth"], 15.0, meshk=(12, 12, 12), shiftk=(1, 1, 1), Nstates=8,
                       use_smearing=True, smearing_kT=0.01)
    H.update(H.guess_rhoe())
    ps = ph.rand_BlochWavefunc(H, seed=20260101)
    out = [ph.DeviceMatrix(H.ctx, p.rows, p.cols) for p in ps]
    nk = H.pw.Nkpt
    res = {"workload": "Al fcc a=7.6, GTH LDA, 15 Ha, 12x12x12 shifted -> %d k-points, 8 bands" % nk,
           "Ngw_min": int(min(H.pw.Ngw)), "Ngw_max": int(max(H.pw.Ngw))}
    for mixed in (False, True):
        tag = "mixed" if mixed else "fp64"
        H.mixed = mixed
        for _ in range(warmup):
            ph.op_H_batch(H, ps, out=out)
        H.ctx.call("pw_synchronize")
        t0 = time.perf_counter()
        for _ in range(steps):
            ph.op_H_batch(H, ps, out=out)
        H.ctx.call("pw_synchronize")
        tb = (time.perf_counter() - t0) / steps
        t0 = time.perf_counter()
        for _ in range(max(1, steps // 4)):
            for ik in range(nk):
                H.ik = ik
                ph.op_H_(H, ps[ik], out[ik])
        H.ctx.call("pw_synchronize")
        tl = (time.perf_counter() - t0) / max(1, steps // 4)
        res[tag] = {"batched_band_apps_per_s": nk * 8 / tb, "batched_ms": tb * 1e3,
                    "per_k_loop_band_apps_per_s": nk * 8 / tl, "per_k_loop_ms": tl * 1e3}
        print(f"# Al 182k {tag}: batched op_H {tb * 1e3:.3f} ms, per-k loop {tl * 1e3:.3f} ms",
              file=sys.stderr, flush=True)
    H.mixed = False
    # batched LOBPCG: every phase of an iteration replayed from its captured hipGraph (the
    # warm-up calls run each phase eagerly once, then capture it), and the same iterations
    # enqueued eagerly for comparison
    from pwdft_hip._lib import PW_TUNE_LOBPCG_GRAPH
    res["lobpcg_batch"] = {}
    # every call starts from the same random block: the solver overwrites psiks with its Ritz
    # vectors, and a start converged by earlier calls would send k-points through the lock
    # branch (|dE| <= 10 eps), which is other work
    start = [p.download() for p in ps]
    for mixed in (False, True):
        tag = "mixed" if mixed else "fp64"
        for graph in (1, 0):
            H.ctx.call("pw_tune_set", PW_TUNE_LOBPCG_GRAPH, graph)
            for _ in range(2):
                for p, a in zip(ps, start):
                    p.upload(a)
                ph.diag_LOBPCG_batch(H, ps, tol=0.0, NiterMax=4, mixed=mixed)
            for p, a in zip(ps, start):
                p.upload(a)
            H.ctx.call("pw_synchronize")
            t0 = time.perf_counter()
            _, st = ph.diag_LOBPCG_batch(H, ps, tol=0.0, NiterMax=4, mixed=mixed)
            H.ctx.call("pw_synchronize")
            dt = time.perf_counter() - t0
            key = tag + ("_graph" if graph else "_eager")
            res["lobpcg_batch"][key] = {"s_per_iter": dt / 4, "band_apps_per_s": sum(x["ncols"] for x in st) / dt,
                                        "locks": sum(x["locks"] for x in st)}
            print(f"# Al 182k batched LOBPCG {key}: {dt / 4 * 1e3:.2f} ms/iteration", file=sys.stderr, flush=True)
    H.ctx.call("pw_tune_set", PW_TUNE_LOBPCG_GRAPH, 1)
    return res


SI8_SHIFT = np.array([[0.041, -0.027, 0.033, -0.012, 0.018, -0.045, 0.006, 0.029],
                      [-0.036, 0.014, 0.022, -0.049, 0.031, 0.008, -0.019, 0.043],
                      [0.025, 0.047, -0.038, 0.011, -0.023, 0.035, -0.044, -0.007]])


def forces_leg():
    """Si8 cubic (a = 10.2631 bohr, GTH LDA, 20 Ha, 2x2x2 k-points, 20 bands), every atom
    moved by a fixed offset <= 0.05 bohr: KS_solve_SCF! to 1e-9 Ha in FP64 and in mixed
    precision (FP32 until |dE| < 1e-6, then FP64), the total forces of both and their
    largest difference, with the SCF seconds per iteration."""
    import pwdft_hip as ph
    L = ph.gen_lattice_sc(A_SI)
    frac = np.array(ph.diamond_supercell(1.0, 1), dtype=float)
    frac = np.linalg.solve(L, L @ frac.T + SI8_SHIFT).T
    res = {"workload": "Si8 sc a=10.2631, GTH LDA, 20 Ha, 2x2x2 k-points (no symmetry), 20 bands, "
                       "atoms displaced <= 0.05 bohr"}
    F = {}
    E = {}
    for mixed in (False, True):
        tag = "mixed" if mixed else "fp64"
        H = ph.Hamiltonian(ph.Atoms(["Si"] * 8, frac, L), ["Si-q4.lda.gth"], 20.0, meshk=(2, 2, 2), Nstates=20,
                           use_symmetry=False, mixed=mixed)
        t0 = time.perf_counter()
        out = ph.KS_solve_SCF_(H, betamix=0.3, etot_conv_thr=1e-9, ethr_evals_last=1e-10, NiterMax=80,
                               mixed_until=1e-6 if mixed else None)
        dt = time.perf_counter() - t0
        rho = ph.calc_rhoe(H, out["psiks"])
        F[tag] = ph.calc_forces(H, out["psiks"], rho)
        E[tag] = out["energies"]["Total"]
        res[tag] = {"s_per_iter": dt / out["iterations"], "iterations": out["iterations"], "Etot": E[tag],
                    "max_abs_force": float(np.abs(F[tag]).max())}
        H.ctx.close()
    res["dE_per_atom_mixed_vs_fp64"] = abs(E["mixed"] - E["fp64"]) / 8
    res["max_dF_mixed_vs_fp64"] = float(np.abs(F["mixed"] - F["fp64"]).max())
    print(f"# Si8 forces: dE/atom {res['dE_per_atom_mixed_vs_fp64']:.2e}, max|dF| {res['max_dF_mixed_vs_fp64']:.2e}, "
          f"SCF {res['fp64']['s_per_iter'] * 1e3:.1f} / {res['mixed']['s_per_iter'] * 1e3:.1f} ms/iteration",
          file=sys.stderr, flush=True)
    return res


def si216_leg():
    """Si216 (simple cubic a = 30.789 bohr, atoms moved by up to 0.02 bohr, GTH PBE, 15 Ha,
    108^3 grid, Gamma, 480 bands): op_H over all bands, diag_LOBPCG! iterations and the first
    two KS_solve_SCF! iterations, FP64 and mixed."""
    import pwdft_hip as ph
    from pwdft_hip._lib import PW_MIXED, PW_TERM_ALL
    from pwdft_hip.solvers import diag_LOBPCG_
    frac = np.array(ph.diamond_supercell(1.0, 3), dtype=float)
    L = ph.gen_lattice_sc(30.789)
    cart = L @ frac.T + np.random.default_rng(216).uniform(-0.02, 0.02, size=(3, len(frac)))
    H = ph.Hamiltonian(ph.Atoms(["Si"] * len(frac), np.linalg.solve(L, cart).T, L), ["Si-q4.pbe.gth"], 15.0,
                       Ns_=(108, 108, 108), Nstates=480, xcfunc="PBE", use_symmetry=False, band_batch=96)
    H.update(H.guess_rhoe())
    X = ph.rand_BlochWavefunc(H, seed=1)[0]
    Y = ph.DeviceMatrix(H.ctx, X.rows, X.cols)
    res = {"workload": "Si216 sc a=30.789, GTH PBE, 15 Ha, 108^3 grid, Gamma, 480 bands",
           "Ngw": int(H.pw.Ngw[0]), "Nbeta": int(H.pspotNL.NbetaNL)}
    for mixed in (False, True):
        tag = "mixed" if mixed else "fp64"
        fl = PW_TERM_ALL | (PW_MIXED if mixed else 0)
        H.ctx.call("pw_op_H_dev", 0, 0, 480, X.p, X.rows, Y.p, Y.rows, fl)
        H.ctx.call("pw_synchronize")
        t = time.perf_counter()
        for _ in range(3):
            H.ctx.call("pw_op_H_dev", 0, 0, 480, X.p, X.rows, Y.p, Y.rows, fl)
        H.ctx.call("pw_synchronize")
        dt = (time.perf_counter() - t) / 3
        H.mixed = mixed
        Xc = ph.DeviceMatrix(H.ctx, X.rows, X.cols)
        H.ctx.call("pw_copy_cols_dev", X.rows, X.cols, X.p, X.rows, Xc.p, Xc.rows)
        diag_LOBPCG_(H, Xc, tol=0.0, NiterMax=2)   # reaches the 3 Nst Rayleigh-Ritz order once
        H.ctx.call("pw_copy_cols_dev", X.rows, X.cols, X.p, X.rows, Xc.p, Xc.rows)
        H.ctx.call("pw_synchronize")
        t = time.perf_counter()
        _, st = diag_LOBPCG_(H, Xc, tol=0.0, NiterMax=3)
        dl = (time.perf_counter() - t) / 3
        Xc.free()
        t = time.perf_counter()
        out = ph.KS_solve_SCF_(H, betamix=0.3, NiterMax=2, etot_conv_thr=1e-12)
        ds = (time.perf_counter() - t) / out["iterations"]
        res[tag] = {"op_H_ms": dt * 1e3, "band_apps_per_s": 480 / dt, "lobpcg_s_per_iter": dl,
                    "scf_s_per_iter_first2": ds}
        print(f"# Si216 {tag}: op_H {dt * 1e3:.1f} ms, LOBPCG {dl * 1e3:.0f} ms/iteration, "
              f"SCF {ds:.2f} s/iteration", file=sys.stderr, flush=True)
    H.mixed = False
    return res


def stages(prof, H, mixed, steps):
    """Achieved rate of every op_H stage: algorithmic GB/s (HBM peak 8 TB/s) for the FFT
    kernels, TFLOP/s (MFMA peak FP32 157.3 / FP64 78.6) for the projector products."""
    out = {}
    nb, ngw, nbeta = NBANDS, H.pw.Ngw[0], H.pspotNL.NbetaNL
    for k, (ms, n, b) in prof.items():
        if not n:
            continue
        us = ms / n * 1e3
        if k.startswith("gemm"):
            fl = 8.0 * ngw * nbeta * nb
            a = fl / (us * 1e-6) / 1e12
            pk = 157.3 if mixed else 78.6
            out[k] = {"us": us, "achieved": a, "unit": "TFLOP/s", "peak": pk, "frac": a / pk}
        else:
            a = b / n / (us * 1e-6) / 1e9
            out[k] = {"us": us, "achieved": a, "unit": "GB/s", "peak": 8000.0, "frac": a / 8000.0}
    return out


def roofline(prof, H, mixed):
    """Roofline object for the kernel with the largest total time in one op_H step."""
    name = max(prof, key=lambda k: prof[k][0])
    ms, n, nbytes = prof[name]
    avg = ms / 1e3 / max(n, 1)
    per = nbytes / max(n, 1)
    ngw, nb = H.pw.Ngw[0], H.pspotNL.NbetaNL
    if name.startswith("gemm"):
        flops = 8.0 * ngw * nb * NBANDS
        peak = 157.3 if mixed else 78.6
        achieved = flops / avg / 1e12
        return {"kernel": name, "bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": None, "avg_launch_us": avg * 1e6,
                "algorithmic": flops}
    achieved = per / avg / 1e9
    tr = _traffic(name, mixed)
    return {"kernel": name, "bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS, "traffic": tr, "avg_launch_us": avg * 1e6,
            "algorithmic": per}


def _traffic(name, mixed):
    """HBM bytes per launch of `name` from the committed rocprofv3 counter summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration), or None."""
    f = os.path.join(ROOT, "profiles", "traffic_si64.json")
    try:
        with open(f) as fh:
            d = json.load(fh)
        return d["kernels"][("mixed/" if mixed else "fp64/") + name]["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run one 256 MiB device-to-device copy (counter calibration)")
    ap.add_argument("--solver-iters", type=int, default=4,
                    help="LOBPCG iterations timed per precision (0 = skip)")
    ap.add_argument("--scf-iters", type=int, default=3, help="SCF iterations timed per precision (0 = skip)")
    ap.add_argument("--no-batched-k", dest="batched_k", action="store_false",
                    help="skip the 182-k-point batched op_H leg")
    ap.add_argument("--si216", action="store_true",
                    help="also run the Si216 PBE leg (108^3, 480 bands; adds about a minute)")
    ap.add_argument("--no-forces", dest="forces", action="store_false",
                    help="skip the Si8 mixed-vs-FP64 force leg")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="run the library's default launch parameters instead of the autotuner's "
                         "(pwdft_hip/autotune.py: FFT radix mode, plane kernel, column chunk, band "
                         "split, GEMM tiles, second stream; tuned once per precision before timing)")
    ap.add_argument("--no-overlap", dest="overlap", action="store_false",
                    help="run op_H's nonlocal products on the main stream (kernels in sequence), "
                         "e.g. under a kernel-trace profiler")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus > 1 is not supported"}))
        sys.exit(2)
    import pwdft_hip as ph
    print("# building Si64 system", file=sys.stderr, flush=True)
    H = build_system()
    X = make_psi(H)
    Y = ph.DeviceMatrix(H.ctx, X.rows, NBANDS)
    print(f"# Ngw={H.pw.Ngw[0]} Nbeta={H.pspotNL.NbetaNL} Ns={H.pw.Ns}", file=sys.stderr, flush=True)
    if args.calib_copy:
        a = ph.DeviceMatrix(H.ctx, 16 * 1024 * 1024, 1)
        b = ph.DeviceMatrix(H.ctx, 16 * 1024 * 1024, 1)
        H.ctx.call("pw_memcpy_d2d", b.p, a.p, a.nbytes)
        H.ctx.call("pw_synchronize")
        a.free()
        b.free()
    # launch parameters per precision: the autotuner's (timed here, outside every timed
    # region, or read from its cache), else the library's defaults with all bands per launch
    from pwdft_hip import autotune as at
    tuned = {}
    for mixed in (False, True):
        tag = "mixed" if mixed else "fp64"
        if args.autotune:
            t0 = time.perf_counter()
            settings, source = at.autotune(H, nbands=NBANDS, mixed=mixed)
            tuned[tag] = dict(settings, source=source, seconds=round(time.perf_counter() - t0, 1))
            print(f"# autotune {tag} ({source}): {json.dumps(settings)}", file=sys.stderr, flush=True)
        else:
            tuned[tag] = dict(at.DEFAULTS, band_batch=NBANDS)

    def use(mixed):
        at.apply_settings(H.ctx, tuned["mixed" if mixed else "fp64"])
        if not args.overlap:
            H.ctx.call("pw_tune_set", 6, 0)

    # accuracy of the mixed path: band energies <psi|H|psi>
    res = {}
    from pwdft_hip.solvers import _Dense
    for mixed in (False, True):
        # headline: the tuned launch parameters, no per-kernel events; then the per-kernel
        # durations for the roofline with the kernels in sequence on one stream (overlapped
        # kernels share the CUs)
        use(mixed)
        sec, wall, _ = time_op_H(H, X, Y, args.steps, args.warmup, mixed, profile=False)
        H.ctx.call("pw_tune_set", 6, 0)
        _, _, prof = time_op_H(H, X, Y, args.steps, args.warmup, mixed, profile=True)
        use(mixed)
        D = _Dense(H, False)
        e = np.real(np.diag(D.gram(X.p, Y.p, NBANDS, NBANDS, X.rows)))
        res["mixed" if mixed else "fp64"] = {"sec": sec, "wall": wall, "eband": e, "prof": prof}
        tag = "mixed" if mixed else "fp64"
        print(f"# {tag}: {sec / args.steps * 1e3:.3f} ms/op_H", file=sys.stderr, flush=True)
        for k, (ms, n, b) in sorted(prof.items(), key=lambda kv: -kv[1][0]):
            if n:
                print(f"#   {tag} {k:11s} {ms / n * 1e3:9.1f} us/launch x{n // args.steps} per step, "
                      f"{b / n / (ms / n / 1e3) / 1e9:8.1f} GB/s algorithmic", file=sys.stderr, flush=True)
    rhoe = {}
    for mixed in (False, True):
        tag = "mixed" if mixed else "fp64"
        use(mixed)
        rhoe[tag] = time_rhoe(H, X, args.steps, args.warmup, mixed)
        print(f"# calc_rhoe {tag}: {rhoe[tag]['ms_per_step']:.3f} ms per 160 bands", file=sys.stderr, flush=True)
    Natoms = H.atoms.Natoms
    dE = abs(np.sum(res["mixed"]["eband"]) - np.sum(res["fp64"]["eband"])) * 2.0 / Natoms
    t_mix = res["mixed"]["sec"] / args.steps
    t_64 = res["fp64"]["sec"] / args.steps
    value = NBANDS / t_mix
    lob = {}
    scf = {}
    if args.solver_iters > 0:
        for mixed in (False, True):
            tag = "mixed" if mixed else "fp64"
            use(mixed)
            lob[tag] = time_lobpcg(H, X, args.solver_iters, mixed)
            print(f"# LOBPCG {tag}: {lob[tag]['s_per_iter'] * 1e3:.2f} ms/iteration, "
                  f"{lob[tag]['band_apps_per_s']:.0f} band-applications/s", file=sys.stderr, flush=True)
        lob["mixed_over_fp64"] = lob["mixed"]["band_apps_per_s"] / lob["fp64"]["band_apps_per_s"]
    if args.scf_iters > 0:
        for mixed in (False, True):
            tag = "mixed" if mixed else "fp64"
            use(mixed)
            scf[tag] = time_scf(H, args.scf_iters, mixed)
            print(f"# SCF {tag}: {scf[tag]['s_per_iter']:.3f} s/iteration", file=sys.stderr, flush=True)
    batched = time_batched_k(args.steps, args.warmup) if args.batched_k else None
    forces = forces_leg() if args.forces else None
    si216 = si216_leg() if args.si216 else None
    # CPU restatement on all of this process's host cores (the baseline) and on one core
    cpu = None if args.no_cpu else cpu_baseline(10.0, cpu_threads())
    cpu1 = None if args.no_cpu else cpu_baseline(10.0, 1)
    out = {
        "metric": "Hpsi band-applications/s (mixed precision; FP64 beside it)",
        "value": value,
        "unit": "band-applications/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_mix * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-fft/f32-mfma (mixed) vs f64",
        "data": "synthetic (seeded random orthonormal psi, atomic-superposition V_loc)",
        "config": {"workload": "Si64 sc a=20.5262, GTH LDA, ecutwfc 15 Ha, 72^3 grid, Gamma, 160 bands",
                   "Ngw": int(H.pw.Ngw[0]), "Nbeta": int(H.pspotNL.NbetaNL), "bands_per_step": NBANDS},
        "fp64": {"value": NBANDS / t_64, "ms_per_step": t_64 * 1e3},
        "mixed_over_fp64": t_64 / t_mix,
        "dE_per_atom_mixed_vs_fp64": dE,
        "op_H_streams": ("nonlocal products on a second stream beside the FFT passes"
                         if args.overlap and tuned["mixed"].get("overlap", 1) else "one stream"),
        "autotune": tuned,
        "roofline": roofline(res["mixed"]["prof"], H, True),
        "roofline_fp64": roofline(res["fp64"]["prof"], H, False),
        "lobpcg": lob or None,
        "batched_k": batched,
        "scf_s_per_iter": scf or None,
        "kernels_us_per_step": {t: {k: v[0] * 1e3 / args.steps for k, v in res[t]["prof"].items() if v[1]}
                                for t in ("fp64", "mixed")},
        "stages": {t: stages(res[t]["prof"], H, t == "mixed", args.steps) for t in ("fp64", "mixed")},
        "calc_rhoe": rhoe,
        "forces_si8": forces,
        "si216": si216,
        "cpu_baseline": cpu,
        "cpu_baseline_1core": cpu1,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
