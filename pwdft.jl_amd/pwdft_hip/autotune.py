"""Autotuner for the op_H launch parameters: the complex-GEMM tile shape and split-K
workgroup target of the two projector products (betaNL^H psi and beta (D B)), and the
number of bands per fused FFT launch.  Each candidate is timed with the library's own
per-kernel events (pw_profile_*) on the caller's Hamiltonian and wave functions; the
fastest setting is applied with pw_tune_set and returned."""
import itertools

from ._lib import PW_MIXED, PW_TERM_ALL

TUNE_GEMM_CN_SHAPE, TUNE_GEMM_NN_SHAPE, TUNE_GEMM_CN_WG, TUNE_GEMM_NN_WG, TUNE_BAND_BATCH = range(5)
SHAPES = {-1: "auto", 0: "64x64", 1: "128x64", 2: "64x128", 3: "128x128", 4: "128x32", 5: "64x32"}


def _profile_step(H, X, Y, flags, steps):
    c = H.ctx
    c.call("pw_op_H_dev", H.ik, H.ispin, X.cols, X.p, X.rows, Y.p, Y.rows, flags)  # warm
    c.profile(True)
    for _ in range(steps):
        c.call("pw_op_H_dev", H.ik, H.ispin, X.cols, X.p, X.rows, Y.p, Y.rows, flags)
    prof = c.profile_read()
    c.profile(False)
    return {k: v[0] / steps for k, v in prof.items()}


def tune_op_H(H, X, Y, mixed=None, steps=3, shapes=(0, 1, 2, 3, 4, 5), targets=(256, 512, 1024, 2048),
              batches=None):
    """Sweep GEMM shapes/targets (both products at once, timed separately) and optionally
    the FFT band batch; apply and return the best settings with the timings (ms/step)."""
    mixed = H.mixed if mixed is None else mixed
    flags = PW_TERM_ALL | (PW_MIXED if mixed else 0)
    c = H.ctx
    best = {"cn": (None, float("inf")), "nn": (None, float("inf"))}
    table = []
    for shape, target in itertools.product(shapes, targets):
        for key, v in ((TUNE_GEMM_CN_SHAPE, shape), (TUNE_GEMM_NN_SHAPE, shape),
                       (TUNE_GEMM_CN_WG, target), (TUNE_GEMM_NN_WG, target)):
            c.call("pw_tune_set", key, v)
        t = _profile_step(H, X, Y, flags, steps)
        table.append({"shape": SHAPES[shape], "target": target, "gemm_cn_ms": t["gemm_cn"],
                      "gemm_nn_ms": t["gemm_nn"]})
        if t["gemm_cn"] < best["cn"][1]:
            best["cn"] = ((shape, target), t["gemm_cn"])
        if t["gemm_nn"] < best["nn"][1]:
            best["nn"] = ((shape, target), t["gemm_nn"])
    (cs, ct), _ = best["cn"]
    (ns, nt), _ = best["nn"]
    for key, v in ((TUNE_GEMM_CN_SHAPE, cs), (TUNE_GEMM_CN_WG, ct), (TUNE_GEMM_NN_SHAPE, ns),
                   (TUNE_GEMM_NN_WG, nt)):
        c.call("pw_tune_set", key, v)
    chosen = {"gemm_cn": {"shape": SHAPES[cs], "target": ct, "ms": best["cn"][1]},
              "gemm_nn": {"shape": SHAPES[ns], "target": nt, "ms": best["nn"][1]}}
    if batches:
        bt = []
        for nb in batches:
            c.call("pw_tune_set", TUNE_BAND_BATCH, int(nb))
            t = _profile_step(H, X, Y, flags, steps)
            fft = t["col_ifft"] + t["plane_vloc"] + t["col_fft"]
            bt.append((fft, nb))
        fbest, nbest = min(bt)
        c.call("pw_tune_set", TUNE_BAND_BATCH, int(nbest))
        chosen["band_batch"] = {"bands": nbest, "fft_ms": fbest}
    return chosen, table
