"""CPU-only checks of the product's host side: its set-up data equals the oracle's, and the
C-ABI library loads and exports every entry point include/pwdft_hip.h declares (no compute
call is made: there is no GPU here)."""
import ctypes
import os
import re

import numpy as np
import pytest

from _systems import A_SI, ROOT, si_fcc_structure

import pwdft_hip.basis as B
import pwdft_hip.potentials as P
from oracle import setup as S
from oracle.density import guess_rhoe as o_guess
from oracle.energy import calc_E_NN


def _pair(ecut=15.0, meshk=(2, 2, 2), shift=(0, 0, 0)):
    sym, frac = si_fcc_structure()
    k, wk = B.kpoint_grid(B.gen_lattice_fcc(A_SI), meshk, shift)
    ko, wko = S.gen_kpoint_grid(S.gen_lattice_fcc(A_SI), meshk, shift)
    assert np.array_equal(k, ko) and np.array_equal(wk, wko)
    pw = B.PWGrid(ecut, B.gen_lattice_fcc(A_SI), kpoints=(k, wk))
    opw = S.PWGrid(ecut, S.gen_lattice_fcc(A_SI), kpoints=(ko, wko))
    return sym, frac, pw, opw


def test_pwgrid_maps_identical():
    _, _, pw, opw = _pair()
    assert pw.Ns == opw.Ns
    assert np.array_equal(pw.idx_g2r, opw.idx_g2r)
    assert np.array_equal(pw.G, opw.G) and np.array_equal(pw.G2_shells, opw.G2_shells)
    for ik in range(pw.Nkpt):
        assert np.array_equal(pw.idx_gw2r[ik], opw.idx_gw2r[ik])
        assert np.array_equal(pw.gk2(ik), opw.gk2(ik))


def test_kpoints_shifted_al_mesh():
    k, wk = B.kpoint_grid(B.gen_lattice_fcc(7.6), (4, 4, 4), (1, 1, 1))
    ko, wko = S.gen_kpoint_grid(S.gen_lattice_fcc(7.6), (4, 4, 4), (1, 1, 1))
    assert np.array_equal(k, ko) and np.allclose(wk, wko)
    assert k.shape[1] == 32   # 64 shifted points pair up under time reversal


def test_projectors_and_local_potential_identical():
    sym, frac, pw, opw = _pair()
    at = B.Atoms(sym, frac, B.gen_lattice_fcc(A_SI))
    oat = S.Atoms(sym, frac, S.gen_lattice_fcc(A_SI))
    psp = [B.PsPot_GTH("Si-q4.lda.gth")]
    opsp = [S.PsPot_GTH(os.path.join(ROOT, "pwdft.jl_amd", "pwdft_hip", "pseudopotentials", "Si-q4.lda.gth"))]
    assert np.array_equal(psp[0].h, opsp[0].h) and np.array_equal(psp[0].c, opsp[0].c)
    nl, onl = B.PsPotNL(at, pw, psp), S.PsPotNL(oat, opw, opsp)
    assert np.array_equal(nl.D, onl.D)
    for ik in range(pw.Nkpt):
        assert np.abs(nl.betaNL[ik] - onl.betaNL[ik]).max() < 1e-13

    class OH:  # the oracle's guess_rhoe reads these fields
        pass
    oh = OH()
    oh.atoms, oh.pw, oh.pspots, oh.Nelectrons = oat, opw, opsp, 8.0
    r, orr = P.guess_rhoe(at, pw, psp, 8.0), o_guess(oh)
    assert np.abs(r - orr).max() < 1e-14
    assert abs(P.calc_E_NN(at, [4.0]) - calc_E_NN(oat.LatVecs, oat, np.array([4.0]))) < 1e-12


def test_host_projectors_d_f_channels_match_oracle():
    """The host projector builder (kept for the lazy PsPotNL.betaNL) against the oracle for
    a species with d and f channels (Ce GTH-PADE-q12, l = 0..3)."""
    L = B.gen_lattice_fcc(A_SI)
    ko, wko = S.gen_kpoint_grid(S.gen_lattice_fcc(A_SI), (2, 2, 2), (0, 0, 0))
    pw = B.PWGrid(12.0, L, kpoints=(ko, wko))
    opw = S.PWGrid(12.0, S.gen_lattice_fcc(A_SI), kpoints=(ko, wko))
    at = B.Atoms(["Ce", "Si"], [[0, 0, 0], [0.25, 0.25, 0.25]], L)
    oat = S.Atoms(["Ce", "Si"], [[0, 0, 0], [0.25, 0.25, 0.25]], S.gen_lattice_fcc(A_SI))
    files = ["Ce-q12.lda.gth", "Si-q4.lda.gth"]
    psp = [B.PsPot_GTH(f) for f in files]
    opsp = [S.PsPot_GTH(os.path.join(ROOT, "pwdft.jl_amd", "pwdft_hip", "pseudopotentials", f)) for f in files]
    assert psp[0].lmax == 3
    nl, onl = B.PsPotNL(at, pw, psp), S.PsPotNL(oat, opw, opsp)
    assert np.array_equal(nl.D, onl.D)
    for ik in range(pw.Nkpt):
        assert np.abs(nl.betaNL[ik] - onl.betaNL[ik]).max() < 1e-13
    assert len(nl.projector_table()) == nl.NbetaNL


def test_electrons_counts():
    from pwdft_hip.hamiltonian import Electrons
    e = Electrons(8.0, 4, extra_states=4)
    assert (e.Nstates, e.Nstates_occ) == (8, 4) and e.Focc[:4].sum() == 32
    a = Electrons(3.0, 2, Nstates=8)
    assert (a.Nstates, a.Nstates_occ) == (8, 2) and a.Focc[1, 0] == 1.0
    with pytest.raises(ValueError):
        Electrons(8.0, 1, Nstates=3)


def _header_symbols():
    txt = open(os.path.join(ROOT, "include", "pwdft_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(pw_\w+)\s*\(", txt, flags=re.M)))


def test_library_exports_every_declared_symbol():
    from pwdft_hip import _lib
    path = _lib.LIB_PATH
    assert os.path.exists(path), "run __graft_entry__.build() first"
    lib = ctypes.CDLL(path)
    syms = _header_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(lib, s), s
    assert sorted(_lib.EXPORTED) == syms


def test_null_context_is_an_error_not_a_crash():
    from pwdft_hip import _lib
    lib = _lib.load()
    assert lib.pw_set_grid(None, 8, 8, 8, 1.0) == -1
    assert lib.pw_last_error(None) == b"null context"


def test_symmetry_kpoint_counts():
    """Crystal operations and the reduced k-meshes: 48 operations for diamond Si and fcc
    Al; IBZ sizes 8 (Si 4x4x4, test/test_spglib.jl:19), 10 (4x4x4 shifted), 60 (8x8x8
    shifted), 4 with weights 1:8:6:12 (Si 3x3x3) and 182 (Al 12x12x12 shifted); a displaced
    Si8 cell keeps only the identity."""
    sym, frac = ["Si", "Si"], [[0, 0, 0], [0.25, 0.25, 0.25]]
    si = B.Atoms(sym, frac, B.gen_lattice_fcc(A_SI))
    ops = B.crystal_symmetry(si)
    assert len(ops) == 48
    rots = B.kpoint_rotations(ops)
    for mesh, sh, n in (((4, 4, 4), (0, 0, 0), 8), ((4, 4, 4), (1, 1, 1), 10), ((8, 8, 8), (1, 1, 1), 60)):
        assert B.kpoint_grid(si.LatVecs, mesh, sh, rots=rots)[0].shape[1] == n
    k, wk = B.kpoint_grid(si.LatVecs, (3, 3, 3), (0, 0, 0), rots=rots)
    assert np.allclose(np.sort(wk * 27), [1, 6, 8, 12])
    al = B.Atoms(["Al"], [[0, 0, 0]], B.gen_lattice_fcc(7.6))
    assert B.kpoint_grid(al.LatVecs, (12, 12, 12), (1, 1, 1),
                         rots=B.kpoint_rotations(B.crystal_symmetry(al)))[0].shape[1] == 182
    s8 = B.Atoms(["Si"] * 8, B.diamond_supercell(1.0, 1), B.gen_lattice_sc(A_SI))
    s8.positions += np.random.default_rng(0).uniform(-0.05, 0.05, size=s8.positions.shape)
    assert len(B.crystal_symmetry(s8)) == 1


def _oracle_ops(ops):
    """The reference's convention for crystal_symmetry's {R|t}: inv_s = R^-T, ft = -t."""
    return [np.rint(np.linalg.inv(R).T).astype(np.int64) for R, _ in ops], [-t for _, t in ops]


def test_rhoe_symmetrizer_and_vectors():
    """The oracle's symmetrize_rhoe! with the operations crystal_symmetry finds for Si fcc
    (48, half of them non-symmorphic): idempotent, charge-conserving, the identity on the
    symmetric atomic-superposition density, equal to the group average
    rho_sym(m) = 1/Nsym sum_s rho(R^-T m) exp(2 pi i t . R^-T m) on the sphere, and its output
    is invariant under every operation (rho(R x + t) = rho(x) on the grid for the
    symmorphic ones).  symmetrize_vector!: a projection onto symmetric fields."""
    from pwdft_hip.potentials import guess_rhoe
    from oracle.symmetry import symmetrize_rhoe
    si = B.Atoms(["Si", "Si"], [[0, 0, 0], [0.25, 0.25, 0.25]], B.gen_lattice_fcc(A_SI))
    ops = B.crystal_symmetry(si)
    assert len(ops) == 48 and sum(np.any(t != 0) for _, t in ops) == 24
    inv_s, ft = _oracle_ops(ops)
    pw = B.PWGrid(8.0, si.LatVecs)
    S = lambda r: symmetrize_rhoe(pw, inv_s, ft, r)
    rho = np.random.default_rng(0).uniform(size=pw.Npoints)
    f = np.fft.fftn(rho.reshape(pw.Ns, order="F")).reshape(-1, order="F")
    keep = np.zeros(pw.Npoints, dtype=bool)
    keep[pw.idx_g2r] = True
    f[~keep] = 0.0   # band-limited to the rho sphere: outside it the reference keeps rho(G)
    rho = np.real(np.fft.ifftn(f.reshape(pw.Ns, order="F")).reshape(-1, order="F"))
    r1 = S(rho)
    assert np.abs(S(r1) - r1).max() < 1e-13
    assert abs(r1.sum() - rho.sum()) < 1e-9 * abs(rho).sum()
    # the group average, entry by entry
    m = np.rint(si.LatVecs.T @ pw.G / (2 * np.pi)).astype(np.int64)
    where = {tuple(m[:, g]): g for g in range(pw.Ng)}
    v = f[pw.idx_g2r]
    avg = np.zeros(pw.Ng, dtype=complex)
    for R, t in ops:
        ms = np.rint(np.linalg.inv(R).T).astype(np.int64) @ m
        avg += v[[where[tuple(c)] for c in ms.T]] * np.exp(2j * np.pi * (t @ ms))
    f1 = np.fft.fftn(r1.reshape(pw.Ns, order="F")).reshape(-1, order="F")
    assert np.abs(f1[pw.idx_g2r] - avg / len(ops)).max() < 1e-10 * np.abs(v).max()
    g = guess_rhoe(si, pw, [B.PsPot_GTH("Si-q4.lda.gth")], 8.0)
    assert np.abs(S(g) - g).max() < 1e-14
    n = np.array(pw.Ns)
    x = np.stack(np.meshgrid(*[np.arange(k) for k in pw.Ns], indexing="ij"), axis=-1).reshape(-1, 3)
    lin = lambda q: q[:, 0] + n[0] * (q[:, 1] + n[1] * q[:, 2])
    for R, t in ops:
        tn = t * n
        if np.abs(tn - np.round(tn)).max() > 1e-9:
            continue
        y = (x @ R.T + np.round(tn).astype(int)) % n
        assert np.abs(r1[lin(y)] - r1[lin(x)]).max() < 1e-12
    # the product's set-up tables: M = R^-T and the sphere's crystal coordinates
    T = B.RhoeSymmetrizer(si, pw, ops)
    assert np.array_equal(T.m, m.T) and all(np.array_equal(T.M[s], inv_s[s]) for s in range(len(ops)))
    vv = np.random.default_rng(1).standard_normal((3, 2))
    vs = B.symmetrize_vector(si, ops, vv)
    assert np.abs(B.symmetrize_vector(si, ops, vs) - vs).max() < 1e-14
    assert np.abs(vs).max() < 1e-14   # no symmetric force on the diamond sites


def test_packed_fp32_guard_catches_packed_instructions(tmp_path):
    """The build's check disassembles every gfx950 code object: a float2 multiply-add
    compiled without -packed-fp32-ops yields v_pk_fma_f32 and is refused; the same source
    with the library's flags passes."""
    import subprocess
    from pwdft_hip import _build as B
    src = tmp_path / "pk.hip"
    src.write_text("#include <hip/hip_runtime.h>\n"
                   "extern \"C\" __global__ void k(float2* a, const float2* b, float s, int n) {\n"
                   "  int i = blockIdx.x * blockDim.x + threadIdx.x;\n"
                   "  if (i < n) { float2 x = a[i], y = b[i]; x.x = x.x * s + y.x; x.y = x.y * s + y.y; a[i] = x; }\n"
                   "}\n")
    base = ["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O3", "-fPIC", "-shared"]
    bad, good = tmp_path / "bad.so", tmp_path / "good.so"
    subprocess.run(base + ["-o", str(bad), str(src)], check=True, capture_output=True)
    subprocess.run(["/opt/rocm/bin/hipcc"] + B.HIPCC_FLAGS + ["-o", str(good), str(src)], check=True,
                   capture_output=True)
    with pytest.raises(RuntimeError, match="packed FP32"):
        B.check_no_packed_fp32(str(bad))
    assert B.check_no_packed_fp32(str(good)) == 1


def test_autotune_cache_round_trip(tmp_path):
    """The tuning cache: shape keys, merged writes, and settings applied in dependency order
    (the FFT radix mode first) -- host logic only, on a stand-in context."""
    from pwdft_hip import autotune as at
    path = str(tmp_path / "sub" / "tune.json")
    k1 = at.shape_key((72, 72, 72), 23871, 320, 160, True)
    k2 = at.shape_key((20, 20, 20), 316, 4, 8, False)
    assert k1 == "72x72x72/ngw23871/nbeta320/nb160/mixed" and k1 != k2
    assert at.load_cache(path) == {}
    at.save_cache({k1: {"fft_order": 1, "band_batch": 96}}, path)
    at.save_cache({k2: {"fft_order": 0}}, path)
    d = at.load_cache(path)
    assert d[k1]["band_batch"] == 96 and d[k2]["fft_order"] == 0

    class Ctx:
        calls = []

        def call(self, name, *args):
            self.calls.append((name, args))

    ctx = Ctx()
    at.apply_settings(ctx, {"band_batch": 32, "gemm_nn_shape": 4, "fft_order": 2, "unknown": 7})
    assert ctx.calls == [("pw_tune_set", (at.TUNE_FFT_ORDER, 2)), ("pw_tune_set", (at.TUNE_BAND_BATCH, 32)),
                         ("pw_tune_set", (at.TUNE_GEMM_NN_SHAPE, 4))]
