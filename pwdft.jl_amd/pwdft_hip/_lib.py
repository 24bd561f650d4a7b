"""ctypes binding of libpwdft_hip.so (include/pwdft_hip.h).

This is the Python counterpart of the Julia `ccall` shim (julia/PWDFTHip.jl): the same
C ABI, the same column-major complex128 layout.  Loading fails loudly when the library
has not been built -- there is no CPU fallback anywhere in this package.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpwdft_hip.so")

PW_OK = 0
PW_ERR_NOT_POSDEF = -4
PW_TERM_K = 1
PW_TERM_VLOC = 2
PW_TERM_NLOC = 4
PW_TERM_ALL = 7
PW_ACCUMULATE = 16
PW_MIXED = 32
PW_XC_VWN = 0
PW_XC_PBE = 1
PW_TUNE_PLANE_COMPRESSED = 5
PW_TUNE_EIG_PATH = 9
PW_TUNE_LOBPCG_LOCK = 10
PW_TUNE_FUSED_BATCH = 11
PW_TUNE_LOBPCG_GRAPH = 12
PW_TUNE_FUSED_NT = 13
PW_TUNE_BATCH_GEMM_SMALL = 16
PW_TUNE_EIG_WG = 17
PW_TUNE_EIG_TRI = 18
PW_TUNE_CHOL_REC = 19

PROF_SLOTS = {"col_ifft": 0, "plane_vloc": 1, "col_fft": 2, "gemm_cn": 3, "gemm_nn": 4, "plane_rhoe": 5}

_c = ctypes
_vp = _c.c_void_p
_i = _c.c_int
_i64 = _c.c_int64
_d = _c.c_double
_dp = _c.POINTER(_c.c_double)

# name -> argtypes (all return int except pw_last_error)
_PROTOS = {
    "pw_ctx_create": [_c.POINTER(_vp), _i],
    "pw_ctx_destroy": [_vp],
    "pw_set_band_batch": [_vp, _i],
    "pw_tune_set": [_vp, _i, _i],
    "pw_set_grid": [_vp, _i, _i, _i, _d],
    "pw_set_kpoint": [_vp, _i, _i, _vp, _i, _vp],
    "pw_set_vloc": [_vp, _i, _vp],
    "pw_set_dij": [_vp, _i, _vp],
    "pw_set_projectors": [_vp, _i, _i, _vp, _i64],
    "pw_op_H": [_vp, _i, _i, _i, _vp, _i64, _vp, _i64, _c.c_uint],
    "pw_op_H_dev": [_vp, _i, _i, _i, _vp, _i64, _vp, _i64, _c.c_uint],
    "pw_op_H_dev_noalloc": [_vp, _i, _i, _i, _vp, _i64, _vp, _i64, _c.c_uint],
    "pw_kprec_dev": [_vp, _i, _i, _vp, _i64, _vp, _i64],
    "pw_betaNL_psi_dev": [_vp, _i, _i, _vp, _i64, _vp, _i64, _c.c_uint],
    "pw_rhoe_zero": [_vp],
    "pw_rhoe_accum": [_vp, _i, _i, _vp, _i64, _vp, _c.c_uint],
    "pw_rhoe_accum_dev": [_vp, _i, _i, _vp, _i64, _vp, _c.c_uint],
    "pw_rhoe_get": [_vp, _vp],
    "pw_op_H_batch_dev": [_vp, _i, _vp, _i, _i, _vp, _vp, _vp, _vp, _c.c_uint],
    "pw_kprec_batch_dev": [_vp, _i, _vp, _i, _vp, _vp, _vp, _vp],
    "pw_copy_cols_batch_dev": [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp],
    "pw_herm_fix_dev": [_vp, _i, _vp, _i64, _i],
    "pw_potrf_dev": [_vp, _i, _vp, _i64],
    "pw_trtri_dev": [_vp, _i, _vp, _i64],
    "pw_heevd_dev": [_vp, _i, _vp, _i64, _vp],
    "pw_heevd_batch_dev": [_vp, _i, _i, _vp, _i64, _i64, _vp],
    "pw_herm_fix_batch_dev": [_vp, _i, _i, _vp, _i64, _i64, _i],
    "pw_chol_inv_batch_dev": [_vp, _i, _i, _vp, _i64, _i64],
    "pw_chol_inv_batch_async_dev": [_vp, _i, _i, _vp, _i64, _i64, _vp],
    "pw_heevd_batch_async_dev": [_vp, _i, _i, _vp, _i64, _i64, _vp, _vp],
    "pw_heev_values_dev": [_vp, _i, _vp, _i64, _vp],
    "pw_heev_lowest_dev": [_vp, _i, _i, _vp, _i64, _vp],
    "pw_hegvd_dev": [_vp, _i, _vp, _i64, _vp, _i64, _vp],
    "pw_diag_lobpcg_dev": [_vp, _i, _i, _i, _vp, _i64, _d, _i, _i, _c.c_uint, _vp, _vp],
    "pw_diag_davidson_dev": [_vp, _i, _i, _i, _vp, _i64, _d, _i, _i, _c.c_uint, _vp, _vp],
    "pw_diag_lobpcg_batch_dev": [_vp, _i, _vp, _i, _i, _vp, _vp, _d, _i, _i, _c.c_uint, _vp, _vp],
    "pw_zgemm_batch_dev": [_vp, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "pw_set_atoms": [_vp, _i, _vp, _vp, _i, _vp],
    "pw_set_rho_sphere": [_vp, _i, _vp, _i, _vp, _i, _vp],
    "pw_E_Ps_loc_grad": [_vp, _vp, _vp, _vp, _vp],
    "pw_calc_forces_Ps_loc": [_vp, _vp, _vp],
    "pw_calc_forces_Ps_nloc_dev": [_vp, _i, _i, _vp, _i64, _vp, _vp, _vp],
    "pw_calc_forces_Ps_nloc": [_vp, _i, _i, _vp, _i64, _vp, _vp, _vp],
    "pw_zgemm_dev": [_vp, _i, _i, _i, _i, _vp, _vp, _i64, _vp, _i64, _vp, _vp, _i64, _c.c_uint],
    "pw_malloc": [_vp, _c.c_size_t, _c.POINTER(_vp)],
    "pw_free": [_vp, _vp],
    "pw_memcpy_h2d": [_vp, _vp, _vp, _c.c_size_t],
    "pw_memcpy_d2h": [_vp, _vp, _vp, _c.c_size_t],
    "pw_memcpy_d2d": [_vp, _vp, _vp, _c.c_size_t],
    "pw_synchronize": [_vp],
    "pw_graph_begin": [_vp],
    "pw_graph_end": [_vp, _c.POINTER(_i)],
    "pw_graph_launch": [_vp, _i],
    "pw_graph_free": [_vp, _i],
    "pw_copy_cols_dev": [_vp, _i, _i, _vp, _i64, _vp, _i64],
    "pw_event_record": [_vp, _i],
    "pw_profile_enable": [_vp, _i],
    "pw_profile_read": [_vp, _i, _c.POINTER(_c.c_double), _c.POINTER(_c.c_long), _c.POINTER(_c.c_double)],
    "pw_event_elapsed_ms": [_vp, _i, _i, _c.POINTER(_c.c_float)],
    "pw_rhoe_get_dev": [_vp, _vp],
    "pw_set_vps_loc": [_vp, _vp],
    "pw_set_symmetry": [_vp, _i, _vp, _vp, _i, _vp],
    "pw_update_potential_dev": [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp],
    "pw_mix_rhoe_dev": [_vp, _d, _vp, _vp, _vp],
    "pw_symmetrize_rhoe_dev": [_vp, _vp],
    "pw_energies_psi_dev": [_vp, _i, _i, _vp, _i64, _vp, _vp],
    "pw_vec_dots_dev": [_vp, _i64, _i, _vp, _vp, _vp],
    "pw_vec_lincomb_dev": [_vp, _i64, _i, _vp, _vp, _d, _vp],
    "pw_mix_adaptive_dev": [_vp, _i64, _d, _d, _vp, _vp, _vp, _vp],
    "pw_vec_clamp_dev": [_vp, _i64, _d, _vp],
    "pw_build_projectors_dev": [_vp, _i, _i, _vp, _i, _vp],
    "pw_build_vps_loc_dev": [_vp, _vp],
    "pw_set_vloc_parts_dev": [_vp, _i, _vp, _vp, _vp],
    "pw_scale_cols_dev": [_vp, _i, _i, _vp, _vp, _i64],
}

EXPORTED = sorted(list(_PROTOS) + ["pw_last_error"])

_lib = None


class PWError(RuntimeError):
    """Error returned by the C ABI (status code + pw_last_error message)."""

    def __init__(self, code, msg):
        super().__init__(f"pwdft_hip error {code}: {msg}")
        self.code = code


class PosDefException(PWError):
    """Counterpart of Julia's LinearAlgebra.PosDefException raised by cholesky()."""


def load(path=None):
    """Load libpwdft_hip.so (built by __graft_entry__.build()); raise if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ImportError(f"libpwdft_hip.so not found at {p}; run __graft_entry__.build() first")
    lib = ctypes.CDLL(p)
    for name, args in _PROTOS.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = _i
    lib.pw_last_error.argtypes = [_vp]
    lib.pw_last_error.restype = _c.c_char_p
    _lib = lib
    return lib


def ptr(a):
    """Host pointer of a contiguous numpy array."""
    return ctypes.c_void_p(a.ctypes.data)


class Context:
    """Owner of one pw_ctx (one HIP stream and its device buffers)."""

    def __init__(self, device=0):
        self.lib = load()
        h = _vp()
        rc = self.lib.pw_ctx_create(ctypes.byref(h), device)
        if rc != PW_OK:
            raise PWError(rc, "pw_ctx_create failed (is a GPU visible?)")
        self.h = h

    def check(self, rc):
        if rc != PW_OK:
            msg = self.lib.pw_last_error(self.h).decode()
            raise (PosDefException if rc == PW_ERR_NOT_POSDEF else PWError)(rc, msg)

    def call(self, name, *args):
        self.check(getattr(self.lib, name)(self.h, *args))

    def profile(self, on):
        self.call("pw_profile_enable", 1 if on else 0)

    def profile_read(self):
        """{slot name: (total_ms, launches, algorithmic bytes)} since profile(True)."""
        out = {}
        for name, slot in PROF_SLOTS.items():
            ms, n, b = ctypes.c_double(), ctypes.c_long(), ctypes.c_double()
            self.call("pw_profile_read", slot, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(b))
            out[name] = (ms.value, n.value, b.value)
        return out

    def close(self):
        if self.h is not None and self.h.value:
            self.lib.pw_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """Stream capture of the device work issued inside a `with` block (pw_graph_begin /
    pw_graph_end); launch() replays it with one hipGraphLaunch."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.id = None

    def __enter__(self):
        self.ctx.call("pw_graph_begin")
        return self

    def __exit__(self, exc_type, exc, tb):
        gid = _i(-1)
        rc = self.ctx.lib.pw_graph_end(self.ctx.h, ctypes.byref(gid))
        if exc_type is not None:
            if rc == PW_OK:
                self.ctx.lib.pw_graph_free(self.ctx.h, gid.value)
            return False
        self.ctx.check(rc)
        self.id = gid.value
        return False

    def launch(self):
        self.ctx.call("pw_graph_launch", self.id)

    def free(self):
        if self.id is not None:
            self.ctx.call("pw_graph_free", self.id)
            self.id = None


class DeviceMatrix:
    """Column-major complex128 matrix in device memory (rows x cols, leading dim = rows)."""

    def __init__(self, ctx, rows, cols):
        self.ctx, self.rows, self.cols = ctx, int(rows), int(cols)
        p = _vp()
        ctx.call("pw_malloc", max(16, 16 * self.rows * self.cols), ctypes.byref(p))
        self.p = p

    @property
    def nbytes(self):
        return 16 * self.rows * self.cols

    def col(self, j):
        """Device pointer of column j."""
        return _vp(self.p.value + 16 * self.rows * int(j))

    @classmethod
    def from_host(cls, ctx, a):
        a = np.asfortranarray(a, dtype=np.complex128)
        m = cls(ctx, a.shape[0], a.shape[1] if a.ndim > 1 else 1)
        m.upload(a)
        return m

    def upload(self, a):
        a = np.asfortranarray(a, dtype=np.complex128)
        assert a.size == self.rows * self.cols
        self.ctx.call("pw_memcpy_h2d", self.p, ptr(a), self.nbytes)

    def download(self):
        a = np.empty((self.rows, self.cols), dtype=np.complex128, order="F")
        self.ctx.call("pw_memcpy_d2h", ptr(a), self.p, self.nbytes)
        return a

    def free(self):
        if self.p is not None and self.p.value:
            self.ctx.call("pw_free", self.p)
            self.p = None

    def __del__(self):
        try:
            if self.ctx.h is not None:
                self.free()
        except Exception:
            pass


class DeviceVector:
    """float64 vector in device memory (a density or a potential on the FFT grid)."""

    def __init__(self, ctx, n):
        self.ctx, self.n = ctx, int(n)
        p = _vp()
        ctx.call("pw_malloc", max(16, 8 * self.n), ctypes.byref(p))
        self.p = p

    @property
    def nbytes(self):
        return 8 * self.n

    @classmethod
    def from_host(cls, ctx, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        v = cls(ctx, a.size)
        v.upload(a)
        return v

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        assert a.size == self.n
        self.ctx.call("pw_memcpy_h2d", self.p, ptr(a), self.nbytes)

    def download(self):
        a = np.empty(self.n)
        self.ctx.call("pw_memcpy_d2h", ptr(a), self.p, self.nbytes)
        return a

    def copy_from(self, other):
        assert other.n == self.n
        self.ctx.call("pw_memcpy_d2d", self.p, other.p, self.nbytes)

    def free(self):
        if self.p is not None and self.p.value:
            self.ctx.call("pw_free", self.p)
            self.p = None

    def __del__(self):
        try:
            if self.ctx.h is not None:
                self.free()
        except Exception:
            pass
