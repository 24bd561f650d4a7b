// C ABI of libpwdft_hip: context, set-up of the state op_H reads, and the entry points
// declared in include/pwdft_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <new>

#include "pw_internal.h"

namespace pw {

int set_error(pw_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

int hip_check(pw_ctx* c, hipError_t e, const char* what) {
  if (e == hipSuccess) return PW_OK;
  return set_error(c, PW_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

int ensure(pw_ctx* c, void** buf, size_t* have, size_t need) {
  if (need <= *have && *buf) return PW_OK;
  if (c->no_alloc)
    return set_error(c, PW_ERR_ARG, "a workspace would grow inside a no-allocation call (" + std::to_string(need) +
                                        " bytes needed, " + std::to_string(*have) + " available)");
  if (c->capturing)
    return set_error(c, PW_ERR_STATE, "a workspace would grow during graph capture; run the sequence once first");
  if (*buf) {
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_check(c, e, "ensure/sync");
    if (c->graphs_live > 0)
      c->retired.push_back(*buf);  // a captured graph may still address it
    else
      hipFree(*buf);
    *buf = nullptr;
    *have = 0;
  }
  size_t sz = std::max<size_t>(need, 256);
  c->graph_gen++;
  hipError_t e = hipMalloc(buf, sz);
  if (e != hipSuccess) {
    *buf = nullptr;
    return set_error(c, PW_ERR_NOMEM, std::string("hipMalloc workspace: ") + hipGetErrorString(e));
  }
  *have = sz;
  return PW_OK;
}

// Radix passes of a length-n line.  mode 0: largest radices first (9, 8, 5, 4, 3, 2; the
// compile-time plans of 20, 27, 45, 72, 108 are of this form); 1: the same radices in
// reverse order; 2: radices 5, 4, 3, 2 only (more passes, more lanes per line).
static bool factorize(int n, FftDim& d, int mode = 0) {
  d.n = n;
  d.npass = 0;
  int rem = n;
  const int big[6] = {9, 8, 5, 4, 3, 2};
  const int small[6] = {5, 4, 3, 2, 0, 0};
  const int* order = mode == 2 ? small : big;
  while (rem > 1) {
    bool found = false;
    for (int q = 0; q < 6; ++q) {
      const int r = order[q];
      if (r == 0) break;
      if (rem % r == 0) {
        if (d.npass >= 8) return false;
        d.radix[d.npass++] = r;
        rem /= r;
        found = true;
        break;
      }
    }
    if (!found) return false;
  }
  if (d.npass == 0) {  // n == 1: identity
    d.radix[d.npass++] = 1;
    return false;
  }
  if (mode == 1) std::reverse(d.radix, d.radix + d.npass);
  return true;
}

// FFT plans of the three grid dimensions for a radix mode (factorize); false when a
// dimension has no plan that fits one wavefront.
static bool plan_grid(const int ns[3], int mode, FftDim dims[3]) {
  for (int d = 0; d < 3; ++d) {
    if (!factorize(ns[d], dims[d], mode)) return false;
    plan_lanes(dims[d], 12);
    if (dims[d].gpw < 1 || dims[d].L > 64) return false;
  }
  return true;
}

void free_k(KData& k) {
  hipFree(k.d_ent_igw);
  hipFree(k.d_ent_pos);
  hipFree(k.d_col_ptr);
  hipFree(k.d_col_ij);
  hipFree(k.d_col_cy);
  hipFree(k.d_xact);
  hipFree(k.d_gk2);
  hipFree(k.d_gk);
  hipFree(k.d_gpos);
  hipFree(k.d_beta64);
  hipFree(k.d_betaD64);
  hipFree(k.d_beta32);
  hipFree(k.d_betaD32);
  k = KData();
}

template <typename T>
int upload(pw_ctx* c, T** dst, const T* src, size_t n) {
  c->graph_gen++;
  if (*dst) {
    hipFree(*dst);
    *dst = nullptr;
  }
  hipError_t e = hipMalloc(reinterpret_cast<void**>(dst), std::max<size_t>(n * sizeof(T), 16));
  if (e != hipSuccess) return set_error(c, PW_ERR_NOMEM, "hipMalloc in set-up");
  e = hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice);
  return hip_check(c, e, "hipMemcpy in set-up");
}

__global__ void k_zero_cols(cplx<double>* __restrict__ a, long ld, int nrows, int ncols) {
  const long tot = (long)nrows * ncols;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < tot; t += (long)gridDim.x * blockDim.x)
    a[(t / nrows) * ld + (t % nrows)] = mk(0.0, 0.0);
}

__global__ void k_copy_cols(const cplx<double>* __restrict__ s, long lds, cplx<double>* __restrict__ d,
                            long ldd, int nrows, int ncols) {
  const long tot = (long)nrows * ncols;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < tot; t += (long)gridDim.x * blockDim.x) {
    const long j = t / nrows, i = t % nrows;
    d[j * ldd + i] = s[j * lds + i];
  }
}

static int launch_blocks(long tot) { return (int)std::max<long>(1, std::min<long>((tot + 255) / 256, 4096)); }

int copy_cols(pw_ctx* c, int nrows, int ncols, const cplx<double>* src, long lds, cplx<double>* dst, long ldd) {
  if (nrows <= 0 || ncols <= 0) return PW_OK;
  hipLaunchKernelGGL(k_copy_cols, dim3(launch_blocks((long)nrows * ncols)), dim3(256), 0, c->stream, src, lds, dst, ldd,
                     nrows, ncols);
  return hip_check(c, hipGetLastError(), "copy_cols");
}

static int check_k(pw_ctx* c, int ik) {
  if (c->N == 0) return set_error(c, PW_ERR_STATE, "pw_set_grid has not been called");
  if (ik < 0 || ik >= (int)c->kd.size() || c->kd[ik].ngw == 0)
    return set_error(c, PW_ERR_STATE, "k-point " + std::to_string(ik) + " is not set up");
  return PW_OK;
}

int op_H_device(pw_ctx* c, int ik, int ispin, int nb, const cplx<double>* psi, long ldpsi, cplx<double>* hpsi,
                long ldh, unsigned flags) {
  int rc = check_k(c, ik);
  if (rc) return rc;
  KData& k = c->kd[ik];
  if (nb < 0 || ldpsi < k.ngw || ldh < k.ngw) return set_error(c, PW_ERR_ARG, "op_H: bad nb or leading dimension");
  if (nb == 0) return PW_OK;
  const unsigned terms = flags & PW_TERM_ALL;
  bool have = (flags & PW_ACCUMULATE) != 0;
  // (small problems -- a few bands of a small sphere -- keep one stream: the fork and join
  // would cost more than the overlap saves)
  if ((terms & PW_TERM_VLOC) && (terms & PW_TERM_NLOC) && k.nbeta > 0 && c->overlap && c->stream2 &&
      (long)k.ngw * nb >= (1L << 18)) {
    // The nonlocal products (MFMA) run on stream2 beside the column / plane FFT passes
    // (VALU / LDS): stream2 writes Hpsi = [Hpsi +] (beta D) (beta^H psi) first, and the gather
    // pass, which waits for it, adds the kinetic and local terms.
    if (ispin < 0 || ispin >= c->nspin || !c->d_v64[ispin])
      return set_error(c, PW_ERR_STATE, "op_V_loc: potential of spin " + std::to_string(ispin) + " not set");
    rc = ensure(c, &c->d_bnl, &c->bnl_bytes, (size_t)k.nbeta * nb * sizeof(cplx<double>));
    if (rc) return rc;
    hipError_t e = hipEventRecord(c->ev_fork, c->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(c->stream2, c->ev_fork, 0);
    if (e != hipSuccess) return hip_check(c, e, "stream fork");
    cplx<double>* bnl = reinterpret_cast<cplx<double>*>(c->d_bnl);
    const cplx<double> one = mk(1.0, 0.0), zero = mk(0.0, 0.0);
    hipStream_t main = c->stream;
    c->stream = c->stream2;  // the GEMM drivers launch on c->stream
    if (flags & PW_MIXED) {
      rc = zgemm_f32in(c, 1, k.nbeta, nb, k.ngw, one, reinterpret_cast<const cplx<float>*>(k.d_beta32), k.ngw,
                       psi, ldpsi, zero, bnl, k.nbeta, 1);
      if (!rc)
        rc = zgemm_f32in(c, 0, k.ngw, nb, k.nbeta, one, reinterpret_cast<const cplx<float>*>(k.d_betaD32), k.ngw,
                         bnl, k.nbeta, have ? one : zero, hpsi, ldh, 1);
    } else {
      rc = zgemm(c, 1, k.nbeta, nb, k.ngw, one, reinterpret_cast<const cplx<double>*>(k.d_beta64), k.ngw, psi,
                 ldpsi, zero, bnl, k.nbeta, 0);
      if (!rc)
        rc = zgemm(c, 0, k.ngw, nb, k.nbeta, one, reinterpret_cast<const cplx<double>*>(k.d_betaD64), k.ngw, bnl,
                   k.nbeta, have ? one : zero, hpsi, ldh, 0);
    }
    e = hipEventRecord(c->ev_join, c->stream2);
    c->stream = main;
    if (!rc && e != hipSuccess) rc = hip_check(c, e, "stream join record");
    if (rc) {
      hipStreamWaitEvent(c->stream, c->ev_join, 0);  // keep the streams ordered on failure
      return rc;
    }
    return vloc_apply(c, ik, ispin, nb, psi, ldpsi, hpsi, ldh, flags | PW_ACCUMULATE, c->ev_join);
  }
  if (terms & PW_TERM_VLOC) {
    if (ispin < 0 || ispin >= c->nspin || !c->d_v64[ispin])
      return set_error(c, PW_ERR_STATE, "op_V_loc: potential of spin " + std::to_string(ispin) + " not set");
    rc = vloc_apply(c, ik, ispin, nb, psi, ldpsi, hpsi, ldh, flags);
    if (rc) return rc;
    have = true;
  } else if (terms & PW_TERM_K) {
    rc = kinetic_apply(c, ik, nb, psi, ldpsi, hpsi, ldh, flags);
    if (rc) return rc;
    have = true;
  }
  const bool nloc = (terms & PW_TERM_NLOC) && k.nbeta > 0;
  if (nloc) {
    rc = ensure(c, &c->d_bnl, &c->bnl_bytes, (size_t)k.nbeta * nb * sizeof(cplx<double>));
    if (rc) return rc;
    cplx<double>* bnl = reinterpret_cast<cplx<double>*>(c->d_bnl);
    const cplx<double> one = mk(1.0, 0.0), zero = mk(0.0, 0.0);
    if (flags & PW_MIXED) {
      rc = zgemm_f32in(c, 1, k.nbeta, nb, k.ngw, one, reinterpret_cast<const cplx<float>*>(k.d_beta32), k.ngw,
                       psi, ldpsi, zero, bnl, k.nbeta, 1);
      if (rc) return rc;
      rc = zgemm_f32in(c, 0, k.ngw, nb, k.nbeta, one, reinterpret_cast<const cplx<float>*>(k.d_betaD32),
                       k.ngw, bnl, k.nbeta, have ? one : zero, hpsi, ldh, 1);
    } else {
      rc = zgemm(c, 1, k.nbeta, nb, k.ngw, one, reinterpret_cast<const cplx<double>*>(k.d_beta64), k.ngw, psi,
                 ldpsi, zero, bnl, k.nbeta, 0);
      if (rc) return rc;
      rc = zgemm(c, 0, k.ngw, nb, k.nbeta, one, reinterpret_cast<const cplx<double>*>(k.d_betaD64), k.ngw,
                 bnl, k.nbeta, have ? one : zero, hpsi, ldh, 0);
    }
    if (rc) return rc;
    have = true;
  }
  if (!have) {
    hipLaunchKernelGGL(k_zero_cols, dim3(launch_blocks((long)k.ngw * nb)), dim3(256), 0, c->stream, hpsi, ldh,
                       k.ngw, nb);
    return hip_check(c, hipGetLastError(), "k_zero_cols");
  }
  return PW_OK;
}

static int stage(pw_ctx* c, size_t bytes) { return ensure(c, &c->d_tmp, &c->tmp_bytes, bytes); }

static hipEvent_t pool_event(pw_ctx* c) {
  if (!c->ev_pool.empty()) {
    hipEvent_t e = c->ev_pool.back();
    c->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

ProfScope::ProfScope(pw_ctx* ctx, int s, double bytes) : c(ctx), slot(s) {
  if (!c->profile || c->capturing) return;
  hipEvent_t e0 = pool_event(c);
  e1 = pool_event(c);
  if (!e0 || !e1) {
    e1 = nullptr;
    return;
  }
  hipEventRecord(e0, c->stream);
  c->ev_used.push_back({slot, {e0, e1}});
  c->prof_bytes[slot] += bytes;
}

ProfScope::~ProfScope() {
  if (e1) hipEventRecord(e1, c->stream);
}

static void drain_events(pw_ctx* c) {
  if (c->ev_used.empty()) return;
  hipStreamSynchronize(c->stream);
  for (auto& u : c->ev_used) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, u.second.first, u.second.second) == hipSuccess) {
      c->prof_ms[u.first] += ms;
      c->prof_count[u.first] += 1;
    }
    c->ev_pool.push_back(u.second.first);
    c->ev_pool.push_back(u.second.second);
  }
  c->ev_used.clear();
}

}  // namespace pw

using namespace pw;

#define CTX_CHECK(c) \
  if (!(c)) return PW_ERR_ARG

extern "C" {

int pw_ctx_create(pw_ctx** out, int device) {
  if (!out) return PW_ERR_ARG;
  *out = nullptr;
  pw_ctx* c = new (std::nothrow) pw_ctx();
  if (!c) return PW_ERR_NOMEM;
  c->device = device;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming);
  if (e != hipSuccess) {
    delete c;
    return PW_ERR_HIP;
  }
  if (noalloc_init(c) != PW_OK) {
    pw_ctx_destroy(c);
    return PW_ERR_HIP;
  }
  *out = c;
  return PW_OK;
}

int pw_ctx_destroy(pw_ctx* c) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  hipStreamSynchronize(c->stream);
  free_lob_graphs(c);
  for (auto& k : c->kd) free_k(k);
  for (int d = 0; d < 3; ++d) {
    hipFree(c->d_tw64[d]);
    hipFree(c->d_tw32[d]);
  }
  for (auto p : c->d_v64) hipFree(p);
  for (auto p : c->d_v32) hipFree(p);
  hipFree(c->d_work);
  hipFree(c->d_yglob);
  hipFree(c->d_gpart);
  hipFree(c->d_bnl);
  hipFree(c->d_rpart);
  hipFree(c->d_rho);
  hipFree(c->d_tmp);
  hipFree(c->d_wts);
  free_forces(c);
  free_scf(c);
  free_desc(c);
  free_dense(c);
  for (void* g : c->graphs)
    if (g) hipGraphExecDestroy(reinterpret_cast<hipGraphExec_t>(g));
  for (void* g : c->graph_defs)
    if (g) hipGraphDestroy(reinterpret_cast<hipGraph_t>(g));
  for (void* p : c->retired) hipFree(p);
  drain_events(c);
  for (auto e : c->ev_pool) hipEventDestroy(e);
  if (c->ev_fork) hipEventDestroy(c->ev_fork);
  if (c->ev_join) hipEventDestroy(c->ev_join);
  if (c->stream2) hipStreamDestroy(c->stream2);
  hipStreamDestroy(c->stream);
  delete c;
  return PW_OK;
}

const char* pw_last_error(pw_ctx* c) { return c ? c->err.c_str() : "null context"; }

int pw_set_band_batch(pw_ctx* c, int nb) {
  CTX_CHECK(c);
  if (nb < 1) return set_error(c, PW_ERR_ARG, "band batch must be >= 1");
  c->band_batch = nb;
  return PW_OK;
}

int pw_tune_set(pw_ctx* c, int key, int value) {
  CTX_CHECK(c);
  switch (key) {
    case PW_TUNE_GEMM_CN_SHAPE:
    case PW_TUNE_GEMM_NN_SHAPE:
      if (value < -1 || value > 5) return set_error(c, PW_ERR_ARG, "pw_tune_set: tile shape must be -1..5");
      c->tune_gemm_shape[key - PW_TUNE_GEMM_CN_SHAPE] = value;
      return PW_OK;
    case PW_TUNE_GEMM_CN_WG:
    case PW_TUNE_GEMM_NN_WG:
      if (value < 0) return set_error(c, PW_ERR_ARG, "pw_tune_set: workgroup target must be >= 0");
      c->tune_gemm_wg[key - PW_TUNE_GEMM_CN_WG] = value;
      return PW_OK;
    case PW_TUNE_BAND_BATCH:
      return pw_set_band_batch(c, value);
    case PW_TUNE_PLANE_COMPRESSED:
      if (value < 0 || value > 2) return set_error(c, PW_ERR_ARG, "pw_tune_set: plane mode must be 0..2");
      c->plane_compressed = value;
      return PW_OK;
    case PW_TUNE_GEMM_SPLIT_BF16:
      if (value < -1 || value > 1) return set_error(c, PW_ERR_ARG, "pw_tune_set: split mode must be -1..1");
      c->gemm_split_bf16 = value;
      return PW_OK;
    case PW_TUNE_COL_CHUNK:
      if (value < 0) return set_error(c, PW_ERR_ARG, "pw_tune_set: column chunk must be >= 0");
      c->tune_col_chunk = value;
      return PW_OK;
    case PW_TUNE_OVERLAP:
      c->overlap = value != 0;
      return PW_OK;
    case PW_TUNE_FUSED_BATCH:
      if (value < 0 || value > 2) return set_error(c, PW_ERR_ARG, "pw_tune_set: fused batch mode must be 0..2");
      c->fused_batch = value;
      return PW_OK;
    case PW_TUNE_FFT_ORDER: {
      if (value < 0 || value > 2) return set_error(c, PW_ERR_ARG, "pw_tune_set: FFT radix mode must be 0..2");
      if (c->N > 0) {
        const int ns[3] = {c->n1, c->n2, c->n3};
        FftDim dims[3];
        if (!plan_grid(ns, value, dims))
          return set_error(c, PW_ERR_ARG, "pw_tune_set: no FFT plan of this radix mode fits the grid");
        for (int d = 0; d < 3; ++d) c->dim[d] = dims[d];
        c->graph_gen++;   // captured launches carry the old plans
      }
      c->fft_order = value;
      return PW_OK;
    }
    case PW_TUNE_BATCH_GEMM_SMALL:
      if (value < 0 || value > 1) return set_error(c, PW_ERR_ARG, "pw_tune_set: batched GEMM mode must be 0..1");
      c->batch_gemm_small = value;
      return PW_OK;
    case PW_TUNE_EIG_WG:
      if (value < 0 || value > 256) return set_error(c, PW_ERR_ARG, "pw_tune_set: reduction workgroups must be 0..256");
      c->eig_min_wg = value;
      return PW_OK;
    case PW_TUNE_CHOL_REC:
      if (value < 0 || value > 1) return set_error(c, PW_ERR_ARG, "pw_tune_set: Cholesky inverse mode must be 0..1");
      c->chol_rec = value;
      return PW_OK;
    case PW_TUNE_EIG_TRI:
      if (value < 0 || value > 2) return set_error(c, PW_ERR_ARG, "pw_tune_set: tridiagonal eigensolver mode must be 0..2");
      c->eig_tri = value;
      return PW_OK;
    case PW_TUNE_EIG_SYNC:
      if (value < 0 || value > 3) return set_error(c, PW_ERR_ARG, "pw_tune_set: eigensolver sync mode must be 0..3");
      c->eig_light = value;
      return PW_OK;
    case PW_TUNE_FUSED_NT:
      if (value != 0 && value != 256 && value != 512 && value != 1024)
        return set_error(c, PW_ERR_ARG, "pw_tune_set: fused workgroup width must be 0, 256, 512 or 1024");
      c->fused_nt = value;
      return PW_OK;
    case PW_TUNE_LOBPCG_GRAPH:
      c->lob_graph = value != 0;
      if (!c->lob_graph) free_lob_graphs(c);
      return PW_OK;
    case PW_TUNE_LOBPCG_LOCK:
      if (value < 0 || value > 300) return set_error(c, PW_ERR_ARG, "pw_tune_set: lock exponent must be 0..300");
      c->lobpcg_lock_exp = value;
      return PW_OK;
    case PW_TUNE_EIG_PATH:
      if (value < 0 || value > 1) return set_error(c, PW_ERR_ARG, "pw_tune_set: eigensolver path must be 0..1");
      c->eig_path = value;
      return PW_OK;
    default:
      return set_error(c, PW_ERR_ARG, "pw_tune_set: unknown key");
  }
}

int pw_set_grid(pw_ctx* c, int n1, int n2, int n3, double omega) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (n1 < 2 || n2 < 2 || n3 < 2 || !(omega > 0.0)) return set_error(c, PW_ERR_ARG, "pw_set_grid: bad sizes");
  const int ns[3] = {n1, n2, n3};
  FftDim dims[3];
  for (int d = 0; d < 3; ++d) {
    if (!factorize(ns[d], dims[d]))
      return set_error(c, PW_ERR_ARG, "pw_set_grid: size " + std::to_string(ns[d]) + " is not a product of 2,3,5");
    plan_lanes(dims[d], 12);
    if (dims[d].gpw < 1)
      return set_error(c, PW_ERR_ARG, "pw_set_grid: size " + std::to_string(ns[d]) + " too long for one wavefront");
  }
  // a radix mode chosen by the tuner applies to the new grid where it has plans
  if (c->fft_order != 0) {
    FftDim alt[3];
    if (plan_grid(ns, c->fft_order, alt))
      for (int d = 0; d < 3; ++d) dims[d] = alt[d];
    else
      c->fft_order = 0;
  }
  for (auto& k : c->kd) free_k(k);
  c->kd.clear();
  free_k(c->krho);
  c->krho_g0 = -1;
  for (auto p : c->d_v64) hipFree(p);
  for (auto p : c->d_v32) hipFree(p);
  c->d_v64.clear();
  c->d_v32.clear();
  c->nspin = 0;
  free_scf(c);
  c->n1 = n1;
  c->n2 = n2;
  c->n3 = n3;
  c->N = (long)n1 * n2 * n3;
  c->omega = omega;
  for (int d = 0; d < 3; ++d) {
    c->dim[d] = dims[d];
    const int n = ns[d];
    std::vector<cplx<double>> t64(n);
    std::vector<cplx<float>> t32(n);
    for (int t = 0; t < n; ++t) {
      const long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)t / (long double)n;
      t64[t] = mk((double)cosl(a), (double)sinl(a));
      t32[t] = mk((float)cosl(a), (float)sinl(a));
    }
    int rc = upload(c, reinterpret_cast<cplx<double>**>(&c->d_tw64[d]), t64.data(), n);
    if (rc) return rc;
    rc = upload(c, reinterpret_cast<cplx<float>**>(&c->d_tw32[d]), t32.data(), n);
    if (rc) return rc;
  }
  hipFree(c->d_rho);
  c->d_rho = nullptr;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&c->d_rho), c->N * sizeof(double));
  if (e != hipSuccess) return set_error(c, PW_ERR_NOMEM, "hipMalloc rho");
  e = hipMemset(c->d_rho, 0, c->N * sizeof(double));
  return hip_check(c, e, "memset rho");
}

}  // extern "C"

namespace pw {

int build_kdata(pw_ctx* c, KData& k, int ngw, const int64_t* idx, int base, const double* gk, const char* who) {
  free_k(k);
  const int n1 = c->n1, n2 = c->n2, n3 = c->n3;
  struct Ent { int col; int z; int igw; };
  std::map<int, int> colmap;  // plane offset -> column id (ordered by offset)
  std::vector<int> keys(ngw), zs(ngw);
  std::vector<char> seen(c->N, 0);
  for (int g = 0; g < ngw; ++g) {
    const int64_t ip = idx[g] - base;
    if (ip < 0 || ip >= c->N) return set_error(c, PW_ERR_ARG, std::string(who) + ": index out of grid");
    if (seen[ip]) return set_error(c, PW_ERR_ARG, std::string(who) + ": duplicate grid index");
    seen[ip] = 1;
    const int i = (int)(ip % n1), j = (int)((ip / n1) % n2), z = (int)(ip / ((int64_t)n1 * n2));
    keys[g] = i + n1 * j;
    zs[g] = z;
    colmap[keys[g]] = 0;
  }
  int ncol = 0;
  std::vector<int> col_ij;
  std::vector<int> xs;
  for (auto& kv : colmap) {
    kv.second = ncol++;
    col_ij.push_back(kv.first);
    xs.push_back(kv.first % n1);
  }
  std::vector<Ent> ent(ngw);
  for (int g = 0; g < ngw; ++g) ent[g] = Ent{colmap[keys[g]], zs[g], g};
  std::sort(ent.begin(), ent.end(), [](const Ent& a, const Ent& b) {
    return a.col != b.col ? a.col < b.col : a.z < b.z;
  });
  std::vector<int> e_igw(ngw), e_pos(ngw), col_ptr(ncol + 1, 0);
  const int cs = odd_stride(n3), ps = odd_stride(n1);
  for (int e = 0; e < ngw; ++e) {
    e_igw[e] = ent[e].igw;
    e_pos[e] = ent[e].col * cs + ent[e].z;
    col_ptr[ent[e].col + 1]++;
  }
  for (int cc = 0; cc < ncol; ++cc) col_ij[cc] = (col_ij[cc] % n1) + ps * (col_ij[cc] / n1);
  for (int cc = 0; cc < ncol; ++cc) col_ptr[cc + 1] += col_ptr[cc];
  std::sort(xs.begin(), xs.end());
  xs.erase(std::unique(xs.begin(), xs.end()), xs.end());
  std::vector<int> col_cy(ncol);
  {
    std::vector<int> apos(n1, -1);
    for (int a = 0; a < (int)xs.size(); ++a) apos[xs[a]] = a;
    const int ys = odd_stride(n2);
    int cc = 0;
    for (auto& kv : colmap) {
      const int x = kv.first % n1, y = kv.first / n1;
      col_cy[cc++] = apos[x] * ys + y;
    }
  }
  std::vector<double> gk2(ngw);
  for (int g = 0; g < ngw; ++g) {
    const double x = gk[3 * g], y = gk[3 * g + 1], z = gk[3 * g + 2];
    gk2[g] = x * x + y * y + z * z;
  }
  // slot of each psi row in a whole padded grid (x stride 1, y stride odd_stride(n1), z stride
  // odd_stride(n1) n2): the layout of the fused single-workgroup op_H
  std::vector<int> gpos(ngw);
  for (int g = 0; g < ngw; ++g) {
    const int64_t ip = idx[g] - base;
    const int i = (int)(ip % n1), j = (int)((ip / n1) % n2), z = (int)(ip / ((int64_t)n1 * n2));
    gpos[g] = i + ps * (j + n2 * z);
  }
  k.ngw = ngw;
  k.ncol = ncol;
  k.nxact = (int)xs.size();
  k.h_xact = xs;
  int rc;
  if ((rc = upload(c, &k.d_ent_igw, e_igw.data(), ngw))) return rc;
  if ((rc = upload(c, &k.d_ent_pos, e_pos.data(), ngw))) return rc;
  if ((rc = upload(c, &k.d_col_ptr, col_ptr.data(), ncol + 1))) return rc;
  if ((rc = upload(c, &k.d_col_ij, col_ij.data(), ncol))) return rc;
  if ((rc = upload(c, &k.d_col_cy, col_cy.data(), ncol))) return rc;
  if ((rc = upload(c, &k.d_xact, xs.data(), xs.size()))) return rc;
  if ((rc = upload(c, &k.d_gk2, gk2.data(), ngw))) return rc;
  if ((rc = upload(c, &k.d_gk, gk, 3 * (size_t)ngw))) return rc;
  if ((rc = upload(c, &k.d_gpos, gpos.data(), ngw))) return rc;
  return PW_OK;
}

template int upload<double>(pw_ctx*, double**, const double*, size_t);
template int upload<int>(pw_ctx*, int**, const int*, size_t);

}  // namespace pw

extern "C" {

int pw_set_kpoint(pw_ctx* c, int ik, int ngw, const int64_t* idx, int base, const double* gk) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (c->N == 0) return set_error(c, PW_ERR_STATE, "pw_set_grid must precede pw_set_kpoint");
  if (ik < 0 || ngw <= 0 || !idx || !gk) return set_error(c, PW_ERR_ARG, "pw_set_kpoint: bad arguments");
  if ((int)c->kd.size() <= ik) c->kd.resize(ik + 1);
  return build_kdata(c, c->kd[ik], ngw, idx, base, gk, "pw_set_kpoint");
}

int pw_set_vloc(pw_ctx* c, int ispin, const double* V) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (c->N == 0) return set_error(c, PW_ERR_STATE, "pw_set_grid must precede pw_set_vloc");
  if (ispin < 0 || ispin > 3 || !V) return set_error(c, PW_ERR_ARG, "pw_set_vloc: bad arguments");
  if (c->nspin <= ispin) {
    c->d_v64.resize(ispin + 1, nullptr);
    c->d_v32.resize(ispin + 1, nullptr);
    c->nspin = ispin + 1;
  }
  std::vector<double> v64(c->N);
  std::vector<float> v32(c->N);
  const double s = 1.0 / (double)c->N;
  for (long i = 0; i < c->N; ++i) {
    v64[i] = V[i] * s;
    v32[i] = (float)v64[i];
  }
  int rc = upload(c, &c->d_v64[ispin], v64.data(), c->N);
  if (rc) return rc;
  return upload(c, &c->d_v32[ispin], v32.data(), c->N);
}

int pw_set_dij(pw_ctx* c, int nbeta, const double* D) {
  CTX_CHECK(c);
  if (nbeta < 0 || (nbeta > 0 && !D)) return set_error(c, PW_ERR_ARG, "pw_set_dij: bad arguments");
  c->nbeta = nbeta;
  c->h_D.assign(D, D + (size_t)nbeta * nbeta);
  c->dcsr_valid = false;
  return PW_OK;
}

int pw_set_projectors(pw_ctx* c, int ik, int nbeta, const void* betaNL, int64_t ld) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  int rc = check_k(c, ik);
  if (rc) return rc;
  KData& k = c->kd[ik];
  if (nbeta != c->nbeta) return set_error(c, PW_ERR_ARG, "pw_set_projectors: nbeta differs from pw_set_dij");
  if (nbeta > 0 && (!betaNL || ld < k.ngw)) return set_error(c, PW_ERR_ARG, "pw_set_projectors: bad arguments");
  k.nbeta = nbeta;
  if (nbeta == 0) return PW_OK;
  const cplx<double>* b = reinterpret_cast<const cplx<double>*>(betaNL);
  const size_t n = (size_t)k.ngw * nbeta;
  std::vector<cplx<double>> beta(n), bD(n);
  std::vector<cplx<float>> b32(n), bD32(n);
  for (int j = 0; j < nbeta; ++j)
    for (int g = 0; g < k.ngw; ++g) beta[(size_t)j * k.ngw + g] = b[(size_t)j * ld + g];
  // beta * D, summing only the non-zero entries of each column of D
  for (int j = 0; j < nbeta; ++j) {
    cplx<double>* out = &bD[(size_t)j * k.ngw];
    for (int g = 0; g < k.ngw; ++g) out[g] = mk(0.0, 0.0);
    for (int i = 0; i < nbeta; ++i) {
      const double d = c->h_D[(size_t)j * nbeta + i];
      if (d == 0.0) continue;
      const cplx<double>* bi = &beta[(size_t)i * k.ngw];
      for (int g = 0; g < k.ngw; ++g) out[g] = out[g] + scal(d, bi[g]);
    }
  }
  for (size_t t = 0; t < n; ++t) {
    b32[t] = cvt<float>(beta[t]);
    bD32[t] = cvt<float>(bD[t]);
  }
  if ((rc = upload(c, reinterpret_cast<cplx<double>**>(&k.d_beta64), beta.data(), n))) return rc;
  if ((rc = upload(c, reinterpret_cast<cplx<double>**>(&k.d_betaD64), bD.data(), n))) return rc;
  if ((rc = upload(c, reinterpret_cast<cplx<float>**>(&k.d_beta32), b32.data(), n))) return rc;
  if ((rc = upload(c, reinterpret_cast<cplx<float>**>(&k.d_betaD32), bD32.data(), n))) return rc;
  return PW_OK;
}

int pw_op_H_dev(pw_ctx* c, int ik, int ispin, int nb, const void* psi, int64_t ldpsi, void* Hpsi, int64_t ldh,
                unsigned flags) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  return op_H_device(c, ik, ispin, nb, reinterpret_cast<const cplx<double>*>(psi), ldpsi,
                     reinterpret_cast<cplx<double>*>(Hpsi), ldh, flags);
}

int pw_op_H(pw_ctx* c, int ik, int ispin, int nb, const void* psi, int64_t ldpsi, void* Hpsi, int64_t ldh,
            unsigned flags) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (nb <= 0) return nb == 0 ? PW_OK : set_error(c, PW_ERR_ARG, "op_H: nb < 0");
  if (!psi || !Hpsi) return set_error(c, PW_ERR_ARG, "op_H: null pointer");
  const size_t bp = (size_t)ldpsi * nb * 16, bh = (size_t)ldh * nb * 16;
  int rc = stage(c, bp + bh);
  if (rc) return rc;
  char* dpsi = reinterpret_cast<char*>(c->d_tmp);
  char* dh = dpsi + bp;
  hipError_t e = hipMemcpyAsync(dpsi, psi, bp, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess && (flags & PW_ACCUMULATE)) e = hipMemcpyAsync(dh, Hpsi, bh, hipMemcpyHostToDevice, c->stream);
  if (e != hipSuccess) return hip_check(c, e, "op_H upload");
  rc = op_H_device(c, ik, ispin, nb, reinterpret_cast<const cplx<double>*>(dpsi), ldpsi,
                   reinterpret_cast<cplx<double>*>(dh), ldh, flags);
  if (rc) return rc;
  e = hipMemcpyAsync(Hpsi, dh, bh, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  return hip_check(c, e, "op_H download");
}

int pw_kprec_dev(pw_ctx* c, int ik, int nb, const void* in, int64_t ldin, void* out, int64_t ldout) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  int rc = check_k(c, ik);
  if (rc) return rc;
  return kprec_apply(c, ik, nb, reinterpret_cast<const cplx<double>*>(in), ldin,
                     reinterpret_cast<cplx<double>*>(out), ldout);
}

int pw_betaNL_psi_dev(pw_ctx* c, int ik, int nb, const void* psi, int64_t ldpsi, void* out, int64_t ldout,
                      unsigned flags) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  int rc = check_k(c, ik);
  if (rc) return rc;
  KData& k = c->kd[ik];
  if (k.nbeta == 0) return PW_OK;
  const cplx<double> one = mk(1.0, 0.0), zero = mk(0.0, 0.0);
  if (flags & PW_MIXED)
    return zgemm_f32in(c, 1, k.nbeta, nb, k.ngw, one, reinterpret_cast<const cplx<float>*>(k.d_beta32), k.ngw,
                       reinterpret_cast<const cplx<double>*>(psi), ldpsi, zero,
                       reinterpret_cast<cplx<double>*>(out), ldout, 1);
  return zgemm(c, 1, k.nbeta, nb, k.ngw, one, reinterpret_cast<const cplx<double>*>(k.d_beta64), k.ngw,
               reinterpret_cast<const cplx<double>*>(psi), ldpsi, zero, reinterpret_cast<cplx<double>*>(out),
               ldout, 0);
}

int pw_rhoe_zero(pw_ctx* c) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (!c->d_rho) return set_error(c, PW_ERR_STATE, "pw_set_grid has not been called");
  return zero_fill(c, c->d_rho, c->N);
}

int pw_rhoe_accum_dev(pw_ctx* c, int ik, int nb, const void* psi, int64_t ldpsi, const double* w,
                      unsigned flags) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  int rc = check_k(c, ik);
  if (rc) return rc;
  if (nb <= 0) return nb == 0 ? PW_OK : set_error(c, PW_ERR_ARG, "rhoe: nb < 0");
  rc = ensure(c, &c->d_wts, &c->wts_bytes, nb * sizeof(double));
  if (rc) return rc;
  double* dw = reinterpret_cast<double*>(c->d_wts);
  hipError_t e = hipMemcpyAsync(dw, w, nb * sizeof(double), hipMemcpyHostToDevice, c->stream);
  if (e != hipSuccess) return hip_check(c, e, "rhoe weights upload");
  rc = rhoe_accum(c, ik, nb, reinterpret_cast<const cplx<double>*>(psi), ldpsi, dw, flags);
  if (rc) return rc;
  return hip_check(c, hipStreamSynchronize(c->stream), "rhoe");
}

int pw_rhoe_accum(pw_ctx* c, int ik, int nb, const void* psi, int64_t ldpsi, const double* w, unsigned flags) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (nb <= 0) return nb == 0 ? PW_OK : set_error(c, PW_ERR_ARG, "rhoe: nb < 0");
  const size_t bp = (size_t)ldpsi * nb * 16;
  int rc = stage(c, bp);
  if (rc) return rc;
  hipError_t e = hipMemcpyAsync(c->d_tmp, psi, bp, hipMemcpyHostToDevice, c->stream);
  if (e != hipSuccess) return hip_check(c, e, "rhoe upload");
  return pw_rhoe_accum_dev(c, ik, nb, c->d_tmp, ldpsi, w, flags);
}

int pw_rhoe_get(pw_ctx* c, double* rho) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (!c->d_rho || !rho) return set_error(c, PW_ERR_STATE, "rho not available");
  hipError_t e = hipMemcpyAsync(rho, c->d_rho, c->N * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  return hip_check(c, e, "rho download");
}

int pw_zgemm_dev(pw_ctx* c, int transa, int m, int n, int k, const double* alpha, const void* A, int64_t lda,
                 const void* B, int64_t ldb, const double* beta, void* C, int64_t ldc, unsigned flags) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (m < 0 || n < 0 || k < 0 || !alpha || !beta) return set_error(c, PW_ERR_ARG, "zgemm: bad arguments");
  if (ldc < m || ldb < k || lda < (transa ? k : m)) return set_error(c, PW_ERR_ARG, "zgemm: bad leading dimension");
  return zgemm(c, transa, m, n, k, mk(alpha[0], alpha[1]), reinterpret_cast<const cplx<double>*>(A), lda,
               reinterpret_cast<const cplx<double>*>(B), ldb, mk(beta[0], beta[1]),
               reinterpret_cast<cplx<double>*>(C), ldc, (flags & PW_MIXED) ? 1 : 0);
}

int pw_malloc(pw_ctx* c, size_t bytes, void** dptr) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (!dptr) return set_error(c, PW_ERR_ARG, "pw_malloc: null");
  hipError_t e = hipMalloc(dptr, std::max<size_t>(bytes, 16));
  if (e != hipSuccess) return set_error(c, PW_ERR_NOMEM, std::string("pw_malloc: ") + hipGetErrorString(e));
  return PW_OK;
}

int pw_free(pw_ctx* c, void* p) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  hipStreamSynchronize(c->stream);
  return hip_check(c, hipFree(p), "pw_free");
}

int pw_memcpy_h2d(pw_ctx* c, void* dst, const void* src, size_t bytes) {
  CTX_CHECK(c);
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  return hip_check(c, e, "h2d");
}

int pw_memcpy_d2h(pw_ctx* c, void* dst, const void* src, size_t bytes) {
  CTX_CHECK(c);
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  return hip_check(c, e, "d2h");
}

int pw_memcpy_d2d(pw_ctx* c, void* dst, const void* src, size_t bytes) {
  CTX_CHECK(c);
  return hip_check(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream), "d2d");
}

int pw_synchronize(pw_ctx* c) {
  CTX_CHECK(c);
  if (c->capturing) return set_error(c, PW_ERR_STATE, "pw_synchronize during graph capture");
  return hip_check(c, hipStreamSynchronize(c->stream), "synchronize");
}

int pw_graph_begin(pw_ctx* c) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (c->capturing) return set_error(c, PW_ERR_STATE, "pw_graph_begin: already capturing");
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return hip_check(c, e, "hipStreamBeginCapture");
  c->capturing = true;
  return PW_OK;
}

int pw_graph_end(pw_ctx* c, int* id) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (!c->capturing) return set_error(c, PW_ERR_STATE, "pw_graph_end: not capturing");
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  c->capturing = false;
  if (e != hipSuccess) return hip_check(c, e, "hipStreamEndCapture");
  hipGraphExec_t x = nullptr;
  e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    hipGraphDestroy(g);
    return hip_check(c, e, "hipGraphInstantiate");
  }
  c->graphs.push_back(reinterpret_cast<void*>(x));
  c->graph_defs.push_back(reinterpret_cast<void*>(g));
  c->graphs_live++;
  if (id) *id = (int)c->graphs.size() - 1;
  return PW_OK;
}

int pw_graph_launch(pw_ctx* c, int id) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (id < 0 || id >= (int)c->graphs.size() || !c->graphs[id]) return set_error(c, PW_ERR_ARG, "pw_graph_launch: bad id");
  return hip_check(c, hipGraphLaunch(reinterpret_cast<hipGraphExec_t>(c->graphs[id]), c->stream), "hipGraphLaunch");
}

int pw_graph_free(pw_ctx* c, int id) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (id < 0 || id >= (int)c->graphs.size() || !c->graphs[id]) return set_error(c, PW_ERR_ARG, "pw_graph_free: bad id");
  hipStreamSynchronize(c->stream);
  hipGraphExecDestroy(reinterpret_cast<hipGraphExec_t>(c->graphs[id]));
  hipGraphDestroy(reinterpret_cast<hipGraph_t>(c->graph_defs[id]));
  c->graphs[id] = nullptr;
  c->graph_defs[id] = nullptr;
  if (--c->graphs_live == 0) {
    for (void* p : c->retired) hipFree(p);
    c->retired.clear();
  }
  return PW_OK;
}

int pw_copy_cols_dev(pw_ctx* c, int nrows, int ncols, const void* src, int64_t lds, void* dst, int64_t ldd) {
  CTX_CHECK(c);
  return copy_cols(c, nrows, ncols, reinterpret_cast<const cplx<double>*>(src), lds, reinterpret_cast<cplx<double>*>(dst),
                   ldd);
}

int pw_profile_enable(pw_ctx* c, int on) {
  CTX_CHECK(c);
  drain_events(c);
  c->profile = on != 0;
  if (on) {
    for (int i = 0; i < PW_PROF_SLOTS; ++i) {
      c->prof_ms[i] = 0.0;
      c->prof_count[i] = 0;
      c->prof_bytes[i] = 0.0;
    }
  }
  return PW_OK;
}

int pw_profile_read(pw_ctx* c, int slot, double* total_ms, long* count, double* bytes) {
  CTX_CHECK(c);
  if (slot < 0 || slot >= PW_PROF_SLOTS || !total_ms || !count || !bytes)
    return set_error(c, PW_ERR_ARG, "pw_profile_read: bad arguments");
  drain_events(c);
  *total_ms = c->prof_ms[slot];
  *count = c->prof_count[slot];
  *bytes = c->prof_bytes[slot];
  return PW_OK;
}

static hipEvent_t g_events[16];
static bool g_events_init = false;

int pw_event_record(pw_ctx* c, int slot) {
  CTX_CHECK(c);
  if (slot < 0 || slot >= 16) return set_error(c, PW_ERR_ARG, "event slot");
  if (!g_events_init) {
    for (auto& ev : g_events) hipEventCreate(&ev);
    g_events_init = true;
  }
  return hip_check(c, hipEventRecord(g_events[slot], c->stream), "event record");
}

int pw_event_elapsed_ms(pw_ctx* c, int s0, int s1, float* ms) {
  CTX_CHECK(c);
  if (!g_events_init || s0 < 0 || s1 < 0 || s0 >= 16 || s1 >= 16 || !ms) return set_error(c, PW_ERR_ARG, "event");
  hipError_t e = hipEventSynchronize(g_events[s1]);
  if (e == hipSuccess) e = hipEventElapsedTime(ms, g_events[s0], g_events[s1]);
  return hip_check(c, e, "event elapsed");
}

}  // extern "C"
