// Batched entry points: many k-points (each with its own plane-wave count) in one launch
// per stage.  The Hamiltonian application runs the three sphere-pruned FFT kernels over all
// k-points together (pw_fft.hip, *_batch kernels) and the nonlocal products through a
// batched complex GEMM for small, variable-size products; the eigensolver's Gram products
// and column updates use the same GEMM.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "pw_internal.h"

namespace pw {

constexpr int BT = 16;   // output tile BT x BT per workgroup
constexpr int BK = 32;   // k-chunk staged in LDS

// C = alpha op(A) B + beta C for every product of the batch; one workgroup of 256 threads
// per 16 x 16 output tile, thread (tx, ty) owns C[i0 + tx, j0 + ty].  A and B chunks are
// staged through LDS with loads that run along the contiguous (column) direction.
template <int TA>
__global__ void __launch_bounds__(256) k_zgemm_batch(const GBatch* __restrict__ gb, int ng, cplx<double> alpha,
                                                     cplx<double> beta, int use_beta) {
  __shared__ cplx<double> As[BK][BT + 1];
  __shared__ cplx<double> Bs[BK][BT + 1];
  const int w = blockIdx.x;
  int lo = 0, hi = ng - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (gb[mid].tile0 <= w) lo = mid; else hi = mid - 1;
  }
  const GBatch& g = gb[lo];
  const int tm = (g.m + BT - 1) / BT;
  const int t = w - g.tile0;
  const int i0 = (t % tm) * BT, j0 = (t / tm) * BT;
  const int tid = threadIdx.x, tx = tid % BT, ty = tid / BT;
  cplx<double> acc = mk(0.0, 0.0);
  // the next chunk's elements are loaded into registers while the current one is multiplied
  constexpr int EPT = BK * BT / 256;
  cplx<double> ra[EPT], rb[EPT];
  auto load = [&](int k0) {
#pragma unroll
    for (int s = 0; s < EPT; ++s) {
      const int e = tid + 256 * s;
      int kc, i;
      if (TA) {
        kc = e % BK;
        i = e / BK;
      } else {
        i = e % BT;
        kc = e / BT;
      }
      const int kk = k0 + kc, ii = i0 + i;
      ra[s] = mk(0.0, 0.0);
      if (kk < g.k && ii < g.m) ra[s] = TA ? g.A[(long)ii * g.lda + kk] : g.A[(long)kk * g.lda + ii];
      const int kb = e % BK, jj = j0 + e / BK;
      rb[s] = mk(0.0, 0.0);
      if (k0 + kb < g.k && jj < g.n) rb[s] = g.B[(long)jj * g.ldb + k0 + kb];
    }
  };
  if (g.k > 0) load(0);
  for (int k0 = 0; k0 < g.k; k0 += BK) {
#pragma unroll
    for (int s = 0; s < EPT; ++s) {
      const int e = tid + 256 * s;
      if (TA) As[e % BK][e / BK] = ra[s];
      else As[e / BT][e % BT] = ra[s];
      Bs[e % BK][e / BK] = rb[s];
    }
    __syncthreads();
    if (k0 + BK < g.k) load(k0 + BK);
#pragma unroll 8
    for (int kc = 0; kc < BK; ++kc) {
      const cplx<double> a = As[kc][tx], b = Bs[kc][ty];
      if (TA) {  // conj(a) * b
        acc.x += a.x * b.x + a.y * b.y;
        acc.y += a.x * b.y - a.y * b.x;
      } else {
        acc.x += a.x * b.x - a.y * b.y;
        acc.y += a.x * b.y + a.y * b.x;
      }
    }
    __syncthreads();
  }
  const int i = i0 + tx, j = j0 + ty;
  if (i < g.m && j < g.n) {
    cplx<double>* cp = g.C + (long)j * g.ldc + i;
    cplx<double> r = cmul(alpha, acc);
    if (use_beta) r = r + cmul(beta, *cp);
    *cp = r;
  }
}

// C = alpha A B + beta C for products with a short inner dimension (k <= SK_KMAX) and few
// columns (n <= SK_NMAX): one thread per row of C, 256 rows per workgroup; B (k x n) is
// staged in LDS and read as a broadcast, the row's k entries of A are loaded together (one
// memory latency per workgroup instead of one per k-chunk).  Workgroup w takes row block
// w % bpp of product w / bpp (bpp: row blocks of the tallest product); blocks past a
// product's rows return at once.
constexpr int SK_KMAX = 64, SK_NMAX = 32, SK_NC = 8;
__global__ void __launch_bounds__(256) k_zgemm_batch_nn_rows(const GBatch* __restrict__ gb, int bpp,
                                                             cplx<double> alpha, cplx<double> beta, int use_beta) {
  __shared__ cplx<double> Bs[SK_KMAX * SK_NMAX];
  const GBatch& g = gb[blockIdx.x / bpp];
  const int rb = blockIdx.x % bpp;
  if (rb * 256 >= g.m) return;
  const int tid = threadIdx.x;
  for (int e = tid; e < g.k * g.n; e += 256) {
    const int kk = e % g.k, j = e / g.k;
    Bs[kk * SK_NMAX + j] = g.B[(long)j * g.ldb + kk];
  }
  __syncthreads();
  const int i = rb * 256 + tid;
  if (i >= g.m) return;
  for (int j0 = 0; j0 < g.n; j0 += SK_NC) {
    cplx<double> acc[SK_NC];
#pragma unroll
    for (int jj = 0; jj < SK_NC; ++jj) acc[jj] = mk(0.0, 0.0);
#pragma unroll 8
    for (int kk = 0; kk < g.k; ++kk) {
      const cplx<double> a = g.A[(long)kk * g.lda + i];
#pragma unroll
      for (int jj = 0; jj < SK_NC; ++jj) {
        const cplx<double> b = Bs[kk * SK_NMAX + j0 + jj];   // columns past n hold stale values, never stored
        acc[jj].x += a.x * b.x - a.y * b.y;
        acc[jj].y += a.x * b.y + a.y * b.x;
      }
    }
#pragma unroll
    for (int jj = 0; jj < SK_NC; ++jj) {
      const int j = j0 + jj;
      if (j < g.n) {
        cplx<double>* cp = g.C + (long)j * g.ldc + i;
        cplx<double> r = cmul(alpha, acc[jj]);
        if (use_beta) r = r + cmul(beta, *cp);
        *cp = r;
      }
    }
  }
}

// C = alpha A^H B + beta C for small products (m, n <= SC_MAX) with a long inner dimension:
// one workgroup per product, every output of C owned by one thread (up to SC_OPT each), the
// k-chunks of A and B staged through LDS with the next chunk loaded into registers while
// the current one is multiplied.
// k-chunks 64 deep (66 KB of LDS): at the block eigensolvers' few hundred rows this halves
// the chunk round trips of 32-deep chunks; 48 and 128 measured no better and slower
constexpr int SC_MAX = 32, SC_KC = 64, SC_OPT = SC_MAX * SC_MAX / 256;
__global__ void __launch_bounds__(256) k_zgemm_batch_cn_small(const GBatch* __restrict__ gb, cplx<double> alpha,
                                                              cplx<double> beta, int use_beta) {
  __shared__ cplx<double> As[SC_KC][SC_MAX + 1];
  __shared__ cplx<double> Bs[SC_KC][SC_MAX + 1];
  const GBatch& g = gb[blockIdx.x];
  const int tid = threadIdx.x;
  const int m = g.m, n = g.n;
  cplx<double> acc[SC_OPT];
#pragma unroll
  for (int o = 0; o < SC_OPT; ++o) acc[o] = mk(0.0, 0.0);
  constexpr int EPT = SC_KC * SC_MAX / 256;   // staged elements per thread and operand
  cplx<double> ra[EPT], rb[EPT];
  // element e of a chunk: k offset e % SC_KC (contiguous in memory), column e / SC_KC
  auto load = [&](int k0) {
#pragma unroll
    for (int s = 0; s < EPT; ++s) {
      const int e = tid + 256 * s, kc = e % SC_KC, col = e / SC_KC, kk = k0 + kc;
      ra[s] = (kk < g.k && col < m) ? g.A[(long)col * g.lda + kk] : mk(0.0, 0.0);
      rb[s] = (kk < g.k && col < n) ? g.B[(long)col * g.ldb + kk] : mk(0.0, 0.0);
    }
  };
  if (g.k > 0) load(0);
  for (int k0 = 0; k0 < g.k; k0 += SC_KC) {
#pragma unroll
    for (int s = 0; s < EPT; ++s) {
      const int e = tid + 256 * s;
      As[e % SC_KC][e / SC_KC] = ra[s];
      Bs[e % SC_KC][e / SC_KC] = rb[s];
    }
    __syncthreads();
    if (k0 + SC_KC < g.k) load(k0 + SC_KC);
#pragma unroll
    for (int o = 0; o < SC_OPT; ++o) {
      const int q = tid + 256 * o;
      if (q < m * n) {
        const int i = q % m, j = q / m;
        cplx<double> x = acc[o];
#pragma unroll 8
        for (int kc = 0; kc < SC_KC; ++kc) {
          const cplx<double> a = As[kc][i], b = Bs[kc][j];   // conj(a) * b
          x.x += a.x * b.x + a.y * b.y;
          x.y += a.x * b.y - a.y * b.x;
        }
        acc[o] = x;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int o = 0; o < SC_OPT; ++o) {
    const int q = tid + 256 * o;
    if (q < m * n) {
      const int i = q % m, j = q / m;
      cplx<double>* cp = g.C + (long)j * g.ldc + i;
      cplx<double> r = cmul(alpha, acc[o]);
      if (use_beta) r = r + cmul(beta, *cp);
      *cp = r;
    }
  }
}

int desc_get(pw_ctx* c, const void* data, size_t bytes, const void** out) {
  const int n = (int)c->desc_host.size();
  for (int i = 0; i < n; ++i)
    if (c->desc_host[i].size() == bytes && std::memcmp(c->desc_host[i].data(), data, bytes) == 0) {
      *out = c->desc_dev[i];
      return PW_OK;
    }
  if (c->capturing)
    return set_error(c, PW_ERR_STATE, "a new descriptor array during graph capture; run the sequence once first");
  constexpr int kSlots = 256;
  int slot;
  if (n < kSlots || c->graphs_live > 0) {  // never overwrite an array a captured graph may read
    // a new slot: no queued kernel reads it, so nothing to wait for
    c->desc_host.emplace_back();
    c->desc_dev.push_back(nullptr);
    c->desc_cap.push_back(0);
    slot = n;
  } else {
    // reusing a slot: kernels still queued may read its old contents (and its host copy
    // may still be the source of a queued upload)
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_check(c, e, "descriptor sync");
    slot = c->desc_next;
    c->desc_next = (c->desc_next + 1) % kSlots;
  }
  // allocated directly, not through ensure(): a new slot leaves every captured graph valid
  // (no slot a live graph reads is ever freed or rewritten)
  if (c->desc_cap[slot] < bytes) {
    hipFree(c->desc_dev[slot]);
    c->desc_dev[slot] = nullptr;
    c->desc_cap[slot] = 0;
    const size_t sz = std::max<size_t>(bytes, 256);
    hipError_t e = hipMalloc(&c->desc_dev[slot], sz);
    if (e != hipSuccess) {
      c->desc_dev[slot] = nullptr;
      return set_error(c, PW_ERR_NOMEM, std::string("hipMalloc descriptor: ") + hipGetErrorString(e));
    }
    c->desc_cap[slot] = sz;
  }
  const char* p = reinterpret_cast<const char*>(data);
  c->desc_host[slot].assign(p, p + bytes);
  // ordered on the stream before the kernels that read it; the host copy it reads from is
  // not changed before the slot is reused, which waits for the stream first
  hipError_t e = hipMemcpyAsync(c->desc_dev[slot], c->desc_host[slot].data(), bytes, hipMemcpyHostToDevice,
                                c->stream);
  if (e != hipSuccess) return hip_check(c, e, "descriptor upload");
  *out = c->desc_dev[slot];
  return PW_OK;
}

void free_desc(pw_ctx* c) {
  for (void* p : c->desc_dev) hipFree(p);
  c->desc_dev.clear();
  c->desc_host.clear();
  c->desc_cap.clear();
  c->desc_next = 0;
  hipFree(c->d_bnlb);
  c->d_bnlb = nullptr;
  c->bnlb_bytes = 0;
}

int zgemm_batch(pw_ctx* c, int transa, const std::vector<GBatch>& g0, cplx<double> alpha, cplx<double> beta) {
  std::vector<GBatch> g;
  g.reserve(g0.size());
  int maxm = 0, maxn = 0, maxk = 0;
  for (const GBatch& x : g0) {
    if (x.m <= 0 || x.n <= 0) continue;
    g.push_back(x);
    maxm = std::max(maxm, x.m);
    maxn = std::max(maxn, x.n);
    maxk = std::max(maxk, x.k);
  }
  if (g.empty()) return PW_OK;
  const int ub = (beta.x != 0.0 || beta.y != 0.0) ? 1 : 0;
  // the shapes of the block eigensolvers' products: A B with a short inner dimension and few
  // columns (the Ritz-vector updates), A^H B with a small result (the Gram matrices)
  const int kind = (!transa && maxk <= SK_KMAX && maxn <= SK_NMAX && c->batch_gemm_small) ? 1
                   : (transa && maxm <= SC_MAX && maxn <= SC_MAX && c->batch_gemm_small) ? 2 : 0;
  long tiles = 0;
  const int bpp = (maxm + 255) / 256;   // row blocks per product (kind 1)
  for (GBatch& y : g) {
    y.tile0 = (int)tiles;
    if (kind == 1) tiles += bpp;
    else if (kind == 2) tiles += 1;
    else tiles += (long)((y.m + BT - 1) / BT) * ((y.n + BT - 1) / BT);
  }
  if (tiles > 0x7fffffffL) return set_error(c, PW_ERR_ARG, "batched GEMM too large for one launch");
  const void* dptr = nullptr;
  int rc = desc_get(c, g.data(), g.size() * sizeof(GBatch), &dptr);
  if (rc) return rc;
  const GBatch* d = reinterpret_cast<const GBatch*>(dptr);
  if (kind == 1) {
    hipLaunchKernelGGL(k_zgemm_batch_nn_rows, dim3((unsigned)tiles), dim3(256), 0, c->stream, d, bpp, alpha, beta,
                       ub);
    return hip_check(c, hipGetLastError(), "k_zgemm_batch_nn_rows");
  }
  if (kind == 2) {
    hipLaunchKernelGGL(k_zgemm_batch_cn_small, dim3((unsigned)tiles), dim3(256), 0, c->stream, d, alpha, beta, ub);
    return hip_check(c, hipGetLastError(), "k_zgemm_batch_cn_small");
  }
  if (transa)
    hipLaunchKernelGGL(k_zgemm_batch<1>, dim3((unsigned)tiles), dim3(256), 0, c->stream, d, (int)g.size(), alpha,
                       beta, ub);
  else
    hipLaunchKernelGGL(k_zgemm_batch<0>, dim3((unsigned)tiles), dim3(256), 0, c->stream, d, (int)g.size(), alpha,
                       beta, ub);
  return hip_check(c, hipGetLastError(), "k_zgemm_batch");
}

__global__ void k_zero_cols_b(cplx<double>* __restrict__ a, long ld, int nrows, int ncols) {
  const long tot = (long)nrows * ncols;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < tot; t += (long)gridDim.x * blockDim.x)
    a[(t / nrows) * ld + (t % nrows)] = mk(0.0, 0.0);
}

constexpr int EPB = 1024;  // elements per workgroup of k_elem_batch

__global__ void __launch_bounds__(256) k_elem_batch(const EBatch* __restrict__ eb, int ne) {
  const int w = blockIdx.x;
  int lo = 0, hi = ne - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (eb[mid].blk0 <= w) lo = mid; else hi = mid - 1;
  }
  const EBatch& d = eb[lo];
  const long tot = (long)d.m * d.n;
  const long base = (long)(w - d.blk0) * EPB;
  for (long t = base + threadIdx.x; t < min(base + EPB, tot); t += 256) {
    const int i = (int)(t % d.m), j = (int)(t / d.m);
    cplx<double> v = d.in[(long)j * d.ldin + i];
    if (d.gk2) v = scal(1.0 / (1.0 + d.gk2[i]), v);
    d.out[(long)j * d.ldout + i] = v;
  }
}

int elem_batch(pw_ctx* c, std::vector<EBatch>& e0) {
  std::vector<EBatch> e;
  long blk = 0;
  for (EBatch x : e0) {
    if (x.m <= 0 || x.n <= 0) continue;
    x.blk0 = (int)blk;
    blk += ((long)x.m * x.n + EPB - 1) / EPB;
    e.push_back(x);
  }
  if (e.empty()) return PW_OK;
  if (blk > 0x7fffffffL) return set_error(c, PW_ERR_ARG, "batched copy too large for one launch");
  const void* dptr = nullptr;
  int rc = desc_get(c, e.data(), e.size() * sizeof(EBatch), &dptr);
  if (rc) return rc;
  hipLaunchKernelGGL(k_elem_batch, dim3((unsigned)blk), dim3(256), 0, c->stream,
                     reinterpret_cast<const EBatch*>(dptr), (int)e.size());
  return hip_check(c, hipGetLastError(), "k_elem_batch");
}

}  // namespace pw

using namespace pw;

#define CTX_CHECK(c) \
  if (!(c)) return PW_ERR_ARG

static int check_batch_k(pw_ctx* c, int nk, const int* iks) {
  if (c->N == 0) return set_error(c, PW_ERR_STATE, "pw_set_grid has not been called");
  if (nk < 0 || (nk > 0 && !iks)) return set_error(c, PW_ERR_ARG, "batch: bad k-point list");
  for (int kk = 0; kk < nk; ++kk)
    if (iks[kk] < 0 || iks[kk] >= (int)c->kd.size() || c->kd[iks[kk]].ngw == 0)
      return set_error(c, PW_ERR_STATE, "batch: k-point " + std::to_string(iks[kk]) + " is not set up");
  return PW_OK;
}

extern "C" {

int pw_zgemm_batch_dev(pw_ctx* c, int transa, int nbatch, const int* m, const int* n, const int* k,
                       const double* alpha, void* const* A, const int64_t* lda, void* const* B, const int64_t* ldb,
                       const double* beta, void* const* C, const int64_t* ldc) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (nbatch < 0 || (nbatch > 0 && (!m || !n || !k || !A || !lda || !B || !ldb || !C || !ldc)) || !alpha || !beta)
    return set_error(c, PW_ERR_ARG, "pw_zgemm_batch_dev: null argument");
  std::vector<GBatch> g(nbatch);
  for (int i = 0; i < nbatch; ++i) {
    if (m[i] < 0 || n[i] < 0 || k[i] < 0 || ldc[i] < std::max(1, m[i]) ||
        ldb[i] < std::max(1, k[i]) || lda[i] < std::max(1, transa ? k[i] : m[i]))
      return set_error(c, PW_ERR_ARG, "pw_zgemm_batch_dev: bad sizes or leading dimensions");
    GBatch x{};
    x.A = reinterpret_cast<const cplx<double>*>(A[i]);
    x.B = reinterpret_cast<const cplx<double>*>(B[i]);
    x.C = reinterpret_cast<cplx<double>*>(C[i]);
    x.lda = lda[i];
    x.ldb = ldb[i];
    x.ldc = ldc[i];
    x.m = m[i];
    x.n = n[i];
    x.k = k[i];
    g[i] = x;
  }
  return zgemm_batch(c, transa ? 1 : 0, g, mk(alpha[0], alpha[1]), mk(beta[0], beta[1]));
}

int pw_op_H_batch_dev(pw_ctx* c, int nk, const int* iks, int ispin, int nb, void* const* psi, const int64_t* ldpsi,
                      void* const* Hpsi, const int64_t* ldh, unsigned flags) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  int rc = check_batch_k(c, nk, iks);
  if (rc) return rc;
  if (nb < 0) return set_error(c, PW_ERR_ARG, "op_H batch: nb < 0");
  if (nk == 0 || nb == 0) return PW_OK;
  if (!psi || !ldpsi || !Hpsi || !ldh) return set_error(c, PW_ERR_ARG, "op_H batch: null argument");
  for (int kk = 0; kk < nk; ++kk)
    if (!psi[kk] || !Hpsi[kk] || ldpsi[kk] < c->kd[iks[kk]].ngw || ldh[kk] < c->kd[iks[kk]].ngw)
      return set_error(c, PW_ERR_ARG, "op_H batch: bad pointer or leading dimension");
  const unsigned terms = flags & PW_TERM_ALL;
  bool have = (flags & PW_ACCUMULATE) != 0;
  const cplx<double>* const* P = reinterpret_cast<const cplx<double>* const*>(psi);
  cplx<double>* const* Hp = reinterpret_cast<cplx<double>* const*>(Hpsi);
  if ((terms & PW_TERM_VLOC) && c->fused_batch && op_H_fused_fits(c, (flags & PW_MIXED) != 0)) {
    // the whole grid of a band fits one workgroup's LDS: every term of every band of every
    // k-point in one launch, nothing of the FFTs in HBM
    if (ispin < 0 || ispin >= c->nspin || !c->d_v64[ispin])
      return set_error(c, PW_ERR_STATE, "op_V_loc: potential of spin " + std::to_string(ispin) + " not set");
    return op_H_fused_batch(c, nk, iks, ispin, nb, P, ldpsi, Hp, ldh, flags);
  }
  if (terms & PW_TERM_VLOC) {
    if (ispin < 0 || ispin >= c->nspin || !c->d_v64[ispin])
      return set_error(c, PW_ERR_STATE, "op_V_loc: potential of spin " + std::to_string(ispin) + " not set");
    rc = vloc_apply_batch(c, nk, iks, ispin, nb, P, ldpsi, Hp, ldh, flags);
    if (rc) return rc;
    have = true;
  } else if (terms & PW_TERM_K) {
    for (int kk = 0; kk < nk; ++kk) {
      rc = kinetic_apply(c, iks[kk], nb, P[kk], ldpsi[kk], Hp[kk], ldh[kk], have ? PW_ACCUMULATE : 0);
      if (rc) return rc;
    }
    have = true;
  }
  const int nbeta = c->kd[iks[0]].nbeta;
  bool nloc = (terms & PW_TERM_NLOC) && nbeta > 0;
  if (nloc) {
    for (int kk = 0; kk < nk; ++kk)
      if (c->kd[iks[kk]].nbeta != nbeta) return set_error(c, PW_ERR_STATE, "op_H batch: projectors not set");
    rc = ensure(c, &c->d_bnlb, &c->bnlb_bytes, (size_t)nk * nbeta * nb * sizeof(cplx<double>));
    if (rc) return rc;
    cplx<double>* bnl = reinterpret_cast<cplx<double>*>(c->d_bnlb);
    std::vector<GBatch> g1(nk), g2(nk);
    for (int kk = 0; kk < nk; ++kk) {
      const KData& k = c->kd[iks[kk]];
      GBatch a{};
      a.A = reinterpret_cast<const cplx<double>*>(k.d_beta64);
      a.lda = k.ngw;
      a.B = P[kk];
      a.ldb = ldpsi[kk];
      a.C = bnl + (size_t)kk * nbeta * nb;
      a.ldc = nbeta;
      a.m = nbeta;
      a.n = nb;
      a.k = k.ngw;
      g1[kk] = a;
      GBatch b{};
      b.A = reinterpret_cast<const cplx<double>*>(k.d_betaD64);
      b.lda = k.ngw;
      b.B = a.C;
      b.ldb = nbeta;
      b.C = Hp[kk];
      b.ldc = ldh[kk];
      b.m = k.ngw;
      b.n = nb;
      b.k = nbeta;
      g2[kk] = b;
    }
    const cplx<double> one = mk(1.0, 0.0), zero = mk(0.0, 0.0);
    rc = zgemm_batch(c, 1, g1, one, zero);
    if (rc) return rc;
    rc = zgemm_batch(c, 0, g2, one, have ? one : zero);
    if (rc) return rc;
    have = true;
  }
  if (!have) {
    for (int kk = 0; kk < nk; ++kk) {
      const long tot = (long)c->kd[iks[kk]].ngw * nb;
      hipLaunchKernelGGL(k_zero_cols_b, dim3((int)std::min<long>((tot + 255) / 256, 4096)), dim3(256), 0, c->stream,
                         Hp[kk], ldh[kk], c->kd[iks[kk]].ngw, nb);
    }
    return hip_check(c, hipGetLastError(), "op_H batch: zero");
  }
  return PW_OK;
}

int pw_kprec_batch_dev(pw_ctx* c, int nk, const int* iks, int nb, void* const* in, const int64_t* ldin,
                       void* const* out, const int64_t* ldout) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  int rc = check_batch_k(c, nk, iks);
  if (rc) return rc;
  if (nk == 0 || nb <= 0) return nb < 0 ? set_error(c, PW_ERR_ARG, "kprec batch: nb < 0") : PW_OK;
  if (!in || !ldin || !out || !ldout) return set_error(c, PW_ERR_ARG, "kprec batch: null argument");
  std::vector<EBatch> e(nk);
  for (int kk = 0; kk < nk; ++kk) {
    const KData& k = c->kd[iks[kk]];
    if (ldin[kk] < k.ngw || ldout[kk] < k.ngw) return set_error(c, PW_ERR_ARG, "kprec batch: leading dimension");
    EBatch x{};
    x.in = reinterpret_cast<const cplx<double>*>(in[kk]);
    x.out = reinterpret_cast<cplx<double>*>(out[kk]);
    x.gk2 = k.d_gk2;
    x.ldin = ldin[kk];
    x.ldout = ldout[kk];
    x.m = k.ngw;
    x.n = nb;
    e[kk] = x;
  }
  return elem_batch(c, e);
}

int pw_copy_cols_batch_dev(pw_ctx* c, int nbatch, const int* rows, const int* cols, void* const* src,
                           const int64_t* lds, void* const* dst, const int64_t* ldd) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (nbatch < 0 || (nbatch > 0 && (!rows || !cols || !src || !lds || !dst || !ldd)))
    return set_error(c, PW_ERR_ARG, "pw_copy_cols_batch_dev: null argument");
  std::vector<EBatch> e(nbatch);
  for (int i = 0; i < nbatch; ++i) {
    if (rows[i] < 0 || cols[i] < 0 || lds[i] < rows[i] || ldd[i] < rows[i])
      return set_error(c, PW_ERR_ARG, "pw_copy_cols_batch_dev: bad sizes");
    EBatch x{};
    x.in = reinterpret_cast<const cplx<double>*>(src[i]);
    x.out = reinterpret_cast<cplx<double>*>(dst[i]);
    x.ldin = lds[i];
    x.ldout = ldd[i];
    x.m = rows[i];
    x.n = cols[i];
    e[i] = x;
  }
  return elem_batch(c, e);
}

}  // extern "C"
