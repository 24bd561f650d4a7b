// Hermitian eigensolver of the Rayleigh-Ritz step (the ZHEEV of diag_LOBPCG.jl:176-203 and
// diag_davidson.jl; ortho_sqrt.jl:6-9), built for the sizes the block eigensolvers meet
// (n = Nst .. 3 Nst, a few hundred to about a thousand):
//
//  1. k_hetrd_dist / k_hetrd_dist1: Householder reduction to real symmetric tridiagonal
//     form, T = Q^H A Q, Q = H(0) H(1) ... H(n-2), H(i) = I - tau_i v_i v_i^H (LAPACK ZHETD2,
//     lower variant).  One persistent cooperative launch: the matrix lives in the LDS of P
//     workgroups (column j in workgroup j mod P), so the n - 1 Householder steps never
//     re-read it from HBM and never pay a kernel launch.  k_hetrd_dist has two cross-workgroup
//     synchronisation points per step: the owner of column i publishes v_i, and after every
//     workgroup has formed p = tau A22 v for its own columns (A22 is Hermitian, so
//     (A22 v)_j = A22[:, j]^H v needs only local columns) the p segments are gathered; the
//     owner of column i + 1 updates that column first and publishes v_{i+1} before it
//     updates the rest of its columns (look-ahead).  k_hetrd_dist1 has one: with its p
//     segments the owner of the next column publishes that column, and every workgroup
//     forms the next reflector itself.
//  2. The eigenpairs of T: for the block eigensolvers (which keep only the lowest third of
//     a Rayleigh-Ritz problem) the wanted ones from the device kernels k_tri_bisect,
//     k_tri_invit and k_tri_cluster below; otherwise every one through rocSOLVER's divide
//     and conquer (ZSTEDC).
//  3. k_larft + k_apply_q_wy: eigenvectors of A = Q (eigenvectors of T), the reflectors
//     applied by blocks of 32 in compact WY form (I - V T V^H) to the wanted columns only,
//     four columns per workgroup staged in LDS (one beyond n ~ 1500), each block's
//     reflectors staged by 64-row chunks.
//
// Only the upper triangle of A is read, as Julia's Hermitian(A) does.  Cross-workgroup
// waits are bounded: a wait that exceeds its spin limit sets a status word, every other
// wait then ends at once, and the host reports the failure.  A step costs ~7.4 us at
// n = 480 with 64 workgroups and one exchange (17 us with 26 workgroups and two): the round
// trips to device-coherent memory on its critical path bound this one-dimensional
// distribution.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <string>

#include "pw_internal.h"

namespace pw {

constexpr int HT_NT = 256;                 // threads per workgroup (4 wavefronts)
constexpr long HT_SPIN = 1L << 22;         // spin limit of one cross-workgroup wait
constexpr size_t HT_LDS = 160 * 1024 - 256;  // LDS bytes per workgroup, less room for the static shared state

__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Thread 0: wait until *ctr >= target; false once the status word is set or the spin limit
// is reached (then it sets the status word itself).  The counter is polled with relaxed
// loads (an acquire load would invalidate the caches on every poll); every thread orders its
// later reads after the barrier that follows (xacquire).
__device__ inline bool wait_geq(unsigned* ctr, unsigned target, int* status) {
  long spins = 0;
  while (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
    if (__hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return false;
    if (++spins > HT_SPIN) {
      __hip_atomic_store(status, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  return true;
}

// Values exchanged between the workgroups of the reduction (v_i, tau_i, the p segments and
// partial sums).  With LIGHT they are written and read as relaxed agent-scope atomics
// (device-coherent accesses that bypass the non-coherent caches), and a producer waits for
// its own stores to complete (s_waitcnt) before the barrier that precedes the flag store, so
// no L2 write-back or invalidation is needed; otherwise plain accesses ordered by agent-scope
// fences.
template <bool LIGHT>
__device__ inline void xstore(cplx<double>* p, cplx<double> v) {
  if (LIGHT) {
    __hip_atomic_store(&p->x, v.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&p->y, v.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  } else {
    *p = v;
  }
}
template <bool LIGHT>
__device__ inline void xstore(double* p, double v) {
  if (LIGHT) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else *p = v;
}
template <bool LIGHT>
__device__ inline cplx<double> xload(const cplx<double>* p) {
  if (LIGHT)
    return mk(__hip_atomic_load(&p->x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
              __hip_atomic_load(&p->y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  return *p;
}
template <bool LIGHT>
__device__ inline double xload(const double* p) {
  if (LIGHT) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return *p;
}
// Every thread, after its exchanged stores and before the barrier that precedes a flag.
template <bool LIGHT>
__device__ inline void xpublish() {
  if (LIGHT) __builtin_amdgcn_s_waitcnt(0);
  else __threadfence();
}
// Every thread, after the barrier that follows a successful wait.
template <bool LIGHT>
__device__ inline void xacquire() {
  if (!LIGHT) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// Block-wide sum (every thread gets the result); red holds HT_NT / 64 doubles.
__device__ inline double block_sum(double v, double* red) {
  v = wave_sum(v);
  const int wave = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[wave] = v;
  __syncthreads();
  double s = 0.0;
  for (int q = 0; q < HT_NT / 64; ++q) s += red[q];
  return s;
}

struct HetrdShared {
  int ok;
  double tr, ti;  // tau_i
  double beta;
};

// Householder vector of column i (owner only; column in LDS, rows i+1..n-1 below the
// diagonal): ZLARFG on (alpha = A[i+1, i], x = A[i+2:n, i]).  Writes v_i (v[i+1] = 1) to
// V[:, i], tau_i, e_i = beta, d_i = Re A[i, i], then publishes sync[0] = i + 1.
template <bool LIGHT>
__device__ void hetrd_reflector(const cplx<double>* col, int i, int n, cplx<double>* V, long ldv,
                                cplx<double>* tau, double* dd, double* ee, unsigned* sync, double* red) {
  const int tid = threadIdx.x;
  double s = 0.0;
  for (int r = i + 2 + tid; r < n; r += HT_NT) s += col[r].x * col[r].x + col[r].y * col[r].y;
  const double xn2 = block_sum(s, red);
  const cplx<double> alpha = col[i + 1];
  double beta, tr, ti;
  cplx<double> scale;
  if (xn2 == 0.0 && alpha.y == 0.0) {
    beta = alpha.x;
    tr = ti = 0.0;
    scale = mk(0.0, 0.0);
  } else {
    const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
    beta = alpha.x >= 0.0 ? -nrm : nrm;
    tr = (beta - alpha.x) / beta;
    ti = -alpha.y / beta;
    // scale = 1 / (alpha - beta)
    const double ar = alpha.x - beta, ai = alpha.y;
    const double den = ar * ar + ai * ai;
    scale = mk(ar / den, -ai / den);
  }
  cplx<double>* v = V + (long)i * ldv;
  for (int r = i + 1 + tid; r < n; r += HT_NT) xstore<LIGHT>(v + r, (r == i + 1) ? mk(1.0, 0.0) : cmul(col[r], scale));
  if (tid == 0) {
    xstore<LIGHT>(tau + i, mk(tr, ti));
    ee[i] = beta;
    dd[i] = col[i].x;
  }
  // every thread's stores are made visible at agent scope before the barrier, so the flag
  // that follows it needs no release of its own
  xpublish<LIGHT>();
  __syncthreads();
  if (tid == 0) __hip_atomic_store(&sync[0], (unsigned)(i + 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A22 -= v w^H + w v^H on local column lc (global column j), rows i+1..n-1.
__device__ inline void hetrd_update_col(cplx<double>* col, int j, int i, int n, const cplx<double>* vb,
                                        const cplx<double>* wb) {
  const cplx<double> vj = vb[j], wj = wb[j];
  for (int r = i + 1 + threadIdx.x; r < n; r += HT_NT) {
    const cplx<double> a = col[r];
    const cplx<double> t = cmulc(vb[r], wj) + cmulc(wb[r], vj);
    col[r] = a - t;
  }
}

// sync[0]: index + 1 of the last published reflector; sync[1]: arrivals of the p gather.
// pg: 2 x n published p (by step parity), sp: 2 x P x 2 partial sums s = sum conj(p_j) v_j.
template <bool LIGHT>
__global__ void __launch_bounds__(HT_NT) k_hetrd_dist(const cplx<double>* __restrict__ A, long lda, int n,
                                                      int ncl_max, cplx<double>* __restrict__ V, long ldv,
                                                      cplx<double>* __restrict__ tau, double* __restrict__ dd,
                                                      double* __restrict__ ee, cplx<double>* pg, double* sp,
                                                      unsigned* sync, int* status) {
  extern __shared__ char smem_raw[];
  const int P = gridDim.x, w = blockIdx.x, tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  cplx<double>* Acol = reinterpret_cast<cplx<double>*>(smem_raw);
  cplx<double>* vb = Acol + (size_t)ncl_max * n;
  cplx<double>* wb = vb + n;
  double* red = reinterpret_cast<double*>(wb + n);
  __shared__ HetrdShared sh;
  const int ncl = w < n ? (n - 1 - w) / P + 1 : 0;

  // load the owned columns of Hermitian(A): rows <= j from the upper triangle, rows > j as
  // the conjugate of row j
  for (long t = tid; t < (long)ncl * n; t += HT_NT) {
    const int lc = (int)(t / n), r = (int)(t % n);
    const int j = w + lc * P;
    cplx<double> a;
    if (r < j) {
      a = A[(long)j * lda + r];
    } else if (r == j) {
      a = mk(A[(long)j * lda + r].x, 0.0);
    } else {
      const cplx<double> u = A[(long)r * lda + j];
      a = mk(u.x, -u.y);
    }
    Acol[(long)lc * n + r] = a;
  }
  __syncthreads();
  if (w == 0) hetrd_reflector<LIGHT>(Acol, 0, n, V, ldv, tau, dd, ee, sync, red);

  for (int i = 0; i <= n - 2; ++i) {
    const int par = i & 1;
    // B: v_i
    if (tid == 0) sh.ok = wait_geq(&sync[0], (unsigned)(i + 1), status) ? 1 : 0;
    __syncthreads();
    if (!sh.ok) return;
    xacquire<LIGHT>();
    for (int r = i + 1 + tid; r < n; r += HT_NT) vb[r] = xload<LIGHT>(V + (long)i * ldv + r);
    if (tid == 0) {
      const cplx<double> t = xload<LIGHT>(tau + i);
      sh.tr = t.x;
      sh.ti = t.y;
    }
    __syncthreads();
    const cplx<double> taui = mk(sh.tr, sh.ti);
    // C: p_j = tau * A22[:, j]^H v for the owned columns j > i
    const int lc0 = (i + 1 <= w) ? 0 : (i + 1 - w + P - 1) / P;  // first local column with j > i
    double spr = 0.0, spi = 0.0;
    for (int lc = lc0 + wave; lc < ncl; lc += HT_NT / 64) {
      const int j = w + lc * P;
      const cplx<double>* col = Acol + (long)lc * n;
      double ar = 0.0, ai = 0.0;
      for (int r = i + 1 + lane; r < n; r += 64) {
        const cplx<double> a = col[r], v = vb[r];
        // conj(a) * v
        ar += a.x * v.x + a.y * v.y;
        ai += a.x * v.y - a.y * v.x;
      }
      ar = wave_sum(ar);
      ai = wave_sum(ai);
      const cplx<double> p = cmul(taui, mk(ar, ai));
      if (lane == 0) {
        xstore<LIGHT>(pg + (long)par * n + j, p);
        const cplx<double> v = vb[j];
        // conj(p) * v
        spr += p.x * v.x + p.y * v.y;
        spi += p.x * v.y - p.y * v.x;
      }
    }
    {
      double* red2 = red + HT_NT / 64;
      __syncthreads();
      if (lane == 0) {
        red[wave] = spr;
        red2[wave] = spi;
      }
      __syncthreads();
      if (tid == 0) {
        double a = 0.0, b = 0.0;
        for (int q = 0; q < HT_NT / 64; ++q) {
          a += red[q];
          b += red2[q];
        }
        xstore<LIGHT>(sp + ((long)par * P + w) * 2 + 0, a);
        xstore<LIGHT>(sp + ((long)par * P + w) * 2 + 1, b);
      }
    }
    xpublish<LIGHT>();
    __syncthreads();
    if (tid == 0) __hip_atomic_fetch_add(&sync[1], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // D: gather p, form w, update
    if (tid == 0) sh.ok = wait_geq(&sync[1], (unsigned)P * (unsigned)(i + 1), status) ? 1 : 0;
    __syncthreads();
    if (!sh.ok) return;
    xacquire<LIGHT>();
    double sr = 0.0, si = 0.0;
    for (int q = tid; q < P; q += HT_NT) {
      sr += xload<LIGHT>(sp + ((long)par * P + q) * 2 + 0);
      si += xload<LIGHT>(sp + ((long)par * P + q) * 2 + 1);
    }
    sr = block_sum(sr, red);
    si = block_sum(si, red);
    // alpha2 = -1/2 tau s
    const cplx<double> alpha2 = cmul(mk(-0.5 * taui.x, -0.5 * taui.y), mk(sr, si));
    for (int r = i + 1 + tid; r < n; r += HT_NT) wb[r] = xload<LIGHT>(pg + (long)par * n + r) + cmul(alpha2, vb[r]);
    __syncthreads();
    int done = -1;
    if (i + 1 <= n - 2 && (i + 1) % P == w) {
      const int lc = (i + 1) / P;
      hetrd_update_col(Acol + (long)lc * n, i + 1, i, n, vb, wb);
      __syncthreads();
      hetrd_reflector<LIGHT>(Acol + (long)lc * n, i + 1, n, V, ldv, tau, dd, ee, sync, red);
      done = lc;
    }
    for (int lc = lc0; lc < ncl; ++lc) {
      if (lc == done) continue;
      hetrd_update_col(Acol + (long)lc * n, w + lc * P, i, n, vb, wb);
    }
    __syncthreads();
  }
  if ((n - 1) % P == w && tid == 0) dd[n - 1] = Acol[(long)((n - 1) / P) * n + (n - 1)].x;
}

// ZLARFG in place on col[i+1 .. n-1] (alpha = col[i+1], x = col[i+2 ..]): col becomes v_i
// (col[i+1] = 1), tau_i and beta into *sh.  Every thread of the workgroup calls it; it ends
// with a barrier.
__device__ void hetrd_reflector_ip(cplx<double>* col, int i, int n, double* red, HetrdShared* sh) {
  const int tid = threadIdx.x;
  double s = 0.0;
  for (int r = i + 2 + tid; r < n; r += HT_NT) s += col[r].x * col[r].x + col[r].y * col[r].y;
  const double xn2 = block_sum(s, red);
  const cplx<double> alpha = col[i + 1];
  double beta, tr, ti;
  cplx<double> scale;
  if (xn2 == 0.0 && alpha.y == 0.0) {
    beta = alpha.x;
    tr = ti = 0.0;
    scale = mk(0.0, 0.0);
  } else {
    const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
    beta = alpha.x >= 0.0 ? -nrm : nrm;
    tr = (beta - alpha.x) / beta;
    ti = -alpha.y / beta;
    const double ar = alpha.x - beta, ai = alpha.y;
    const double den = ar * ar + ai * ai;
    scale = mk(ar / den, -ai / den);
  }
  __syncthreads();   // every thread has read alpha before col[i + 1] is overwritten
  for (int r = i + 1 + tid; r < n; r += HT_NT) col[r] = (r == i + 1) ? mk(1.0, 0.0) : cmul(col[r], scale);
  if (tid == 0) {
    sh->tr = tr;
    sh->ti = ti;
    sh->beta = beta;
  }
  __syncthreads();
}

// The same reduction with one cross-workgroup exchange per column (PW_TUNE_EIG_SYNC 2).
// Beside its p segments, the owner of column i + 1 publishes that column as it stands before
// step i's update; after the gather every workgroup applies step i's rank-2 update to that
// one column and forms v_{i+1} itself -- the same data and the same arithmetic in every
// workgroup, so the same bits -- instead of waiting for the owner to publish the reflector.
// v_0 comes from column 0 of A, which every workgroup reads.  Workgroup 0 writes V, tau, d
// and e for the eigenvector back-transform and ZSTEDC.
// pg: 2 x n published p, sp: 2 x P x 2 partial sums, cg: 2 x n published next column (all by
// step parity); sync[1]: arrivals.
template <bool LIGHT>
__global__ void __launch_bounds__(HT_NT) k_hetrd_dist1(const cplx<double>* __restrict__ A, long lda, int n,
                                                       int ncl_max, cplx<double>* __restrict__ V, long ldv,
                                                       cplx<double>* __restrict__ tau, double* __restrict__ dd,
                                                       double* __restrict__ ee, cplx<double>* pg, double* sp,
                                                       cplx<double>* cg, unsigned* sync, int* status) {
  extern __shared__ char smem_raw[];
  const int P = gridDim.x, w = blockIdx.x, tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  cplx<double>* Acol = reinterpret_cast<cplx<double>*>(smem_raw);
  cplx<double>* vb = Acol + (size_t)ncl_max * n;   // v_i
  cplx<double>* wb = vb + n;                        // w_i
  cplx<double>* vn = wb + n;                        // column i + 1, then v_{i+1}
  double* red = reinterpret_cast<double*>(vn + n);
  __shared__ HetrdShared sh;
  const int ncl = w < n ? (n - 1 - w) / P + 1 : 0;

  for (long t = tid; t < (long)ncl * n; t += HT_NT) {
    const int lc = (int)(t / n), r = (int)(t % n);
    const int j = w + lc * P;
    cplx<double> a;
    if (r < j) {
      a = A[(long)j * lda + r];
    } else if (r == j) {
      a = mk(A[(long)j * lda + r].x, 0.0);
    } else {
      const cplx<double> u = A[(long)r * lda + j];
      a = mk(u.x, -u.y);
    }
    Acol[(long)lc * n + r] = a;
  }
  // column 0 of Hermitian(A) below the diagonal (conjugated row 0 of the upper triangle)
  for (int r = 1 + tid; r < n; r += HT_NT) {
    const cplx<double> u = A[(long)r * lda];
    vn[r] = mk(u.x, -u.y);
  }
  __syncthreads();
  hetrd_reflector_ip(vn, 0, n, red, &sh);
  if (w == 0) {
    for (int r = 1 + tid; r < n; r += HT_NT) V[r] = vn[r];
    if (tid == 0) {
      tau[0] = mk(sh.tr, sh.ti);
      ee[0] = sh.beta;
      dd[0] = A[0].x;
    }
  }
  {
    cplx<double>* t = vb;
    vb = vn;
    vn = t;
  }

  for (int i = 0; i <= n - 2; ++i) {
    const int par = i & 1;
    const cplx<double> taui = mk(sh.tr, sh.ti);
    // p_j = tau A22[:, j]^H v for the owned columns j > i; the owner of column i + 1
    // publishes it as well
    const int lc0 = (i + 1 <= w) ? 0 : (i + 1 - w + P - 1) / P;
    double spr = 0.0, spi = 0.0;
    for (int lc = lc0 + wave; lc < ncl; lc += HT_NT / 64) {
      const int j = w + lc * P;
      const cplx<double>* col = Acol + (long)lc * n;
      double ar = 0.0, ai = 0.0;
      for (int r = i + 1 + lane; r < n; r += 64) {
        const cplx<double> a = col[r], v = vb[r];
        ar += a.x * v.x + a.y * v.y;
        ai += a.x * v.y - a.y * v.x;
      }
      ar = wave_sum(ar);
      ai = wave_sum(ai);
      const cplx<double> p = cmul(taui, mk(ar, ai));
      if (lane == 0) {
        xstore<LIGHT>(pg + (long)par * n + j, p);
        const cplx<double> v = vb[j];
        spr += p.x * v.x + p.y * v.y;
        spi += p.x * v.y - p.y * v.x;
      }
    }
    if (i + 1 <= n - 2 && (i + 1) % P == w) {
      const cplx<double>* col = Acol + (long)((i + 1) / P) * n;
      for (int r = i + 1 + tid; r < n; r += HT_NT) xstore<LIGHT>(cg + (long)par * n + r, col[r]);
    }
    {
      double* red2 = red + HT_NT / 64;
      __syncthreads();
      if (lane == 0) {
        red[wave] = spr;
        red2[wave] = spi;
      }
      __syncthreads();
      if (tid == 0) {
        double a = 0.0, b = 0.0;
        for (int q = 0; q < HT_NT / 64; ++q) {
          a += red[q];
          b += red2[q];
        }
        xstore<LIGHT>(sp + ((long)par * P + w) * 2 + 0, a);
        xstore<LIGHT>(sp + ((long)par * P + w) * 2 + 1, b);
      }
    }
    xpublish<LIGHT>();
    __syncthreads();
    if (tid == 0) __hip_atomic_fetch_add(&sync[1], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (tid == 0) sh.ok = wait_geq(&sync[1], (unsigned)P * (unsigned)(i + 1), status) ? 1 : 0;
    __syncthreads();
    if (!sh.ok) return;
    xacquire<LIGHT>();
    double sr = 0.0, si = 0.0;
    for (int q = tid; q < P; q += HT_NT) {
      sr += xload<LIGHT>(sp + ((long)par * P + q) * 2 + 0);
      si += xload<LIGHT>(sp + ((long)par * P + q) * 2 + 1);
    }
    sr = block_sum(sr, red);
    si = block_sum(si, red);
    const cplx<double> alpha2 = cmul(mk(-0.5 * taui.x, -0.5 * taui.y), mk(sr, si));
    for (int r = i + 1 + tid; r < n; r += HT_NT) wb[r] = xload<LIGHT>(pg + (long)par * n + r) + cmul(alpha2, vb[r]);
    const bool more = i + 1 <= n - 2;
    if (more)
      for (int r = i + 1 + tid; r < n; r += HT_NT) vn[r] = xload<LIGHT>(cg + (long)par * n + r);
    __syncthreads();
    // step i's update of the owned columns j > i, and of the next column in every workgroup
    for (int lc = lc0; lc < ncl; ++lc) hetrd_update_col(Acol + (long)lc * n, w + lc * P, i, n, vb, wb);
    if (more) {
      hetrd_update_col(vn, i + 1, i, n, vb, wb);
      __syncthreads();
      const double d1 = vn[i + 1].x;
      hetrd_reflector_ip(vn, i + 1, n, red, &sh);
      if (w == 0) {
        for (int r = i + 2 + tid; r < n; r += HT_NT) V[(long)(i + 1) * ldv + r] = vn[r];
        if (tid == 0) {
          tau[i + 1] = mk(sh.tr, sh.ti);
          ee[i + 1] = sh.beta;
          dd[i + 1] = d1;
        }
      }
      cplx<double>* t = vb;
      vb = vn;
      vn = t;
    }
    __syncthreads();
  }
  if ((n - 1) % P == w && tid == 0) dd[n - 1] = Acol[(long)((n - 1) / P) * n + (n - 1)].x;
}

// Back-transform by blocks of WY_NB reflectors (compact WY form, LAPACK ZLARFT / ZLARFB,
// forward, column-wise): H(i0) H(i0+1) ... H(i0+nr-1) = I - V_b T_b V_b^H with T_b upper
// triangular, T[j][j] = tau_j, T[0:j, j] = -tau_j T[0:j, 0:j] (V_b[:, 0:j]^H v_j).  Rows of
// v_i above i + 1 are not stored: they are zero, and v_i[i + 1] = 1.
constexpr int WY_NB = 32;
constexpr int WY_NT_T = 1024;   // k_larft: one workgroup per block of reflectors
constexpr int WY_CW = 4;        // k_apply_q_wy: columns of C per workgroup

__global__ void __launch_bounds__(WY_NT_T) k_larft(const cplx<double>* __restrict__ V, long ldv,
                                                  const cplx<double>* __restrict__ tau, int n,
                                                  cplx<double>* __restrict__ Tout) {
  __shared__ cplx<double> S[WY_NB][WY_NB + 1];
  __shared__ cplx<double> T[WY_NB][WY_NB + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwave = WY_NT_T / 64;
  const int i0 = blockIdx.x * WY_NB;
  const int nr = min(WY_NB, n - 1 - i0);
  // S[l][j] = v_l^H v_j for l < j (rows from i0 + j + 1, where v_j starts with its 1)
  for (int p = wave; p < nr * nr; p += nwave) {
    const int l = p % nr, j = p / nr;
    if (l >= j) continue;
    const cplx<double>* vl = V + (long)(i0 + l) * ldv;
    const cplx<double>* vj = V + (long)(i0 + j) * ldv;
    const int r0 = i0 + j + 1;
    double sr = 0.0, si = 0.0;
    for (int r = r0 + lane; r < n; r += 64) {
      const cplx<double> a = vl[r];
      const cplx<double> b = r == r0 ? mk(1.0, 0.0) : vj[r];
      sr += a.x * b.x + a.y * b.y;
      si += a.x * b.y - a.y * b.x;
    }
    sr = wave_sum(sr);
    si = wave_sum(si);
    if (lane == 0) S[l][j] = mk(sr, si);
  }
  for (int t = tid; t < WY_NB * (WY_NB + 1); t += WY_NT_T) T[t / (WY_NB + 1)][t % (WY_NB + 1)] = mk(0.0, 0.0);
  __syncthreads();
  for (int j = 0; j < nr; ++j) {
    const cplx<double> tj = tau[i0 + j];
    if (tid < j) {
      cplx<double> acc = mk(0.0, 0.0);
      for (int l = tid; l < j; ++l) acc = acc + cmul(T[tid][l], S[l][j]);
      T[tid][j] = mk(-(tj.x * acc.x - tj.y * acc.y), -(tj.x * acc.y + tj.y * acc.x));
    }
    if (tid == 0) T[j][j] = tj;
    __syncthreads();
  }
  cplx<double>* to = Tout + (long)blockIdx.x * WY_NB * WY_NB;
  for (int t = tid; t < WY_NB * WY_NB; t += WY_NT_T) to[t] = T[t % WY_NB][t / WY_NB];
}

// C[:, c] <- Q C[:, c] for the columns c < ncols, Q = B_0 B_1 ... (B_b = I - V_b T_b V_b^H):
// the blocks are applied last to first to CW columns per workgroup staged in LDS.  Each
// block's reflectors are read from HBM / L2 in chunks of WY_RC rows staged in LDS with
// coalesced loads; from a chunk every thread accumulates its share of W = V_b^H C (thread:
// reflector l, column q, half of the chunk's rows), and after W2 = T_b W the same chunks give
// C -= V_b W2 (thread: row, column).  V_b's implicit entries: v_l[i0 + l + 1] = 1, zero above.
constexpr int WY_RC = 64;      // rows of a staged reflector chunk
template <int CW>
__global__ void __launch_bounds__(HT_NT) k_apply_q_wy(const cplx<double>* __restrict__ V, long ldv,
                                                     const cplx<double>* __restrict__ Tb, int n,
                                                     cplx<double>* __restrict__ C, long ldc, int ncols) {
  static_assert(HT_NT == 256 && WY_NB == 32 && WY_RC == 64 && (CW == 4 || CW == 1), "thread maps below");
  extern __shared__ char smem_raw[];
  cplx<double>* Cs = reinterpret_cast<cplx<double>*>(smem_raw);   // [CW][n]
  __shared__ cplx<double> Ts[WY_NB][WY_NB + 1];
  __shared__ cplx<double> Vs[WY_RC][WY_NB + 1];
  __shared__ cplx<double> W[WY_NB][CW], W2[WY_NB][CW];
  __shared__ cplx<double> Wh[2][WY_NB][CW];
  const int tid = threadIdx.x;
  const int c0 = blockIdx.x * CW;
  const int nc = min(CW, ncols - c0);
  for (int t = tid; t < CW * n; t += HT_NT) {
    const int q = t / n, r = t % n;
    Cs[t] = q < nc ? C[(long)(c0 + q) * ldc + r] : mk(0.0, 0.0);
  }
  // phase-1 map: reflector l, column q, row half h (CW = 1: four row quarters)
  const int l1 = tid & 31;
  const int q1 = CW == 4 ? (tid >> 5) & 3 : 0;
  const int h1 = CW == 4 ? tid >> 7 : tid >> 5;            // 0..1 or 0..7
  constexpr int NH = CW == 4 ? 2 : 8;                         // row groups of a chunk
  constexpr int RH = WY_RC / NH;                              // rows per group
  const int nblk = (n - 1 + WY_NB - 1) / WY_NB;
  for (int b = nblk - 1; b >= 0; --b) {
    const int i0 = b * WY_NB;
    const int nr = min(WY_NB, n - 1 - i0);
    for (int t = tid; t < WY_NB * WY_NB; t += HT_NT) Ts[t % WY_NB][t / WY_NB] = Tb[(long)b * WY_NB * WY_NB + t];
    // W = V_b^H C over the rows i0 + 1 .. n - 1, chunk by chunk
    cplx<double> acc = mk(0.0, 0.0);
    for (int rb = i0 + 1; rb < n; rb += WY_RC) {
      __syncthreads();   // the previous chunk's readers are done
      for (int e = tid; e < WY_RC * WY_NB; e += HT_NT) {
        const int rr = e & (WY_RC - 1), l = e >> 6, r = rb + rr;
        cplx<double> v = mk(0.0, 0.0);
        if (r < n && l < nr) {
          const int r1 = i0 + l + 1;
          v = r == r1 ? mk(1.0, 0.0) : (r > r1 ? V[(long)(i0 + l) * ldv + r] : mk(0.0, 0.0));
        }
        Vs[rr][l] = v;
      }
      __syncthreads();
      if (l1 < nr && q1 < nc) {
        for (int k = 0; k < RH; ++k) {
          const int rr = h1 * RH + k, r = rb + rr;
          if (r < n) {
            const cplx<double> a = Vs[rr][l1], x = Cs[q1 * n + r];
            acc.x += a.x * x.x + a.y * x.y;
            acc.y += a.x * x.y - a.y * x.x;
          }
        }
      }
    }
    __syncthreads();
    // combine the row groups
    if (CW == 4) {
      Wh[h1][l1][q1] = acc;
      __syncthreads();
      if (tid < WY_NB * CW) {
        const int l = tid & 31, q = tid >> 5;
        W[l][q] = Wh[0][l][q] + Wh[1][l][q];
      }
    } else {
      // CW = 1: eight row groups of 8 rows; reduce through Wh[.][l][0] in two halves
      if (h1 < 2) Wh[h1][l1][0] = acc;
      __syncthreads();
      for (int g = 2; g < 8; g += 2) {
        if (h1 == g || h1 == g + 1) Wh[h1 - g][l1][0] = Wh[h1 - g][l1][0] + acc;
        __syncthreads();
      }
      if (tid < WY_NB) W[tid][0] = Wh[0][tid][0] + Wh[1][tid][0];
    }
    __syncthreads();
    // W2 = T_b W
    for (int t = tid; t < nr * CW; t += HT_NT) {
      const int k = t % nr, q = t / nr;
      cplx<double> s = mk(0.0, 0.0);
      for (int l = k; l < nr; ++l) s = s + cmul(Ts[k][l], W[l][q]);
      W2[k][q] = s;
    }
    // C -= V_b W2, chunk by chunk (thread: row of the chunk, column)
    for (int rb = i0 + 1; rb < n; rb += WY_RC) {
      __syncthreads();   // W2 written / the previous chunk's readers are done
      for (int e = tid; e < WY_RC * WY_NB; e += HT_NT) {
        const int rr = e & (WY_RC - 1), l = e >> 6, r = rb + rr;
        cplx<double> v = mk(0.0, 0.0);
        if (r < n && l < nr) {
          const int r1 = i0 + l + 1;
          v = r == r1 ? mk(1.0, 0.0) : (r > r1 ? V[(long)(i0 + l) * ldv + r] : mk(0.0, 0.0));
        }
        Vs[rr][l] = v;
      }
      __syncthreads();
      for (int t = tid; t < WY_RC * CW; t += HT_NT) {
        const int rr = t & (WY_RC - 1), q = t >> 6, r = rb + rr;
        if (r < n && q < nc) {
          cplx<double> s = mk(0.0, 0.0);
          for (int l = 0; l < nr; ++l) s = s + cmul(Vs[rr][l], W2[l][q]);
          Cs[q * n + r] = Cs[q * n + r] - s;
        }
      }
    }
    __syncthreads();
  }
  for (int t = tid; t < nc * n; t += HT_NT) {
    const int q = t / n, r = t % n;
    C[(long)(c0 + q) * ldc + r] = Cs[t];
  }
}

__global__ void k_merge_status(int* info, const int* status) {
  if (*status != 0) *info = -1;
}

__global__ void k_fill_u32(unsigned* p, int n, unsigned v) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

// ---- Lowest eigenpairs of the real symmetric tridiagonal T = tridiag(e, d, e) (the
// reduction's output), for the block eigensolvers, which keep the Nst lowest of a Rayleigh-
// Ritz problem of order 2 Nst or 3 Nst (diag_LOBPCG.jl:176-185): rocSOLVER's ZSTEDC computes
// every eigenpair, the kernels below only the ones wanted.
//  1. k_tri_bisect: one wavefront per eigenvalue index i.  Sturm counts (the number of
//     negative pivots of the LDL^T factorization of T - x I, pivots below pivmin replaced by
//     -pivmin) at 64 points of the bracket per round cut it 65-fold, from the Gershgorin
//     interval to eps ||T|| in about 9 rounds.
//  2. k_tri_invit: one wavefront per eigenvector: LU factorization of T - lambda I with
//     partial pivoting (pivots below eps ||T|| raised to it) and two inverse-iteration
//     solves from a hashed start vector, normalised.
//  3. k_tri_cluster: one workgroup per run of eigenvalues closer than TE_CTOL ||T|| to their
//     neighbour: the run's vectors are orthonormalised (modified Gram-Schmidt, twice), then
//     T is diagonalised within their span (cyclic Jacobi on the projected matrix, chunks of
//     at most TE_MAXC vectors), which gives the run's eigenvectors even when the eigenvalues
//     are degenerate or nearly so.
// TE_EXTRA eigenpairs past the wanted ones are computed so that a run crossing the cut is
// resolved whole.
constexpr int TE_EXTRA = 32;
constexpr int TE_MAXC = 32;
constexpr double TE_CTOL = 1e-5;
constexpr double TE_EPS = 2.220446049250313e-16;
constexpr double TE_SAFMIN = 2.2250738585072014e-308;
constexpr int TE_CL_NT = 256;

__device__ inline double wave_min(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ inline double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Gershgorin bounds and max e^2 of T, from LDS copies (one wavefront, every lane gets them).
__device__ inline void tri_bounds(const double* sd, const double* se, int n, double& gl, double& gu,
                                  double& emax2) {
  const int lane = threadIdx.x & 63;
  double lo = 1e308, hi = -1e308, em = 0.0;
  for (int j = lane; j < n; j += 64) {
    const double el = j > 0 ? fabs(se[j - 1]) : 0.0, er = j < n - 1 ? fabs(se[j]) : 0.0;
    lo = fmin(lo, sd[j] - el - er);
    hi = fmax(hi, sd[j] + el + er);
    em = fmax(em, er * er);
  }
  gl = wave_min(lo);
  gu = wave_max(hi);
  emax2 = wave_max(em);
}

// 1 / q to full double precision: the hardware reciprocal refined by two Newton steps (each
// at least doubles the correct bits), a shorter dependency chain than IEEE division.
__device__ inline double rcp_nr(double q) {
  double r = __builtin_amdgcn_rcp(q);
  r = fma(fma(-q, r, 1.0), r, r);
  r = fma(fma(-q, r, 1.0), r, r);
  return r;
}

// Number of eigenvalues of T not above x (negative or zero LDL^T pivots of T - x I).
__device__ inline int tri_sturm(const double* sd, const double* se2, int n, double x, double pivmin) {
  double q = sd[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  int cnt = q <= 0.0 ? 1 : 0;
  // the serial chain runs on registers: each block's LDS operands are read ahead of it
  int j = 1;
  for (; j + 8 <= n; j += 8) {
    double dv[8], ev[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      dv[u] = sd[j + u] - x;
      ev[u] = se2[j + u - 1];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      q = fma(-ev[u], rcp_nr(q), dv[u]);
      if (fabs(q) < pivmin) q = -pivmin;
      cnt += q <= 0.0 ? 1 : 0;
    }
  }
  for (; j < n; ++j) {
    q = fma(-se2[j - 1], rcp_nr(q), sd[j] - x);
    if (fabs(q) < pivmin) q = -pivmin;
    cnt += q <= 0.0 ? 1 : 0;
  }
  return cnt;
}

__global__ void __launch_bounds__(64) k_tri_bisect(const double* __restrict__ d, const double* __restrict__ e,
                                                   int n, double* __restrict__ w) {
  extern __shared__ double tsh[];
  double* sd = tsh;
  double* se = tsh + n;
  double* se2 = se + n;
  const int i = blockIdx.x, lane = threadIdx.x;
  for (int j = lane; j < n; j += 64) {
    sd[j] = d[j];
    const double ej = j < n - 1 ? e[j] : 0.0;
    se[j] = ej;
    se2[j] = ej * ej;
  }
  __syncthreads();
  double gl, gu, emax2;
  tri_bounds(sd, se, n, gl, gu, emax2);
  const double bnorm = fmax(fabs(gl), fabs(gu));
  const double pivmin = TE_SAFMIN * fmax(1.0, emax2);
  double lo = gl - 2.0 * TE_EPS * bnorm * n - 2.0 * pivmin;
  double hi = gu + 2.0 * TE_EPS * bnorm * n + 2.0 * pivmin;
  const double tol = TE_EPS * bnorm + pivmin;
  // count(lo) <= i < count(hi): eigenvalue i (ascending, from 0) lies in (lo, hi]
  for (int round = 0; round < 40 && hi - lo > tol; ++round) {
    const double x = lo + (hi - lo) * (double)(lane + 1) * (1.0 / 65.0);
    const int cnt = tri_sturm(sd, se2, n, x, pivmin);
    const unsigned long long above = __ballot(cnt >= i + 1);
    if (above == 0ull) {
      lo = __shfl(x, 63, 64);
    } else {
      const int f = __ffsll((long long)above) - 1;
      const double xf = __shfl(x, f, 64);
      const double xp = __shfl(x, f > 0 ? f - 1 : 0, 64);
      hi = xf;
      if (f > 0) lo = xp;
    }
  }
  if (lane == 0) w[i] = 0.5 * (lo + hi);
}

// 32-bit integer hash (two multiply-xorshift rounds) to a start-vector entry in [-1, 1).
__device__ inline double tri_start(unsigned i, unsigned r) {
  unsigned h = i * 0x9E3779B1u + r * 0x85EBCA77u + 0x165667B1u;
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  h *= 0x297A2D39u;
  h ^= h >> 15;
  return (double)h * (2.0 / 4294967296.0) - 1.0;
}

__global__ void __launch_bounds__(64) k_tri_invit(const double* __restrict__ d, const double* __restrict__ e, int n,
                                                  const double* __restrict__ w, cplx<double>* __restrict__ Z,
                                                  long ldz) {
  extern __shared__ double tsh[];
  double* sd = tsh;          // d, then (after the factorization) 1 / u0
  double* se = sd + n;       // e
  double* sl = se + n;       // multipliers
  double* su1 = sl + n;      // first superdiagonal of U
  double* su2 = su1 + n;     // second superdiagonal of U (row swaps only)
  double* sx = su2 + n;      // iterate
  int* sw = reinterpret_cast<int*>(sx + n);   // row j swapped with row j + 1
  const int i = blockIdx.x, lane = threadIdx.x;
  for (int j = lane; j < n; j += 64) {
    sd[j] = d[j];
    se[j] = j < n - 1 ? e[j] : 0.0;
    sx[j] = tri_start((unsigned)i, (unsigned)j);
  }
  __syncthreads();
  double gl, gu, emax2;
  tri_bounds(sd, se, n, gl, gu, emax2);
  const double bnorm = fmax(fabs(gl), fabs(gu));
  const double tiny = bnorm > 0.0 ? TE_EPS * bnorm : TE_SAFMIN;
  const double lam = w[i];
  if (lane == 0) {
    // Gaussian elimination of T - lam I with partial pivoting; row j of U is
    // (u0_j, su1_j, su2_j) on columns j, j + 1, j + 2.
    double p = sd[0] - lam, q = n > 1 ? se[0] : 0.0;
    // blocks of 8 steps whose operands (d, e of the rows ahead, not yet overwritten) are read
    // into registers before the serial chain; se[n - 1] is 0, so b1 needs no bound test
    for (int j0 = 0; j0 < n - 1; j0 += 8) {
      double dblk[8], eblk[9];
#pragma unroll
      for (int t = 0; t < 8; ++t) dblk[t] = sd[min(j0 + 1 + t, n - 1)];
#pragma unroll
      for (int t = 0; t < 9; ++t) eblk[t] = se[min(j0 + t, n - 1)];
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const int j = j0 + t;
        if (j >= n - 1) break;
        const double s = eblk[t], a1 = dblk[t] - lam, b1 = eblk[t + 1];
        double u0;
        if (fabs(p) >= fabs(s)) {
          if (fabs(p) < tiny) p = p >= 0.0 ? tiny : -tiny;
          const double lj = s * rcp_nr(p);
          u0 = p;
          su1[j] = q;
          su2[j] = 0.0;
          sw[j] = 0;
          sl[j] = lj;
          p = a1 - lj * q;
          q = b1;
        } else {
          const double lj = p * rcp_nr(s);
          u0 = s;
          su1[j] = a1;
          su2[j] = b1;
          sw[j] = 1;
          sl[j] = lj;
          p = q - lj * a1;
          q = -lj * b1;
        }
        sd[j] = 1.0 / u0;
      }
    }
    if (fabs(p) < tiny) p = p >= 0.0 ? tiny : -tiny;
    sd[n - 1] = 1.0 / p;
  }
  __syncthreads();
  for (int it = 0; it < 2; ++it) {
    if (lane == 0) {
      // forward: the row interchanges and multipliers applied to the right-hand side, the
      // running entry carried in a register
      double yj = sx[0];
      for (int j0 = 0; j0 < n - 1; j0 += 8) {
        double yb[8], lb[8];
        int wb[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          const int j = min(j0 + t, n - 2);
          yb[t] = sx[j + 1];
          lb[t] = sl[j];
          wb[t] = sw[j];
        }
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          const int j = j0 + t;
          if (j >= n - 1) break;
          const bool s = wb[t] != 0;
          const double a = s ? yb[t] : yj, b = s ? yj : yb[t];
          sx[j] = a;
          yj = fma(-lb[t], a, b);
        }
      }
      sx[n - 1] = yj;
      // back substitution through U
      double x2 = 0.0, x1 = sx[n - 1] * sd[n - 1];
      sx[n - 1] = x1;
      for (int j0 = n - 2; j0 >= 0; j0 -= 8) {
        double yb[8], ub1[8], ub2[8], rb[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          const int j = max(j0 - t, 0);
          yb[t] = sx[j];
          ub1[t] = su1[j];
          ub2[t] = su2[j];
          rb[t] = sd[j];
        }
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          const int j = j0 - t;
          if (j < 0) break;
          const double xj = (yb[t] - ub1[t] * x1 - ub2[t] * x2) * rb[t];
          sx[j] = xj;
          x2 = x1;
          x1 = xj;
        }
      }
    }
    __syncthreads();
    double s2 = 0.0, mx = 0.0;
    for (int j = lane; j < n; j += 64) mx = fmax(mx, fabs(sx[j]));
    mx = wave_max(mx);
    const double sc = mx > 0.0 ? 1.0 / mx : 1.0;
    for (int j = lane; j < n; j += 64) {
      const double v = sx[j] * sc;
      s2 += v * v;
    }
    s2 = wave_sum(s2);
    const double f = sc / sqrt(s2);
    __syncthreads();
    for (int j = lane; j < n; j += 64) sx[j] *= f;
    __syncthreads();
  }
  for (int j = lane; j < n; j += 64) Z[(long)i * ldz + j] = mk(sx[j], 0.0);
}

// Block-wide sum (TE_CL_NT threads); red: TE_CL_NT / 64 doubles of LDS.
__device__ inline double block_sum_cl(double v, double* red) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int q = 0; q < TE_CL_NT / 64; ++q) s += red[q];
  return s;
}

// Runs of close eigenvalues among the m2 computed ones that start below m.  kc_max: the
// chunk size that fits LDS for this n (at most TE_MAXC).
__global__ void __launch_bounds__(TE_CL_NT) k_tri_cluster(const double* __restrict__ d, const double* __restrict__ e,
                                                          int n, const double* __restrict__ w, int m, int m2,
                                                          int kc_max, cplx<double>* __restrict__ Z, long ldz) {
  extern __shared__ double tsh[];
  __shared__ double red[TE_CL_NT / 64];
  __shared__ double rot_c[TE_MAXC / 2], rot_s[TE_MAXC / 2];
  __shared__ int rot_p[TE_MAXC / 2], rot_q[TE_MAXC / 2], nrot_sh, rank_sh[TE_MAXC];
  const int j0 = blockIdx.x, tid = threadIdx.x;
  // max(|lower|, |upper| Gershgorin bound) of T, as the other two kernels form it
  double bl = 0.0;
  for (int j = tid; j < n; j += TE_CL_NT) {
    const double el = j > 0 ? fabs(e[j - 1]) : 0.0, er = j < n - 1 ? fabs(e[j]) : 0.0;
    bl = fmax(bl, fmax(fabs(d[j] - el - er), fabs(d[j] + el + er)));
  }
  bl = wave_max(bl);
  if ((tid & 63) == 0) red[tid >> 6] = bl;
  __syncthreads();
  double bnorm = red[0];
  for (int q = 1; q < TE_CL_NT / 64; ++q) bnorm = fmax(bnorm, red[q]);
  __syncthreads();
  const double ctol = TE_CTOL * bnorm;
  if (j0 > 0 && w[j0] - w[j0 - 1] <= ctol) return;   // not the first of its run
  int k = 1;
  while (j0 + k < m2 && w[j0 + k] - w[j0 + k - 1] <= ctol) ++k;
  if (k == 1) return;
  // 1. modified Gram-Schmidt over the run, twice per vector
  for (int c = j0; c < j0 + k; ++c) {
    cplx<double>* zc = Z + (long)c * ldz;
    for (int pass = 0; pass < 2; ++pass) {
      for (int p = j0; p < c; ++p) {
        const cplx<double>* zp = Z + (long)p * ldz;
        double s = 0.0;
        for (int r = tid; r < n; r += TE_CL_NT) s += zp[r].x * zc[r].x;
        s = block_sum_cl(s, red);
        for (int r = tid; r < n; r += TE_CL_NT) zc[r].x -= s * zp[r].x;
        __syncthreads();
      }
    }
    double s = 0.0;
    for (int r = tid; r < n; r += TE_CL_NT) s += zc[r].x * zc[r].x;
    s = block_sum_cl(s, red);
    const double f = 1.0 / sqrt(s);
    for (int r = tid; r < n; r += TE_CL_NT) zc[r].x *= f;
    __syncthreads();
  }
  // 2. Rayleigh-Ritz within chunks of at most kc_max vectors
  double* V = tsh;                              // [n][kc]
  for (int c0 = j0; c0 < j0 + k; c0 += kc_max) {
    const int kc = min(kc_max, j0 + k - c0);
    const int kp = kc + (kc & 1);                 // even player count (one dummy when odd)
    double* M = V + (size_t)n * kc;               // [kc][kc]
    double* Y = M + kc * kc;                      // [kc][kc]
    for (int t = tid; t < n * kc; t += TE_CL_NT) {
      const int r = t / kc, q = t - r * kc;
      V[t] = Z[(long)(c0 + q) * ldz + r].x;
    }
    __syncthreads();
    for (int t = tid; t < kc * kc; t += TE_CL_NT) {
      const int p = t / kc, q = t - p * kc;
      Y[t] = p == q ? 1.0 : 0.0;
      if (p > q) continue;
      double s = 0.0;
      for (int r = 0; r < n; ++r) {
        double tv = d[r] * V[r * kc + q];
        if (r > 0) tv += e[r - 1] * V[(r - 1) * kc + q];
        if (r < n - 1) tv += e[r] * V[(r + 1) * kc + q];
        s += V[r * kc + p] * tv;
      }
      M[p * kc + q] = s;
      M[q * kc + p] = s;
    }
    __syncthreads();
    double fro = 0.0;
    for (int t = tid; t < kc * kc; t += TE_CL_NT) fro += M[t] * M[t];
    fro = sqrt(block_sum_cl(fro, red));
    const double thr = 2.0 * TE_EPS * fro;
    for (int sweep = 0; sweep < 30; ++sweep) {
      if (tid == 0) nrot_sh = 0;
      __syncthreads();
      for (int rnd = 0; rnd < kp - 1; ++rnd) {
        // round-robin pairs: player 0 fixed, the others rotate
        if (tid < kp / 2) {
          const int a = tid == 0 ? 0 : ((tid - 1 + rnd) % (kp - 1)) + 1;
          const int b = ((kp - 2 - tid + rnd) % (kp - 1)) + 1;
          int p = min(a, b), q = max(a, b);
          double cs = 1.0, sn = 0.0;
          if (q < kc) {
            const double apq = M[p * kc + q];
            if (fabs(apq) > thr) {
              const double th = (M[q * kc + q] - M[p * kc + p]) / (2.0 * apq);
              const double t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(1.0 + th * th));
              cs = 1.0 / sqrt(1.0 + t * t);
              sn = t * cs;
              atomicAdd(&nrot_sh, 1);
            }
          } else {
            p = q = -1;
          }
          rot_p[tid] = p;
          rot_q[tid] = q;
          rot_c[tid] = cs;
          rot_s[tid] = sn;
        }
        __syncthreads();
        // rows: M <- J^T M
        for (int t = tid; t < (kp / 2) * kc; t += TE_CL_NT) {
          const int g = t / kc, col = t - g * kc, p = rot_p[g], q = rot_q[g];
          if (p < 0) continue;
          const double cs = rot_c[g], sn = rot_s[g], mp = M[p * kc + col], mq = M[q * kc + col];
          M[p * kc + col] = cs * mp - sn * mq;
          M[q * kc + col] = sn * mp + cs * mq;
        }
        __syncthreads();
        // columns: M <- M J, Y <- Y J
        for (int t = tid; t < (kp / 2) * kc; t += TE_CL_NT) {
          const int g = t / kc, row = t - g * kc, p = rot_p[g], q = rot_q[g];
          if (p < 0) continue;
          const double cs = rot_c[g], sn = rot_s[g];
          const double mp = M[row * kc + p], mq = M[row * kc + q];
          M[row * kc + p] = cs * mp - sn * mq;
          M[row * kc + q] = sn * mp + cs * mq;
          const double yp = Y[row * kc + p], yq = Y[row * kc + q];
          Y[row * kc + p] = cs * yp - sn * yq;
          Y[row * kc + q] = sn * yp + cs * yq;
        }
        __syncthreads();
      }
      const int nr = nrot_sh;
      __syncthreads();
      if (nr == 0) break;
    }
    // ascending order of the Ritz values: rank of each (ties by index)
    if (tid < kc) {
      const double v = M[tid * kc + tid];
      int rk = 0;
      for (int q = 0; q < kc; ++q) {
        const double u = M[q * kc + q];
        rk += (u < v || (u == v && q < tid)) ? 1 : 0;
      }
      rank_sh[rk] = tid;
    }
    __syncthreads();
    for (int t = tid; t < n * kc; t += TE_CL_NT) {
      const int r = t / kc, c = t - r * kc, src = rank_sh[c];
      double s = 0.0;
      for (int p = 0; p < kc; ++p) s += V[r * kc + p] * Y[p * kc + src];
      Z[(long)(c0 + c) * ldz + r] = mk(s, 0.0);
    }
    __syncthreads();
  }
}

// Largest chunk of k_tri_cluster whose V, M and Y fit the dynamic LDS budget at order n.
static int tri_cluster_chunk(int n, size_t* lds) {
  const size_t budget = 144 * 1024;
  int kc = TE_MAXC;
  while (kc > 1 && ((size_t)n * kc + 2 * (size_t)kc * kc) * sizeof(double) > budget) --kc;
  *lds = ((size_t)n * kc + 2 * (size_t)kc * kc) * sizeof(double);
  return kc;
}

// Eigenvalues of the m lowest eigenpairs of T (d, e: n and n - 1 doubles) into w and, when
// vectors, their eigenvectors into Z's first m columns (complex, imaginary parts zero).
// wt: workspace of m + TE_EXTRA doubles (the computed eigenvalues).
int tri_lowest(pw_ctx* c, int n, int m, const double* d, const double* e, double* w, double* wt, bool vectors,
               cplx<double>* Z, long ldz) {
  const int m2 = std::min(n, m + (vectors ? TE_EXTRA : 0));
  hipLaunchKernelGGL(k_tri_bisect, dim3(m2), dim3(64), 3 * (size_t)n * sizeof(double), c->stream, d, e, n, wt);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return hip_check(c, err, "k_tri_bisect");
  err = hipMemcpyAsync(w, wt, (size_t)m * sizeof(double), hipMemcpyDeviceToDevice, c->stream);
  if (err != hipSuccess) return hip_check(c, err, "eigenvalue copy");
  if (!vectors) return PW_OK;
  const size_t lds_inv = 6 * (size_t)n * sizeof(double) + (size_t)n * sizeof(int);
  hipLaunchKernelGGL(k_tri_invit, dim3(m2), dim3(64), lds_inv, c->stream, d, e, n, wt, Z, ldz);
  err = hipGetLastError();
  if (err != hipSuccess) return hip_check(c, err, "k_tri_invit");
  size_t lds_cl = 0;
  const int kc = tri_cluster_chunk(n, &lds_cl);
  hipLaunchKernelGGL(k_tri_cluster, dim3(m), dim3(TE_CL_NT), lds_cl, c->stream, d, e, n, wt, m, m2, kc, Z, ldz);
  return hip_check(c, hipGetLastError(), "k_tri_cluster");
}

// Workgroups of the distributed reduction for order n: the fewest whose column share,
// the three staging vectors and the reduction scratch fit in LDS; 0 when n is out of range.
int hetrd_workgroups(int n, int* ncl_max, size_t* lds_bytes, int min_wg) {
  if (n < 2) return 0;
  const size_t extra = 3 * (size_t)n * sizeof(cplx<double>) + 2 * (HT_NT / 64) * sizeof(double);
  if (extra >= HT_LDS) return 0;
  const long per = (long)((HT_LDS - extra) / ((size_t)n * sizeof(cplx<double>)));
  if (per < 1) return 0;
  int P = (int)((n + per - 1) / per);
  // more workgroups than LDS needs: each holds fewer columns, so the per-column work of
  // every step (its share of p = A v and of the rank-2 update) shrinks
  P = std::max(P, std::min(n, std::max(8, min_wg)));
  if (P > 256) return 0;
  const int ncl = (n + P - 1) / P;
  *ncl_max = ncl;
  *lds_bytes = (size_t)ncl * n * sizeof(cplx<double>) + extra;
  return P;
}

// Largest order the device reduction takes (beyond it, rocSOLVER's ZHEEVD runs).
int heevd_dist_max_n() {
  int lo = 2, hi = 4096, best = 0;
  while (lo <= hi) {
    const int mid = (lo + hi) / 2;
    int ncl;
    size_t b;
    if (hetrd_workgroups(mid, &ncl, &b, 0) > 0) {
      best = mid;
      lo = mid + 1;
    } else {
      hi = mid - 1;
    }
  }
  return best;
}

// Eigenvalues (ascending, into d_w on the device) and, when nvec > 0, the eigenvectors of
// the nvec lowest into A's first nvec columns.  rocSOLVER's handle runs on the context stream.
// Enqueues everything on the context stream; d_info gets 0, or the ZSTEDC failure code, or
// -1 when a cross-workgroup wait of the reduction timed out.  lowest (with nvec > 0): only
// the nvec lowest eigenvalues are wanted, and the tridiagonal eigenpairs come from the
// device kernels above instead of ZSTEDC (d_w then holds nvec values; with nvec = 0 all n).
int heevd_dist(pw_ctx* c, void* rb, int n, cplx<double>* A, long lda, double* d_w, int nvec, int* d_info,
               bool lowest) {
  int ncl = 0;
  size_t lds = 0;
  const int P = hetrd_workgroups(n, &ncl, &lds, c->eig_min_wg);
  if (P <= 0) return set_error(c, PW_ERR_ARG, "heevd_dist: order " + std::to_string(n) + " out of range");
  // workspace: V (n x n), T blocks, tau (n), pg (2n), cg (2n) complex; e (n), sp (4P)
  // doubles; sync (2) + status
  const size_t vbytes = (size_t)n * n * sizeof(cplx<double>);
  const int nblk = (n - 1 + WY_NB - 1) / WY_NB;
  const size_t tbytes = (size_t)nblk * WY_NB * WY_NB * sizeof(cplx<double>);
  // (+ 2n doubles: the tridiagonal's diagonal and the computed eigenvalues of the device path;
  // + n nvec + nvec^2 complex: its copy of the vectors and their Gram matrix)
  const size_t lw = lowest && nvec > 0 ? ((size_t)n * nvec + (size_t)nvec * nvec) * sizeof(cplx<double>) : 0;
  const size_t off_lw = (vbytes + tbytes + 5 * (size_t)n * sizeof(cplx<double>) +
                         ((size_t)3 * n + 4 * P) * sizeof(double) + 64 + 255) & ~(size_t)255;
  const size_t need = off_lw + lw;
  int rc = ensure(c, &c->d_eig, &c->eig_bytes, need);
  if (rc) return rc;
  char* base = reinterpret_cast<char*>(c->d_eig);
  cplx<double>* V = reinterpret_cast<cplx<double>*>(base);
  cplx<double>* Tb = V + (size_t)n * n;
  cplx<double>* tau = Tb + (size_t)nblk * WY_NB * WY_NB;
  cplx<double>* pg = tau + n;
  cplx<double>* cg = pg + 2 * (size_t)n;
  double* E = reinterpret_cast<double*>(cg + 2 * (size_t)n);
  double* sp = E + n;
  double* Dt = sp + 4 * (size_t)P;
  double* Wt = Dt + n;
  unsigned* sync = reinterpret_cast<unsigned*>(Wt + n);
  int* status = reinterpret_cast<int*>(sync + 2);
  cplx<double>* Zc = reinterpret_cast<cplx<double>*>(base + off_lw);
  cplx<double>* Gz = Zc + (size_t)n * (nvec > 0 ? nvec : 0);
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, c->stream, sync, 3, 0u);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_check(c, e, "k_fill_u32");
  const cplx<double>* Ac = A;
  long ldal = lda, ldv = n;
  int nn = n, nclm = ncl;
  double* dd = lowest ? Dt : d_w;
  void* args[] = {(void*)&Ac, (void*)&ldal, (void*)&nn, (void*)&nclm, (void*)&V, (void*)&ldv, (void*)&tau,
                  (void*)&dd, (void*)&E, (void*)&pg, (void*)&sp, (void*)&sync, (void*)&status};
  void* args1[] = {(void*)&Ac, (void*)&ldal, (void*)&nn, (void*)&nclm, (void*)&V, (void*)&ldv, (void*)&tau,
                   (void*)&dd, (void*)&E, (void*)&pg, (void*)&sp, (void*)&cg, (void*)&sync, (void*)&status};
  // by order: the single-exchange form is faster up to n ~ 480 (n = 160: 2.33 -> 2.19 ms),
  // slower at n = 960 (15.8 -> 16.6 ms), where each step's local work outweighs an exchange
  const int mode = c->eig_light == 3 ? (n <= 640 ? 2 : 1) : c->eig_light;
  if (mode == 2)
    e = hipLaunchCooperativeKernel((const void*)k_hetrd_dist1<true>, dim3(P), dim3(HT_NT), args1, (unsigned)lds,
                                   c->stream);
  else
    e = hipLaunchCooperativeKernel(mode ? (const void*)k_hetrd_dist<true> : (const void*)k_hetrd_dist<false>,
                                   dim3(P), dim3(HT_NT), args, (unsigned)lds, c->stream);
  if (e != hipSuccess) return hip_check(c, e, "k_hetrd_dist (cooperative launch)");
  // the timeout status of the reduction is folded into d_info below
  if (lowest) {
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, c->stream, reinterpret_cast<unsigned*>(d_info), 1, 0u);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_check(c, e, "k_fill_u32");
    int rc2 = tri_lowest(c, n, nvec > 0 ? nvec : n, Dt, E, d_w, Wt, nvec > 0, A, lda);
    if (!rc2 && nvec > 0) {
      // inverse-iteration vectors of eigenvalues closer than ~1e-4 ||T|| (and outside one
      // run) overlap by ~eps ||T|| / gap: one first-order Loewdin step, Z <- Z (3 I - G) / 2
      // with G = Z^H Z, leaves overlaps of second order and residuals at eps ||T||
      rc2 = copy_cols(c, n, nvec, A, lda, Zc, n);
      if (!rc2) rc2 = zgemm(c, 1, nvec, nvec, n, mk(1.0, 0.0), Zc, n, Zc, n, mk(0.0, 0.0), Gz, nvec, 0);
      if (!rc2) rc2 = zgemm(c, 0, n, nvec, nvec, mk(-0.5, 0.0), Zc, n, Gz, nvec, mk(1.5, 0.0), A, lda, 0);
    }
    if (rc2) return rc2;
  } else {
    rocblas_handle h = reinterpret_cast<rocblas_handle>(rb);
    // ZSTEDC with the tridiagonal's eigenvectors even when only values are wanted: without
    // them rocSOLVER runs its single-thread QL iteration (ZSTERF), which costs ~13 ms at n = 160
    rocblas_status st = rocsolver_zstedc(h, rocblas_evect_tridiagonal, n, d_w, E,
                                         reinterpret_cast<rocblas_double_complex*>(A), (int)lda, d_info);
    if (st != rocblas_status_success)
      return set_error(c, PW_ERR_HIP, "zstedc: rocSOLVER status " + std::to_string((int)st));
  }
  if (nvec > 0) {
    hipLaunchKernelGGL(k_larft, dim3(nblk), dim3(WY_NT_T), 0, c->stream, V, ldv, tau, n, Tb);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_check(c, e, "k_larft");
    // four columns per workgroup while they fit LDS beside the static staging (~60 KiB)
    const bool cw4 = (size_t)WY_CW * n * sizeof(cplx<double>) <= 96 * 1024;
    const int cw = cw4 ? WY_CW : 1;
    const int blocks = (nvec + cw - 1) / cw;
    const size_t sh = (size_t)cw * n * sizeof(cplx<double>);
    if (cw4)
      hipLaunchKernelGGL(k_apply_q_wy<WY_CW>, dim3(blocks), dim3(HT_NT), sh, c->stream, V, ldv, Tb, n, A, lda, nvec);
    else
      hipLaunchKernelGGL(k_apply_q_wy<1>, dim3(blocks), dim3(HT_NT), sh, c->stream, V, ldv, Tb, n, A, lda, nvec);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_check(c, e, "k_apply_q_wy");
  }
  // d_info = -1 when the reduction timed out (one tiny kernel keeps it on the stream)
  hipLaunchKernelGGL(k_merge_status, dim3(1), dim3(1), 0, c->stream, d_info, status);
  return hip_check(c, hipGetLastError(), "k_merge_status");
}

}  // namespace pw
