// Small dense problems of the block eigensolvers (order n <= 64), one workgroup per matrix
// with the matrix in LDS, for many matrices in one launch: the Rayleigh-Ritz and Gram
// matrices of every k-point of a batched LOBPCG step (2 Nst, 3 Nst = 16, 24 at 8 bands), and
// those of small single k-point problems.  They replace the LAPACK calls of
// diag_LOBPCG.jl:106-118 (cholesky, inv) and :176-190 (eigen) for these orders, launch
// nothing else and need no library workspace, so a step that uses them can be captured
// into a graph whole.
//
//  - k_heev_jacobi: Hermitian eigenpairs by cyclic Jacobi rotations in tournament order
//    (n/2 disjoint rotations per round, all applied at once), a pair
//    rotated while |A_pq| > max(1e-15 sqrt|A_pp A_qq|, 2 eps ||A||_F), until no pair passes
//    that test (checked on every pair after each sweep, in place of a sweep that would only
//    confirm it); eigenvalues sorted ascending with their vectors.  Complex 2x2 step: the pair (p, q) with A_pq = |g| e^{i phi} is made
//    real by diag(1, e^{-i phi}) and diagonalised by the classical rotation
//    t = sign(tau) / (|tau| + sqrt(1 + tau^2)), tau = (A_qq - A_pp) / (2 |g|).
//  - k_chol_inv_small: U = chol(Hermitian(A)).U by right-looking elimination, then U^-1 by
//    back substitution over rows, the matrix held in registers (a 4 x 4 tile per thread) and
//    one barrier per step; the strict lower triangle zeroed.  Status: 0, or the index
//    (1-based) of the first non-positive pivot.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "pw_internal.h"

namespace pw {

constexpr int SM_NT = 256;
constexpr int SM_NMAX = 64;
constexpr int SM_SWEEPS = 30;

__device__ inline double sm_block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int q = 0; q < SM_NT / 64; ++q) s += red[q];
  return s;
}

// Matrix b of the stack at A + b * stride (leading dimension lda): eigenvalues (ascending)
// into w + b * n, the nvec lowest eigenvectors into the matrix's first nvec columns.
// A <- J^H A J is applied per 2 x 2 block (rows of pair k, columns of pair k'; the pairs of
// a round are disjoint, so every element belongs to one block): the columns' rotation, then
// the rows', as a two-pass form would apply them, in one pass beside the update of V.
__device__ inline void jac_rot(cplx<double>& x0, cplx<double>& x1, double c, double s, cplx<double> ph) {
  // (x0, x1) <- (c x0 - s ph x1, s x0 + c ph x1)
  const cplx<double> xp = x0, xq = cmul(x1, ph);
  x0 = mk(c * xp.x - s * xq.x, c * xp.y - s * xq.y);
  x1 = mk(s * xp.x + c * xq.x, s * xp.y + c * xq.y);
}

__global__ void __launch_bounds__(SM_NT) k_heev_jacobi(cplx<double>* __restrict__ A, long lda, long stride, int n,
                                                    double* __restrict__ w, int nvec, int* __restrict__ info) {
  extern __shared__ char smem_raw[];
  cplx<double>* a = reinterpret_cast<cplx<double>*>(smem_raw);   // n x n
  cplx<double>* v = a + n * n;                                     // n x n
  __shared__ double rc[SM_NMAX / 2], rs[SM_NMAX / 2];
  __shared__ cplx<double> rph[SM_NMAX / 2];       // e^{-i phi} of each pair
  __shared__ int rp[SM_NMAX / 2], rq[SM_NMAX / 2];   // rq < 0: the index paired with the dummy player
  __shared__ double dsort[SM_NMAX];
  __shared__ int rank[SM_NMAX];
  cplx<double>* M = A + (long)blockIdx.x * stride;
  const int lane = threadIdx.x;
  const int nn = n * n;
  __shared__ double red[SM_NT / 64];
  __shared__ int flag, flag2;
  // Hermitian(M): upper triangle, real diagonal
  double fro = 0.0;
  for (int t = lane; t < nn; t += SM_NT) {
    const int i = t % n, j = t / n;
    cplx<double> x;
    if (i < j) x = M[(long)j * lda + i];
    else if (i == j) x = mk(M[(long)j * lda + i].x, 0.0);
    else { const cplx<double> u = M[(long)i * lda + j]; x = mk(u.x, -u.y); }
    a[i + n * j] = x;
    v[i + n * j] = mk(i == j ? 1.0 : 0.0, 0.0);
    fro += x.x * x.x + x.y * x.y;
  }
  fro = sm_block_sum(fro, red);
  // below 2 eps ||A||_F an off-diagonal entry is rounding noise of the rotations themselves
  // (rotating it changes eigenvalues by less than the backward error of the decomposition);
  // a tighter floor keeps pairs with small or disparate diagonals rotating on noise for
  // sweep after sweep
  const double gfloor = 2.0 * 2.220446049250313e-16 * sqrt(fro);
  const int m = (n + 1) / 2 * 2;     // players of the tournament (one dummy when n is odd)
  const int npair = m / 2;
  bool converged = false;
  for (int sweep = 0; sweep < SM_SWEEPS; ++sweep) {
    if (lane == 0) flag = 0;
    int nrot = 0;
    for (int r = 0; r < m - 1; ++r) {
      if (lane < npair) {
        int p, q;
        if (lane == 0) {
          p = m - 1;
          q = r;
        } else {
          // (r + lane) mod (m - 1) and (r - lane) mod (m - 1) by one conditional subtraction
          // each (both sums lie in [0, 2 (m - 1))): this runs on every round's critical path
          p = r + lane;
          if (p >= m - 1) p -= m - 1;
          q = r - lane + (m - 1);
          if (q >= m - 1) q -= m - 1;
        }
        if (p > q) { const int t = p; p = q; q = t; }
        double c = 1.0, s = 0.0;
        cplx<double> ph = mk(1.0, 0.0);
        if (q < n) {
          const cplx<double> g = a[p + n * q];
          const double ag2 = g.x * g.x + g.y * g.y;
          const double app = a[p + n * p].x, aqq = a[q + n * q].x;
          // the tests of |A_pq| on squares, and 1 / |A_pq| by one reciprocal square root: the
          // rotation parameters are a serial chain on every round's critical path
          if (ag2 > 1e-30 * fabs(app * aqq) && ag2 > gfloor * gfloor) {
            ++nrot;
            const double rag = rsqrt(ag2);   // 1 / |A_pq|
            const double tau = (aqq - app) * (0.5 * rag);
            const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            c = rsqrt(1.0 + t * t);
            s = t * c;
            ph = mk(g.x * rag, -g.y * rag);   // e^{-i phi}
          }
        } else {
          q = -1;   // p meets the dummy: not rotated this round
        }
        rp[lane] = p;
        rq[lane] = q;
        rc[lane] = c;
        rs[lane] = s;
        rph[lane] = ph;
      }
      __syncthreads();
      // A <- J^H A J, block (rows of pair k, columns of pair k2), one block per thread (a
      // lane map on 32-wide rows of blocks idles most lanes at the batched orders 16 and 24)
      for (int t = lane; t < npair * npair; t += SM_NT) {
        const int k = t % npair, k2 = t / npair;
        const int p = rp[k], q = rq[k], p2 = rp[k2], q2 = rq[k2];
        const double s = rs[k], s2 = rs[k2];
        if (s == 0.0 && s2 == 0.0) continue;
        const bool qv = q >= 0, q2v = q2 >= 0;
        cplx<double> b00 = a[p + n * p2];
        cplx<double> b01 = q2v ? a[p + n * q2] : mk(0.0, 0.0);
        cplx<double> b10 = qv ? a[q + n * p2] : mk(0.0, 0.0);
        cplx<double> b11 = (qv && q2v) ? a[q + n * q2] : mk(0.0, 0.0);
        if (s2 != 0.0) {   // columns p2, q2
          const double c2 = rc[k2];
          const cplx<double> ph2 = rph[k2];
          jac_rot(b00, b01, c2, s2, ph2);
          jac_rot(b10, b11, c2, s2, ph2);
        }
        if (s != 0.0) {    // rows p, q: e^{i phi}
          const double c1 = rc[k];
          const cplx<double> phc = mk(rph[k].x, -rph[k].y);
          jac_rot(b00, b10, c1, s, phc);
          jac_rot(b01, b11, c1, s, phc);
        }
        a[p + n * p2] = b00;
        if (q2v) a[p + n * q2] = b01;
        if (qv) a[q + n * p2] = b10;
        if (qv && q2v) a[q + n * q2] = b11;
      }
      // V <- V J, columns of pair k2
      for (int t = lane; t < npair * n; t += SM_NT) {
        const int k2 = t / n, i = t % n;
        const int p2 = rp[k2], q2 = rq[k2];
        const double s2 = rs[k2];
        if (q2 < 0 || s2 == 0.0) continue;
        cplx<double> x0 = v[i + n * p2], x1 = v[i + n * q2];
        jac_rot(x0, x1, rc[k2], s2, rph[k2]);
        v[i + n * p2] = x0;
        v[i + n * q2] = x1;
      }
      __syncthreads();
    }
    if (nrot) flag = 1;   // any thread that rotated this sweep
    if (lane == 0) flag2 = 0;
    __syncthreads();
    const int more = flag;
    if (!more) {
      converged = true;
      break;
    }
    // would the next sweep rotate any pair?  The same test as the rotations', on every pair
    // of the matrix as it stands: when none passes, that sweep would change nothing, so the
    // sweep that only confirms convergence is skipped (the result is the same to the bit)
    for (int t = lane; t < nn; t += SM_NT) {
      const int i = t % n, j = t / n;
      if (i < j) {
        const cplx<double> g = a[i + n * j];
        const double ag2 = g.x * g.x + g.y * g.y;
        if (ag2 > 1e-30 * fabs(a[i + n * i].x * a[j + n * j].x) && ag2 > gfloor * gfloor) flag2 = 1;
      }
    }
    __syncthreads();
    const int again = flag2;
    __syncthreads();
    if (!again) {
      converged = true;
      break;
    }
  }
  // eigenvalues: the diagonal; sort ascending (ties by index) with the vectors
  if (lane < n) dsort[lane] = a[lane + n * lane].x;
  __syncthreads();
  if (lane < n) {
    const double d = dsort[lane];
    int r = 0;
    for (int j = 0; j < n; ++j) r += (dsort[j] < d || (dsort[j] == d && j < lane)) ? 1 : 0;
    rank[lane] = r;
    w[(long)blockIdx.x * n + r] = d;
  }
  __syncthreads();
  for (int t = lane; t < nn; t += SM_NT) {
    const int i = t % n, j = t / n;
    const int r = rank[j];
    if (r < nvec) M[(long)r * lda + i] = v[i + n * j];
  }
  if (lane == 0 && info) info[blockIdx.x] = converged ? 0 : 1;
}

// Matrix b of the stack: inv(chol(Hermitian(A)).U) in place, strict lower triangle zeroed;
// info[b] (potrf) and info[nbatch + b] (trtri) as rocSOLVER would set them.
// Thread (ti, tj) of a G x G grid (G = 16 or 32) holds the entries (ti + G a, tj + G b), a, b < 64 / G, in
// registers from the load to the store.  Factorization (right-looking): at step k the holders
// of row k publish it to an LDS row (two rows alternate), one barrier, then every thread
// scales what it needs of row k by 1 / sqrt(A_kk) and updates its own trailing entries.
// Inverse (X = U^-1, rows from the last): the holders of row i finish X_ij = (delta_ij -
// S_ij) / U_ii and publish it, one barrier, then every thread adds U_ri X_ij to its sums S_rj
// of the rows r < i (U's column i read from LDS, where U was stored after the factorization).
// One barrier per step in both phases.
template <int G>
__global__ void __launch_bounds__(G * G) k_chol_inv_small(cplx<double>* __restrict__ A, long lda, long stride, int n,
                                                           int* __restrict__ info, int nbatch) {
  constexpr int E = SM_NMAX / G;   // entries per thread along each axis
  extern __shared__ char smem_raw[];
  cplx<double>* u = reinterpret_cast<cplx<double>*>(smem_raw);   // n x n: U after the factorization
  __shared__ cplx<double> rowbuf[2][SM_NMAX];
  __shared__ double rdiag[SM_NMAX];
  static_assert((G == 16 || G == 32) && SM_NMAX == 64, "G x G grid, E x E entries per thread");
  cplx<double>* M = A + (long)blockIdx.x * stride;
  const int tid = threadIdx.x, ti = tid % G, tj = tid / G;
  cplx<double> r[E][E];   // r[b][a]: entry (ti + G a, tj + G b)
#pragma unroll
  for (int b = 0; b < E; ++b)
#pragma unroll
    for (int a = 0; a < E; ++a) {
      const int i = ti + G * a, j = tj + G * b;
      cplx<double> v = mk(0.0, 0.0);
      if (j < n && i <= j) {
        v = M[(long)j * lda + i];
        if (i == j) v.y = 0.0;
      }
      r[b][a] = v;
    }
  int bad = 0;
  for (int k = 0; k < n; ++k) {
    const int buf = k & 1, ak = k / G;
    if (ti == k % G) {
#pragma unroll
      for (int a = 0; a < E; ++a)
        if (a == ak) {
#pragma unroll
          for (int b = 0; b < E; ++b) {
            const int j = tj + G * b;
            if (j >= k && j < n) rowbuf[buf][j] = r[b][a];
          }
        }
    }
    __syncthreads();
    const double dkk = rowbuf[buf][k].x;
    if (!(dkk > 0.0)) {   // the same value on every thread: a uniform exit
      bad = k + 1;
      break;
    }
    // 1 / sqrt and sqrt from one reciprocal square root; the row reads are unconditional
    // (every index < SM_NMAX is a valid LDS slot) so they issue together and wait once
    const double rd = rsqrt(dkk), d = dkk * rd;
    cplx<double> uki[E], ukj[E];
#pragma unroll
    for (int a = 0; a < E; ++a) {
      const int i = ti + G * a;
      const cplx<double> v = rowbuf[buf][i];
      uki[a] = (i > k && i < n) ? scal(rd, v) : mk(0.0, 0.0);
    }
#pragma unroll
    for (int b = 0; b < E; ++b) {
      const int j = tj + G * b;
      const cplx<double> v = rowbuf[buf][j];
      ukj[b] = (j > k && j < n) ? scal(rd, v) : mk(0.0, 0.0);
    }
    // tile rows a and columns b wholly before k are finished: skipped on every thread alike
#pragma unroll
    for (int b = 0; b < E; ++b) {
      if (G * b + G - 1 < k || G * b >= n) continue;
#pragma unroll
      for (int a = 0; a < E; ++a) {
        if (G * a + G - 1 < k || G * a > G * b + G - 1) continue;
        const int i = ti + G * a, j = tj + G * b;
        if (j < n && i <= j) {
          if (i == k) r[b][a] = j == k ? mk(d, 0.0) : scal(rd, r[b][a]);            // row k of U
          else if (i > k) r[b][a] = r[b][a] - cmulc(ukj[b], uki[a]);               // A_ij -= conj(U_ki) U_kj
        }
      }
    }
  }
  const int bb = blockIdx.x;
  if (bad) {
    if (tid == 0) {
      info[bb] = bad;
      info[nbatch + bb] = 0;
    }
    return;
  }
  // U into LDS (its columns feed the inverse), 1 / U_ii
#pragma unroll
  for (int b = 0; b < E; ++b)
#pragma unroll
    for (int a = 0; a < E; ++a) {
      const int i = ti + G * a, j = tj + G * b;
      if (j < n && i <= j) {
        u[i + n * j] = r[b][a];
        if (i == j) rdiag[i] = 1.0 / r[b][a].x;
      }
      r[b][a] = mk(0.0, 0.0);   // from here: the sums S, then X
    }
  __syncthreads();
  for (int i = n - 1; i >= 0; --i) {
    const int buf = i & 1, ai = i / G;
    if (ti == i % G) {
#pragma unroll
      for (int a = 0; a < E; ++a)
        if (a == ai) {
#pragma unroll
          for (int b = 0; b < E; ++b) {
            const int j = tj + G * b;
            if (j >= i && j < n) {
              const cplx<double> x = scal(rdiag[i], mk((j == i ? 1.0 : 0.0) - r[b][a].x, -r[b][a].y));
              r[b][a] = x;
              rowbuf[buf][j] = x;
            }
          }
        }
    }
    __syncthreads();
    cplx<double> uri[E], xij[E];
#pragma unroll
    for (int a = 0; a < E; ++a) {
      const int rr = ti + G * a;
      const cplx<double> v = u[min(rr, n - 1) + n * i];   // unconditional (clamped) read
      uri[a] = rr < i ? v : mk(0.0, 0.0);
    }
#pragma unroll
    for (int b = 0; b < E; ++b) {
      const int j = tj + G * b;
      const cplx<double> v = rowbuf[buf][j];
      xij[b] = (j >= i && j < n) ? v : mk(0.0, 0.0);
    }
#pragma unroll
    for (int b = 0; b < E; ++b) {
      if (G * b + G - 1 < i || G * b >= n) continue;
#pragma unroll
      for (int a = 0; a < E; ++a) {
        if (G * a >= i) continue;
        const int rr = ti + G * a, j = tj + G * b;
        if (rr < i && j >= i && j < n) r[b][a] = r[b][a] + cmul(uri[a], xij[b]);
      }
    }
  }
  // X (upper triangle) and zeros below the diagonal
#pragma unroll
  for (int b = 0; b < E; ++b)
#pragma unroll
    for (int a = 0; a < E; ++a) {
      const int i = ti + G * a, j = tj + G * b;
      if (i < n && j < n) M[(long)j * lda + i] = i <= j ? cplx<double>(r[b][a]) : mk(0.0, 0.0);
    }
  if (tid == 0) {
    info[bb] = 0;
    info[nbatch + bb] = 0;
  }
}

int heev_small_batch(pw_ctx* c, int n, int nbatch, cplx<double>* A, long lda, long stride, double* d_w, int nvec,
                     int* d_info) {
  if (n < 1 || n > SM_NMAX || nbatch < 1) return set_error(c, PW_ERR_ARG, "heev_small_batch: order out of range");
  const size_t lds = 2 * (size_t)n * n * sizeof(cplx<double>);
  hipLaunchKernelGGL(k_heev_jacobi, dim3(nbatch), dim3(SM_NT), lds, c->stream, A, lda, stride, n, d_w, nvec, d_info);
  return hip_check(c, hipGetLastError(), "k_heev_jacobi");
}

int chol_inv_small_batch(pw_ctx* c, int n, int nbatch, cplx<double>* A, long lda, long stride, int* d_info) {
  if (n < 1 || n > SM_NMAX || nbatch < 1) return set_error(c, PW_ERR_ARG, "chol_inv_small_batch: order out of range");
  const size_t lds = (size_t)n * n * sizeof(cplx<double>);
  // one matrix above order 32 (the recursive inverse's blocks): a 32 x 32 grid of 2 x 2 tiles,
  // a quarter of the arithmetic per thread and step; stacks of small matrices keep 16 x 16
  if (nbatch == 1 && n > 32)
    hipLaunchKernelGGL(k_chol_inv_small<32>, dim3(1), dim3(1024), lds, c->stream, A, lda, stride, n, d_info, nbatch);
  else
    hipLaunchKernelGGL(k_chol_inv_small<16>, dim3(nbatch), dim3(256), lds, c->stream, A, lda, stride, n, d_info, nbatch);
  return hip_check(c, hipGetLastError(), "k_chol_inv_small");
}

int small_dense_max_n() { return SM_NMAX; }

}  // namespace pw
