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
//    rotated while |A_pq| > max(1e-15 sqrt|A_pp A_qq|, 2 eps ||A||_F), until a sweep rotates
//    no pair; eigenvalues sorted ascending with their vectors.  Complex 2x2 step: the pair (p, q) with A_pq = |g| e^{i phi} is made
//    real by diag(1, e^{-i phi}) and diagonalised by the classical rotation
//    t = sign(tau) / (|tau| + sqrt(1 + tau^2)), tau = (A_qq - A_pp) / (2 |g|).
//  - k_chol_inv_small: U = chol(Hermitian(A)).U by right-looking elimination, then U^-1 by
//    column-parallel back substitution; the strict lower triangle zeroed.  Status: 0, or
//    the index (1-based) of the first non-positive pivot.
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
  __shared__ int flag;
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
      // A <- J^H A J, block (rows of pair k, columns of pair k2); k = lane mod 32 and k2 stepping
      // by SM_NT / 32 (npair <= 32), no run-time division
      static_assert(SM_NMAX / 2 <= 32 && SM_NMAX <= 64, "lane maps of the round updates");
      for (int k2 = lane >> 5, k = lane & 31; k < npair && k2 < npair; k2 += SM_NT / 32) {
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
      // V <- V J, columns of pair k2; row i = lane mod 64 (n <= 64)
      for (int k2 = lane >> 6, i = lane & 63; i < n && k2 < npair; k2 += SM_NT / 64) {
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
    __syncthreads();
    const int more = flag;
    __syncthreads();
    if (!more) {
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
__global__ void __launch_bounds__(SM_NT) k_chol_inv_small(cplx<double>* __restrict__ A, long lda, long stride, int n,
                                                          int* __restrict__ info, int nbatch) {
  extern __shared__ char smem_raw[];
  cplx<double>* u = reinterpret_cast<cplx<double>*>(smem_raw);   // n x n
  cplx<double>* x = u + n * n;                                     // n x n
  __shared__ int bad;
  cplx<double>* M = A + (long)blockIdx.x * stride;
  const int tid = threadIdx.x;
  const int nn = n * n;
  for (int t = tid; t < nn; t += SM_NT) {
    const int i = t % n, j = t / n;
    u[i + n * j] = i <= j ? (i == j ? mk(M[(long)j * lda + i].x, 0.0) : M[(long)j * lda + i]) : mk(0.0, 0.0);
  }
  if (tid == 0) bad = 0;
  __syncthreads();
  // right-looking: row k of U, then the trailing upper triangle
  for (int k = 0; k < n; ++k) {
    const double dkk = u[k + n * k].x;
    if (!(dkk > 0.0)) {
      if (tid == 0 && bad == 0) bad = k + 1;
      __syncthreads();
      break;
    }
    const double d = sqrt(dkk), rd = 1.0 / d;
    __syncthreads();
    for (int j = k + tid; j < n; j += SM_NT) u[k + n * j] = j == k ? mk(d, 0.0) : scal(rd, u[k + n * j]);
    __syncthreads();
    // A_ij -= conj(U_ki) U_kj for k < i <= j, threads on a fixed 16 x 16 grid of (i, j)
    static_assert(SM_NT == 256, "16 x 16 update grid");
    for (int j = k + 1 + (tid >> 4); j < n; j += 16) {
      const cplx<double> ukj = u[k + n * j];
      for (int i = k + 1 + (tid & 15); i <= j; i += 16) u[i + n * j] = u[i + n * j] - cmulc(ukj, u[k + n * i]);
    }
    __syncthreads();
  }
  const int b = blockIdx.x;
  if (bad) {
    if (tid == 0) {
      info[b] = bad;
      info[nbatch + b] = 0;
    }
    return;
  }
  // X = U^-1: U X[:, j] = e_j by back substitution, column j on the four lanes 4 j .. 4 j + 3
  // of one wavefront (n <= SM_NT / 4): each lane sums every fourth term of a row's dot
  // product, the four partial sums are combined by lane shuffles, and every lane of the
  // group writes the (equal) result, so each lane's later reads follow its own store
  static_assert(SM_NMAX * 4 <= SM_NT, "four lanes per column");
  {
    const int j = tid >> 2, q = tid & 3;
    const bool col = j < n;
    for (int i = n - 1; i >= 0; --i) {
      cplx<double> s = mk(0.0, 0.0);
      if (col && i < j)
        for (int k = i + 1 + q; k <= j; k += 4) s = s + cmul(u[i + n * k], x[k + n * j]);
      s.x += __shfl_xor(s.x, 1, 64);
      s.y += __shfl_xor(s.y, 1, 64);
      s.x += __shfl_xor(s.x, 2, 64);
      s.y += __shfl_xor(s.y, 2, 64);
      if (col) x[i + n * j] = i <= j ? scal(1.0 / u[i + n * i].x, mk((i == j ? 1.0 : 0.0) - s.x, -s.y)) : mk(0.0, 0.0);
    }
  }
  __syncthreads();
  for (int t = tid; t < nn; t += SM_NT) {
    const int i = t % n, j = t / n;
    M[(long)j * lda + i] = x[i + n * j];
  }
  if (tid == 0) {
    info[b] = 0;
    info[nbatch + b] = 0;
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
  const size_t lds = 2 * (size_t)n * n * sizeof(cplx<double>);
  hipLaunchKernelGGL(k_chol_inv_small, dim3(nbatch), dim3(SM_NT), lds, c->stream, A, lda, stride, n, d_info, nbatch);
  return hip_check(c, hipGetLastError(), "k_chol_inv_small");
}

int small_dense_max_n() { return SM_NMAX; }

}  // namespace pw
