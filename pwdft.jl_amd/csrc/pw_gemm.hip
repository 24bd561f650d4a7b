// Complex GEMM on real MFMA for the projector products of op_V_Ps_nloc
// (calc_betaNL_psi, PsPotNL.jl:242-254; beta * (D * proj), op_V_Ps_nloc.jl:59-76) and the
// Rayleigh-Ritz / Gram products of diag_LOBPCG (diag_LOBPCG.jl:85-203) and diag_davidson.
//
//   C (M x N) = alpha * op(A) * B + beta * C,  op(A) = A (M x K) or A^H (A is K x M)
//
// The complex product is the real product  C_raw (2M x N) = At (2M x 2K) * B_raw (2K x N)
// where B_raw / C_raw are the interleaved (re, im) storage of B / C and At replaces each
// complex entry a of op(A) by the 2x2 block [[re a, -im a], [im a, re a]].  Every element of
// At is a signed copy of one element of A, so one MFMA stream computes 8 real flops per
// complex multiply-add with no wasted lanes.
//
// Tiles: a 256-thread workgroup owns 32 complex rows (64 real rows) x 64 columns of C; each
// of its 4 waves owns 32 real rows x 32 columns as 2 x 2 MFMA 16x16 accumulators.  K is
// consumed in steps of 16 complex (32 real) staged through LDS.  Long K (the Ngw-long
// inner products) is split over blockIdx.z; partial tiles are reduced by a second kernel
// in a fixed order, so results are bitwise repeatable.
#include "pw_internal.h"

namespace pw {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename TC>
struct Mfma;
template <>
struct Mfma<double> {
  typedef v4d acc_t;
  __device__ static inline acc_t run(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // f64 16x16x4 leaves D[(l/16) + 4r][l%16] in register r of lane l; A-row i is fed with
  // logical real row 2*(2*(i%4) + (i/4)/2) + (i/4)%2 so that registers (0,1) and (2,3) of
  // lane group g hold the (re, im) pairs of complex rows 2g and 2g+1.
  __device__ static inline int row(int i) { return 2 * (2 * (i & 3) + ((i >> 2) >> 1)) + ((i >> 2) & 1); }
};
template <>
struct Mfma<float> {
  typedef v4f acc_t;
  __device__ static inline acc_t run(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // f32 16x16x4 leaves D[4*(l/16) + r][l%16] in register r of lane l: identity row map.
  __device__ static inline int row(int i) { return i; }
};

constexpr int GBM = 32;  // complex rows of C per workgroup
constexpr int GBN = 64;  // columns of C per workgroup
constexpr int GBK = 16;  // complex k per LDS stage
constexpr int GPAD = 4;

template <typename TC, typename TA, typename TB, int TRANSA>
__global__ void __launch_bounds__(256)
k_zgemm(int M, int N, int K, const cplx<TA>* __restrict__ A, long lda, const cplx<TB>* __restrict__ B,
        long ldb, int kchunk, cplx<double>* __restrict__ C, long ldc, cplx<double> alpha,
        cplx<double> beta, int direct, int use_beta, cplx<double>* __restrict__ part) {
  __shared__ TC As[GBK][2 * GBM + GPAD];
  __shared__ TC Bs[2 * GBK][GBN + GPAD];
  typedef typename Mfma<TC>::acc_t acc_t;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int m0 = blockIdx.x * GBM, n0 = blockIdx.y * GBN, z = blockIdx.z;
  const int kbeg = z * kchunk;
  const int kend = min(K, kbeg + kchunk);
  acc_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[i][j][r] = (TC)0;

  const int li = lane & 15, lg = lane >> 4;
  for (int k0 = kbeg; k0 < kend; k0 += GBK) {
#pragma unroll
    for (int s = 0; s < (GBM * GBK) / 256; ++s) {
      const int e = tid + 256 * s;
      int mm, kk;
      if (TRANSA == 0) { mm = e % GBM; kk = e / GBM; }
      else { kk = e % GBK; mm = e / GBK; }
      const int gm = m0 + mm, gk = k0 + kk;
      cplx<TC> v = mk((TC)0, (TC)0);
      if (gm < M && gk < kend) v = cvt<TC>(TRANSA == 0 ? A[(long)gk * lda + gm] : A[(long)gm * lda + gk]);
      As[kk][2 * mm] = v.x;
      As[kk][2 * mm + 1] = v.y;
    }
#pragma unroll
    for (int s = 0; s < (GBN * GBK) / 256; ++s) {
      const int e = tid + 256 * s;
      const int kk = e % GBK, nn = e / GBK;
      const int gk = k0 + kk, gn = n0 + nn;
      cplx<TC> v = mk((TC)0, (TC)0);
      if (gk < kend && gn < N) v = cvt<TC>(B[(long)gn * ldb + gk]);
      Bs[2 * kk][nn] = v.x;
      Bs[2 * kk + 1][nn] = v.y;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < (2 * GBK) / 4; ++q) {
      const int kp = 4 * q + lg;
      const int kc = kp >> 1, t = kp & 1;
      TC a[2], b[2];
#pragma unroll
      for (int rt = 0; rt < 2; ++rt) {
        const int ip = wm * 32 + rt * 16 + Mfma<TC>::row(li);
        const int s = ip & 1, ml = ip >> 1;
        TC av = As[kc][2 * ml + (s ^ t)];
        const bool neg = (TRANSA == 0) ? (s == 0 && t == 1) : (s == 1 && t == 0);
        a[rt] = neg ? -av : av;
      }
#pragma unroll
      for (int ct = 0; ct < 2; ++ct) b[ct] = Bs[kp][wn * 32 + ct * 16 + li];
#pragma unroll
      for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) acc[rt][ct] = Mfma<TC>::run(a[rt], b[ct], acc[rt][ct]);
    }
    __syncthreads();
  }
  // with the row maps above, registers (0, 1) and (2, 3) of lane group lg hold the
  // (re, im) pairs of complex rows 2*lg and 2*lg + 1 of each 16-row tile.
#pragma unroll
  for (int rt = 0; rt < 2; ++rt)
#pragma unroll
    for (int ct = 0; ct < 2; ++ct)
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        const int ip = wm * 32 + rt * 16 + 4 * lg + 2 * p;
        const int m = m0 + (ip >> 1);
        const int n = n0 + wn * 32 + ct * 16 + li;
        if (m < M && n < N) {
          cplx<double> v = mk((double)acc[rt][ct][2 * p], (double)acc[rt][ct][2 * p + 1]);
          if (direct) {
            cplx<double> r = cmul(alpha, v);
            if (use_beta) r = r + cmul(beta, C[(long)n * ldc + m]);
            C[(long)n * ldc + m] = r;
          } else {
            part[((long)z * N + n) * M + m] = v;
          }
        }
      }
}

__global__ void k_zgemm_reduce(int M, int N, int nz, const cplx<double>* __restrict__ part,
                               cplx<double>* __restrict__ C, long ldc, cplx<double> alpha,
                               cplx<double> beta, int use_beta) {
  const long tot = (long)M * N;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < tot; t += (long)gridDim.x * blockDim.x) {
    const int m = (int)(t % M), n = (int)(t / M);
    cplx<double> s = mk(0.0, 0.0);
    for (int z = 0; z < nz; ++z) s = s + part[((long)z * N + n) * M + m];
    cplx<double> r = cmul(alpha, s);
    if (use_beta) r = r + cmul(beta, C[(long)n * ldc + m]);
    C[(long)n * ldc + m] = r;
  }
}

template <typename TC, typename TA, typename TB>
static int zgemm_t(pw_ctx* c, int transa, int M, int N, int K, cplx<double> alpha, const cplx<TA>* A,
                   long lda, const cplx<TB>* B, long ldb, cplx<double> beta, cplx<double>* C, long ldc) {
  if (M <= 0 || N <= 0) return PW_OK;
  const int use_beta = (beta.x != 0.0 || beta.y != 0.0) ? 1 : 0;
  const int tm = (M + GBM - 1) / GBM, tn = (N + GBN - 1) / GBN;
  const long tiles = (long)tm * tn;
  int nz = 1;
  const int ksteps = (K + GBK - 1) / GBK;
  if (tiles < 1024 && ksteps > 8) {
    nz = (int)std::min<long>((1024 + tiles - 1) / tiles, ksteps / 8);
    if (nz < 1) nz = 1;
  }
  int kchunk = ((ksteps + nz - 1) / nz) * GBK;
  if (K <= 0) {
    kchunk = GBK;
    nz = 1;
  } else {
    nz = (K + kchunk - 1) / kchunk;
  }
  cplx<double>* part = nullptr;
  if (nz > 1) {
    int rc = ensure(c, &c->d_gpart, &c->gpart_bytes, (size_t)nz * M * N * sizeof(cplx<double>));
    if (rc) return rc;
    part = reinterpret_cast<cplx<double>*>(c->d_gpart);
  }
  dim3 grid(tm, tn, nz);
  const int direct = nz == 1 ? 1 : 0;
  // algorithmic bytes: A and B read once, C written (and read when beta != 0)
  const double gbytes = (double)M * K * sizeof(cplx<TA>) + (double)K * N * sizeof(cplx<TB>) +
                        (double)M * N * 16.0 * (use_beta ? 2 : 1);
  ProfScope ps(c, transa ? PW_PROF_GEMM_CN : PW_PROF_GEMM_NN, gbytes);
  if (transa)
    hipLaunchKernelGGL((k_zgemm<TC, TA, TB, 1>), grid, dim3(256), 0, c->stream, M, N, K, A, lda, B, ldb,
                       kchunk, C, ldc, alpha, beta, direct, use_beta, part);
  else
    hipLaunchKernelGGL((k_zgemm<TC, TA, TB, 0>), grid, dim3(256), 0, c->stream, M, N, K, A, lda, B, ldb,
                       kchunk, C, ldc, alpha, beta, direct, use_beta, part);
  if (nz > 1) {
    const long tot = (long)M * N;
    const int blocks = (int)std::min<long>((tot + 255) / 256, 2048);
    hipLaunchKernelGGL(k_zgemm_reduce, dim3(blocks), dim3(256), 0, c->stream, M, N, nz, part, C, ldc,
                       alpha, beta, use_beta);
  }
  return hip_check(c, hipGetLastError(), "zgemm");
}

int zgemm(pw_ctx* c, int transa, int M, int N, int K, cplx<double> alpha, const cplx<double>* A, long lda,
          const cplx<double>* B, long ldb, cplx<double> beta, cplx<double>* C, long ldc, int prec) {
  if (prec) return zgemm_t<float, double, double>(c, transa, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
  return zgemm_t<double, double, double>(c, transa, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
}

int zgemm_f32in(pw_ctx* c, int transa, int M, int N, int K, cplx<double> alpha, const cplx<float>* A,
                long lda, const cplx<double>* B, long ldb, cplx<double> beta, cplx<double>* C, long ldc,
                int prec) {
  (void)prec;
  return zgemm_t<float, float, double>(c, transa, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
}

}  // namespace pw
