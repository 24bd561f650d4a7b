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
// Tiles: a workgroup of WM x WN waves owns 32 WM complex rows (64 WM real rows) x 32 WN
// columns of C; each wave owns 64 real rows x 32 columns as 4 x 2 MFMA 16x16 accumulators.
// K is consumed in stages of 16 (FP32) or 8 (FP64) complex k staged through k-contiguous,
// padded LDS rows in two buffers: the next stage's global loads are issued before the
// current stage's MFMAs and written to the other buffer after them, one barrier per stage.
// Long K (the Ngw-long inner products) is split into chunks; partial tiles are reduced by a
// second kernel in a fixed order, so results are bitwise repeatable.
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

// Workgroup of WM x WN wavefronts; each wavefront owns 64 real rows (32 complex) x 32
// columns of C as 4 x 2 MFMA 16x16 accumulators, so the workgroup tile is
// (32 WM complex rows) x (32 WN columns) while per-lane registers stay fixed.
constexpr int GBK = 16;           // split-K chunks are multiples of this (and of every stage)

// Workgroup id -> (m tile, n tile, k chunk).  Consecutive workgroup ids land on different
// XCDs (round-robin over 8), so the mapping sends every tile of one k chunk (which share
// the chunk of A^H / B rows) to one XCD and its L2.  Without split-K the n tiles of one
// m tile (which share the rows of A) go to one XCD.
__device__ inline void tile_of(int w, int tm, int tn, int nz, int& mt, int& nt, int& z) {
  const int xcd = w & 7, idx = w >> 3;
  if (nz > 1) {
    const int tiles = tm * tn;
    const int zq = idx / tiles, t = idx - zq * tiles;
    z = zq * 8 + xcd;
    mt = t % tm;
    nt = t / tm;
  } else {
    // the tn tiles of one m tile run back to back on one XCD; m tiles rotate over XCDs
    const int mq = idx / tn;
    nt = idx - mq * tn;
    mt = mq * 8 + xcd;
    z = 0;
  }
}

template <typename TC, typename TA, typename TB, int TRANSA, int WM, int WN>
__global__ void __launch_bounds__(64 * WM * WN)
k_zgemm(int M, int N, int K, const cplx<TA>* __restrict__ A, long lda, const cplx<TB>* __restrict__ B,
        long ldb, int kchunk, cplx<double>* __restrict__ C, long ldc, cplx<double> alpha,
        cplx<double> beta, int direct, int use_beta, cplx<double>* __restrict__ part, int tm, int tn,
        int nz) {
  constexpr int NT = 64 * WM * WN;
  constexpr int BMC = 32 * WM;   // complex rows per workgroup
  constexpr int BNC = 32 * WN;   // columns per workgroup
  // complex k per LDS stage: two stages (double buffer) of 16 (FP32) or 8 (FP64) complex k
  // keep the LDS of a 64 x 64 tile at 36 KiB, so four workgroups fit a CU
  constexpr int SBK = sizeof(TC) == 8 ? 8 : 16;
  constexpr int SKS = 2 * SBK + 4;  // LDS row stride (reals): k-contiguous rows, padded
  constexpr int NA = (BMC * SBK) / NT, NB = (BNC * SBK) / NT;
  static_assert(NA >= 1 && NB >= 1 && (BMC * SBK) % NT == 0 && (BNC * SBK) % NT == 0, "tile");
  __shared__ TC As[2][BMC][SKS];   // [buffer][complex row m][real k'] raw (re, im) pairs along k
  __shared__ TC Bs[2][BNC][SKS];   // [buffer][column n][real k']
  typedef typename Mfma<TC>::acc_t acc_t;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;
  int mt, nt, z;
  tile_of(blockIdx.x, tm, tn, nz, mt, nt, z);
  if (z >= nz || mt >= tm || nt >= tn) return;
  const int m0 = mt * BMC, n0 = nt * BNC;
  const int kbeg = z * kchunk;
  const int kend = min(K, kbeg + kchunk);
  acc_t acc[4][2];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[i][j][r] = (TC)0;
  const int li = lane & 15, lg = lane >> 4;

  // global -> register staging: NA complex of A and NB of B per thread per stage, held in
  // their storage type and converted only when written to LDS, so the loads of the next
  // stage stay in flight during the MFMA phase.  Out-of-range elements load a valid
  // address (element 0) and are zeroed at the LDS write through the valid mask.
  cplx<TA> ra[NA];
  cplx<TB> rb[NB];
  unsigned va = 0, vb = 0;
  auto a_idx = [&](int s, int& mm, int& kk) {
    if (TRANSA == 0) { mm = tid % BMC; kk = tid / BMC + (NT / BMC) * s; }
    else { kk = tid % SBK; mm = tid / SBK + (NT / SBK) * s; }
  };
  auto b_idx = [&](int s, int& kb, int& nn) { kb = tid % SBK; nn = tid / SBK + (NT / SBK) * s; };
  // a stage wholly inside the matrices (every stage of an interior tile but the last of its
  // K chunk) is loaded and stored without per-element checks, in workgroups of at most four
  // wavefronts except FP32-MFMA A^H B: measured 2-5 % faster there at the Si64 projector
  // shapes, but 6-10 % slower for FP32-MFMA A^H B and ~10 % slower in the 8-wavefront tiles
  // of the Si216 shapes (the extra path's registers)
  constexpr bool FAST = WM * WN <= 4 && !(sizeof(TC) == sizeof(float) && TRANSA == 1);
  bool sfull = false;
  auto load_stage = [&](int k0) {
    sfull = FAST && m0 + BMC <= M && n0 + BNC <= N && k0 + SBK <= kend;
    if (sfull) {
#pragma unroll
      for (int s = 0; s < NA; ++s) {
        int mm, kk;
        a_idx(s, mm, kk);
        const int gm = m0 + mm, gk = k0 + kk;
        ra[s] = A[TRANSA == 0 ? (long)gk * lda + gm : (long)gm * lda + gk];
      }
#pragma unroll
      for (int s = 0; s < NB; ++s) {
        int kb, nn;
        b_idx(s, kb, nn);
        rb[s] = B[(long)(n0 + nn) * ldb + k0 + kb];
      }
      return;
    }
    va = 0;
    vb = 0;
#pragma unroll
    for (int s = 0; s < NA; ++s) {
      int mm, kk;
      a_idx(s, mm, kk);
      const int gm = m0 + mm, gk = k0 + kk;
      const bool ok = gm < M && gk < kend;
      const long off = ok ? (TRANSA == 0 ? (long)gk * lda + gm : (long)gm * lda + gk) : 0;
      ra[s] = A[off];
      va |= (unsigned)ok << s;
    }
#pragma unroll
    for (int s = 0; s < NB; ++s) {
      int kb, nn;
      b_idx(s, kb, nn);
      const int gkb = k0 + kb, gn = n0 + nn;
      const bool ok = gkb < kend && gn < N;
      rb[s] = B[ok ? (long)gn * ldb + gkb : 0];
      vb |= (unsigned)ok << s;
    }
  };
  auto store_stage = [&](int buf) {
    if (sfull) {
#pragma unroll
      for (int s = 0; s < NA; ++s) {
        int mm, kk;
        a_idx(s, mm, kk);
        *reinterpret_cast<cplx<TC>*>(&As[buf][mm][2 * kk]) = cvt<TC>(ra[s]);
      }
#pragma unroll
      for (int s = 0; s < NB; ++s) {
        int kb, nn;
        b_idx(s, kb, nn);
        *reinterpret_cast<cplx<TC>*>(&Bs[buf][nn][2 * kb]) = cvt<TC>(rb[s]);
      }
      return;
    }
#pragma unroll
    for (int s = 0; s < NA; ++s) {
      int mm, kk;
      a_idx(s, mm, kk);
      const cplx<TC> v = ((va >> s) & 1u) ? cvt<TC>(ra[s]) : mk((TC)0, (TC)0);
      *reinterpret_cast<cplx<TC>*>(&As[buf][mm][2 * kk]) = v;
    }
#pragma unroll
    for (int s = 0; s < NB; ++s) {
      int kb, nn;
      b_idx(s, kb, nn);
      const cplx<TC> v = ((vb >> s) & 1u) ? cvt<TC>(rb[s]) : mk((TC)0, (TC)0);
      *reinterpret_cast<cplx<TC>*>(&Bs[buf][nn][2 * kb]) = v;
    }
  };

  int buf = 0;
  if (kbeg < kend) {
    load_stage(kbeg);
    store_stage(0);
  }
  __syncthreads();
  for (int k0 = kbeg; k0 < kend; k0 += SBK) {
    const bool more = k0 + SBK < kend;
    if (more) load_stage(k0 + SBK);  // in flight during the MFMA phase
    // operand fragments of k-step q: read from LDS one step ahead of the MFMAs that use
    // them (two register sets), so the LDS latency overlaps the previous step's MFMAs
    // Ã's sign pattern depends on the lane and the row tile only (k-step parity is lg's)
    constexpr int NQ = (2 * SBK) / 4;
    TC a[2][4], b[2][2];
    auto frag = [&](int q, int p) {
      const int kp = 4 * q + lg;
      const int kc = kp >> 1, t = kp & 1;
#pragma unroll
      for (int rt = 0; rt < 4; ++rt) {
        const int ip = wm * 64 + rt * 16 + Mfma<TC>::row(li);
        a[p][rt] = As[buf][ip >> 1][2 * kc + ((ip & 1) ^ t)];
      }
#pragma unroll
      for (int ct = 0; ct < 2; ++ct) b[p][ct] = Bs[buf][wn * 32 + ct * 16 + li][kp];
    };
    bool neg[4];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
      const int sgn = (wm * 64 + rt * 16 + Mfma<TC>::row(li)) & 1, t = lg & 1;
      neg[rt] = (TRANSA == 0) ? (sgn == 0 && t == 1) : (sgn == 1 && t == 0);
    }
    frag(0, 0);
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int p = q & 1;
      if (q + 1 < NQ) frag(q + 1, p ^ 1);
      // keep the next step's LDS reads ahead of this step's MFMAs (the scheduler would
      // otherwise sink them to their use), so their latency overlaps these MFMAs
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int rt = 0; rt < 4; ++rt) {
        const TC av = neg[rt] ? -a[p][rt] : a[p][rt];
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) acc[rt][ct] = Mfma<TC>::run(av, b[p][ct], acc[rt][ct]);
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    // one barrier per stage: the other buffer is written while this one is read, and no
    // wave writes this buffer again before every wave has passed this barrier
    if (more) store_stage(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
  // with the row maps above, registers (0, 1) and (2, 3) of lane group lg hold the
  // (re, im) pairs of complex rows 2*lg and 2*lg + 1 of each 16-row tile.
#pragma unroll
  for (int rt = 0; rt < 4; ++rt)
#pragma unroll
    for (int ct = 0; ct < 2; ++ct)
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        const int ip = wm * 64 + rt * 16 + 4 * lg + 2 * p;
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

// FP32-accurate complex GEMM on BF16 MFMA (the mixed path's alternative to the FP32 MFMA
// kernel above).  Every real operand x is split into three BF16 parts x0 + x1 + x2 (x0 =
// bf16(x), x1 = bf16(x - x0), x2 = bf16(x - x0 - x1): 24 significant bits, as FP32), and
// a b = sum of the six products with a combined order above 2^-24 (a0 b0, a0 b1, a1 b0, a0 b2,
// a1 b1, a2 b0), accumulated in FP32 by v_mfma_f32_16x16x32_bf16.  The parts are formed
// once per element when a stage is written to LDS, with the 2x2 real form of op(A) already
// applied (no sign work in the MFMA loop).  Lane l feeds row (or column) l % 16 with the 8
// consecutive real k of slot block l / 16; the accumulator layout equals that of the FP32
// 16x16x4 MFMA, so the tiles, split-K and epilogue are those of k_zgemm.
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
constexpr int HBK = 16;          // complex k per stage = 32 real k = one MFMA k step
constexpr int HKS = 2 * HBK + 8; // LDS row stride in bf16 (16-byte aligned rows, padded)

__device__ inline void split3(float x, __bf16& h0, __bf16& h1, __bf16& h2) {
  h0 = (__bf16)x;
  const float r1 = x - (float)h0;
  h1 = (__bf16)r1;
  h2 = (__bf16)(r1 - (float)h1);
}

template <typename TA, typename TB, int TRANSA, int WM, int WN>
__global__ void __launch_bounds__(64 * WM * WN)
k_zgemm_b3(int M, int N, int K, const cplx<TA>* __restrict__ A, long lda, const cplx<TB>* __restrict__ B,
           long ldb, int kchunk, cplx<double>* __restrict__ C, long ldc, cplx<double> alpha, cplx<double> beta,
           int direct, int use_beta, cplx<double>* __restrict__ part, int tm, int tn, int nz) {
  constexpr int NT = 64 * WM * WN;
  constexpr int BMC = 32 * WM, BNC = 32 * WN;
  constexpr int NA = (BMC * HBK) / NT, NB = (BNC * HBK) / NT;
  static_assert(NA >= 1 && NB >= 1 && (BMC * HBK) % NT == 0 && (BNC * HBK) % NT == 0, "tile");
  __shared__ __attribute__((aligned(16))) __bf16 As[3][2 * BMC][HKS];  // [part][real row][real k]
  __shared__ __attribute__((aligned(16))) __bf16 Bs[3][BNC][HKS];      // [part][column][real k]
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;
  int mt, nt, z;
  tile_of(blockIdx.x, tm, tn, nz, mt, nt, z);
  if (z >= nz || mt >= tm || nt >= tn) return;
  const int m0 = mt * BMC, n0 = nt * BNC;
  const int kbeg = z * kchunk;
  const int kend = min(K, kbeg + kchunk);
  v4f acc[4][2];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[i][j][r] = 0.0f;
  const int li = lane & 15, lg = lane >> 4;

  cplx<TA> ra[NA];
  cplx<TB> rb[NB];
  unsigned va = 0, vb = 0;
  auto a_idx = [&](int s, int& mm, int& kk) {
    if (TRANSA == 0) { mm = tid % BMC; kk = tid / BMC + (NT / BMC) * s; }
    else { kk = tid % HBK; mm = tid / HBK + (NT / HBK) * s; }
  };
  auto b_idx = [&](int s, int& kb, int& nn) { kb = tid % HBK; nn = tid / HBK + (NT / HBK) * s; };
  auto load_stage = [&](int k0) {
    va = 0;
    vb = 0;
#pragma unroll
    for (int s = 0; s < NA; ++s) {
      int mm, kk;
      a_idx(s, mm, kk);
      const int gm = m0 + mm, gk = k0 + kk;
      const bool ok = gm < M && gk < kend;
      ra[s] = A[ok ? (TRANSA == 0 ? (long)gk * lda + gm : (long)gm * lda + gk) : 0];
      va |= (unsigned)ok << s;
    }
#pragma unroll
    for (int s = 0; s < NB; ++s) {
      int kb, nn;
      b_idx(s, kb, nn);
      const int gkb = k0 + kb, gn = n0 + nn;
      const bool ok = gkb < kend && gn < N;
      rb[s] = B[ok ? (long)gn * ldb + gkb : 0];
      vb |= (unsigned)ok << s;
    }
  };
  // 2x2 real form of op(A)[m][k] = (x + i y): rows 2m, 2m+1, columns 2k, 2k+1 hold
  // [[x, -y], [y, x]] (x, y of A[m][k] for op = identity; of conj(A[k][m]) for op = ^H)
  auto store_stage = [&]() {
#pragma unroll
    for (int s = 0; s < NA; ++s) {
      int mm, kk;
      a_idx(s, mm, kk);
      float x = 0.0f, y = 0.0f;
      if ((va >> s) & 1u) {
        x = (float)ra[s].x;
        y = TRANSA == 0 ? (float)ra[s].y : -(float)ra[s].y;
      }
      // row 2m holds (x, -y), row 2m+1 holds (y, x) at columns 2k, 2k+1: one 4-byte write
      // per row and part
      __bf16 p0[2][2], p1[2][2], p2[2][2];
      split3(x, p0[0][0], p1[0][0], p2[0][0]);
      split3(-y, p0[0][1], p1[0][1], p2[0][1]);
      split3(y, p0[1][0], p1[1][0], p2[1][0]);
      p0[1][1] = p0[0][0];
      p1[1][1] = p1[0][0];
      p2[1][1] = p2[0][0];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
        bf16x2_t w0 = {p0[h][0], p0[h][1]}, w1 = {p1[h][0], p1[h][1]}, w2 = {p2[h][0], p2[h][1]};
        *reinterpret_cast<bf16x2_t*>(&As[0][2 * mm + h][2 * kk]) = w0;
        *reinterpret_cast<bf16x2_t*>(&As[1][2 * mm + h][2 * kk]) = w1;
        *reinterpret_cast<bf16x2_t*>(&As[2][2 * mm + h][2 * kk]) = w2;
      }
    }
#pragma unroll
    for (int s = 0; s < NB; ++s) {
      int kb, nn;
      b_idx(s, kb, nn);
      float x = 0.0f, y = 0.0f;
      if ((vb >> s) & 1u) {
        x = (float)rb[s].x;
        y = (float)rb[s].y;
      }
      typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
      __bf16 x0, x1, x2, y0, y1, y2;
      split3(x, x0, x1, x2);
      split3(y, y0, y1, y2);
      const bf16x2_t w0 = {x0, y0}, w1 = {x1, y1}, w2 = {x2, y2};
      *reinterpret_cast<bf16x2_t*>(&Bs[0][nn][2 * kb]) = w0;
      *reinterpret_cast<bf16x2_t*>(&Bs[1][nn][2 * kb]) = w1;
      *reinterpret_cast<bf16x2_t*>(&Bs[2][nn][2 * kb]) = w2;
    }
  };

  if (kbeg < kend) load_stage(kbeg);
  for (int k0 = kbeg; k0 < kend; k0 += HBK) {
    store_stage();
    __syncthreads();
    if (k0 + HBK < kend) load_stage(k0 + HBK);  // in flight during the MFMA phase
    // six products, smallest first; the 8 accumulators interleave so that dependent MFMAs
    // on one accumulator are 8 instructions apart
    constexpr int PA[6] = {2, 1, 0, 1, 0, 0}, PB[6] = {0, 1, 2, 0, 1, 0};
#pragma unroll
    for (int ks = 0; ks < HBK / 16; ++ks) {
      bf16x8_t a[3][4], b[3][2];
#pragma unroll
      for (int p = 0; p < 3; ++p) {
#pragma unroll
        for (int rt = 0; rt < 4; ++rt)
          a[p][rt] = *reinterpret_cast<const bf16x8_t*>(&As[p][wm * 64 + rt * 16 + li][32 * ks + 8 * lg]);
#pragma unroll
        for (int ct = 0; ct < 2; ++ct)
          b[p][ct] = *reinterpret_cast<const bf16x8_t*>(&Bs[p][wn * 32 + ct * 16 + li][32 * ks + 8 * lg]);
      }
#pragma unroll
      for (int pr = 0; pr < 6; ++pr)
#pragma unroll
        for (int rt = 0; rt < 4; ++rt)
#pragma unroll
          for (int ct = 0; ct < 2; ++ct)
            acc[rt][ct] =
                __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[PA[pr]][rt], b[PB[pr]][ct], acc[rt][ct], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int rt = 0; rt < 4; ++rt)
#pragma unroll
    for (int ct = 0; ct < 2; ++ct)
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        const int ip = wm * 64 + rt * 16 + 4 * lg + 2 * p;
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

// Launch one (WM x WN) instance: tiles, split-K chunks, optional ordered reduction.
template <typename TC, typename TA, typename TB, int WM, int WN>
static int zgemm_launch(pw_ctx* c, int transa, int M, int N, int K, cplx<double> alpha, const cplx<TA>* A,
                        long lda, const cplx<TB>* B, long ldb, cplx<double> beta, cplx<double>* C, long ldc,
                        int target_wg) {
  constexpr int BMC = 32 * WM, BNC = 32 * WN;
  const int use_beta = (beta.x != 0.0 || beta.y != 0.0) ? 1 : 0;
  const int tm = (M + BMC - 1) / BMC, tn = (N + BNC - 1) / BNC;
  const long tiles = (long)tm * tn;
  const int ksteps = (K + GBK - 1) / GBK;
  int nz = 1;
  if (tiles < target_wg && ksteps > 8) {
    nz = (int)std::min<long>((target_wg + tiles - 1) / tiles, ksteps / 8);
    nz = std::max(8, (nz + 7) / 8 * 8);  // whole rounds over the 8 XCDs
    // a no-allocation call splits K only as far as its fixed partial-sum array holds
    if (c->no_alloc) nz = std::max<int>(1, std::min<long>(nz, (long)(c->gpart_bytes / ((size_t)M * N * 16))));
  }
  int kchunk = GBK;
  if (K > 0) {
    kchunk = ((ksteps + nz - 1) / nz) * GBK;
    nz = (K + kchunk - 1) / kchunk;
  } else {
    nz = 1;
  }
  cplx<double>* part = nullptr;
  if (nz > 1) {
    int rc = ensure(c, &c->d_gpart, &c->gpart_bytes, (size_t)nz * M * N * sizeof(cplx<double>));
    if (rc) return rc;
    part = reinterpret_cast<cplx<double>*>(c->d_gpart);
  }
  const int direct = nz == 1 ? 1 : 0;
  // ids past the work (the last partial round of 8) return at once
  const long nwg = nz > 1 ? ((nz + 7) / 8) * 8 * tiles : (long)8 * ((tm + 7) / 8) * tn;
  const double gbytes = (double)M * K * sizeof(cplx<TA>) + (double)K * N * sizeof(cplx<TB>) +
                        (double)M * N * 16.0 * (use_beta ? 2 : 1);
  ProfScope ps(c, transa ? PW_PROF_GEMM_CN : PW_PROF_GEMM_NN, gbytes);
  bool split_bf16 = false;
  // the three-part operands of a 128 x 128 tile do not fit in LDS: that shape keeps FP32 MFMA
  // default (-1): by size -- the split form pays off on large products (Si216's projector and
  // Gram products: 4e10 complex multiply-adds, 3.49 -> 3.00 ms), not beside the FFT passes
  // at Si64 (1.2e9)
  if constexpr (sizeof(TC) == 4 && WM * WN < 16)
    split_bf16 = c->gemm_split_bf16 > 0 || (c->gemm_split_bf16 < 0 && (double)M * N * K >= 8589934592.0);
  if constexpr (sizeof(TC) == 4 && WM * WN < 16) {
    if (split_bf16) {
      if (transa)
        hipLaunchKernelGGL((k_zgemm_b3<TA, TB, 1, WM, WN>), dim3(nwg), dim3(64 * WM * WN), 0, c->stream, M, N, K, A,
                           lda, B, ldb, kchunk, C, ldc, alpha, beta, direct, use_beta, part, tm, tn, nz);
      else
        hipLaunchKernelGGL((k_zgemm_b3<TA, TB, 0, WM, WN>), dim3(nwg), dim3(64 * WM * WN), 0, c->stream, M, N, K, A,
                           lda, B, ldb, kchunk, C, ldc, alpha, beta, direct, use_beta, part, tm, tn, nz);
    }
  }
  if (split_bf16) {
    // launched above
  } else if (transa)
    hipLaunchKernelGGL((k_zgemm<TC, TA, TB, 1, WM, WN>), dim3(nwg), dim3(64 * WM * WN), 0, c->stream, M, N, K, A,
                       lda, B, ldb, kchunk, C, ldc, alpha, beta, direct, use_beta, part, tm, tn, nz);
  else
    hipLaunchKernelGGL((k_zgemm<TC, TA, TB, 0, WM, WN>), dim3(nwg), dim3(64 * WM * WN), 0, c->stream, M, N, K, A,
                       lda, B, ldb, kchunk, C, ldc, alpha, beta, direct, use_beta, part, tm, tn, nz);
  if (nz > 1) {
    const long tot = (long)M * N;
    const int blocks = (int)std::min<long>((tot + 255) / 256, 2048);
    hipLaunchKernelGGL(k_zgemm_reduce, dim3(blocks), dim3(256), 0, c->stream, M, N, nz, part, C, ldc,
                       alpha, beta, use_beta);
  }
  return hip_check(c, hipGetLastError(), "zgemm");
}

// Tile choice: big workgroup tiles cut the re-reads of A and B from HBM (each element is
// read once per tile along the other dimension); small problems keep 64 x 64 tiles so the
// grid still fills the chip.
template <typename TC, typename TA, typename TB>
static int zgemm_t(pw_ctx* c, int transa, int M, int N, int K, cplx<double> alpha, const cplx<TA>* A,
                   long lda, const cplx<TB>* B, long ldb, cplx<double> beta, cplx<double>* C, long ldc) {
  if (M <= 0 || N <= 0) return PW_OK;
  // tuning knobs (pw_tune_set): tile shape and the split-K target workgroup count
  const int slot = transa ? 0 : 1;
  int shape = c->tune_gemm_shape[slot];
  if (shape < 0) {
    // 32-column tiles when they cover N with less padding than 64-column ones (N = 160
    // bands: 160 columns instead of 192, a fifth of the MFMA work saved).  Measured on the
    // Si64 projector products of op_H, which run beside the FFT passes: 64 x 32 for A^H B
    // (split-K fills the chip); for A B 128 x 32 in FP64, while in FP32 the 64 x 64 tile
    // (less LDS per workgroup) shares the CUs with the FFT kernels better
    // With many columns (N >= 256: Si216's 480 bands) the 128 x 64 tile halves the re-reads
    // of the other operand and wins for both products (FP64 A^H B 8.2 -> 6.4 ms, A B 7.5 ->
    // 5.8 ms at Ngw = 80 797, Nbeta = 1080).
    const int pad64 = (N + 63) / 64 * 64, pad32 = (N + 31) / 32 * 32;
    if (N >= 256) shape = 1;
    else if (pad32 < pad64) shape = transa ? 5 : (sizeof(TC) == 8 ? 4 : 0);
    else shape = 0;
  }
  // A B with a short K splits only when the tiles leave most of the chip idle; A^H B with
  // many columns splits K over more workgroups
  const int target = c->tune_gemm_wg[slot] > 0 ? c->tune_gemm_wg[slot] : (transa ? (N >= 256 ? 2048 : 1024) : 512);
  switch (shape) {
    case 1: return zgemm_launch<TC, TA, TB, 4, 2>(c, transa, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, target);
    case 4: return zgemm_launch<TC, TA, TB, 4, 1>(c, transa, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, target);
    case 5: return zgemm_launch<TC, TA, TB, 2, 1>(c, transa, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, target);
    case 2: return zgemm_launch<TC, TA, TB, 2, 4>(c, transa, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, target);
    case 3: return zgemm_launch<TC, TA, TB, 4, 4>(c, transa, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, target);
    default: return zgemm_launch<TC, TA, TB, 2, 2>(c, transa, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, target);
  }
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
