// Sphere-pruned batched 3-D FFT for op_V_loc (op_V_loc.jl:33-92) and calc_rhoe
// (calc_rhoe.jl:51-78), with the G-sphere scatter/gather, the local-potential multiply,
// the kinetic term (op_K.jl:55-75) and the density accumulation fused in.
//
// Grid layout: x fastest, ip = i + n1*(j + n2*k).  A band's wave function occupies a
// sphere of radius sqrt(2 ecutwfc), half the grid's radius, so only ~1/5 of the
// (x, y) columns and ~1/2 of the x values are ever non-zero on input or read on output.
// Three kernels per batch of bands:
//   1. col_ifft  : scatter sphere entries into their z-columns (LDS), inverse z-FFT,
//                  write W[b][z][col]                      (work array, in place below)
//   2. plane_vloc: per (z-plane, band): inverse y-FFT on active x lines, inverse x-FFT on
//                  all rows, * V(r)/N, forward x-FFT, forward y-FFT on active x lines,
//                  write W[b][z][col] back                 (whole plane in LDS)
//   3. col_fft   : forward z-FFT of each column, gather the sphere entries, add the
//                  kinetic term and write/accumulate Hpsi.
// For calc_rhoe, kernel 2 is replaced by plane_rhoe (inverse y/x FFT, |.|^2 weighted
// accumulation over a group of bands in registers) and a deterministic group reduction.
//
// Every 1-D FFT runs as Stockham radix passes (radices 2,3,4,5,8,9) on lines held in LDS;
// each thread keeps its butterflies in registers across the pass barrier, so a pass is
// done in place.  Backward (G->R) uses exp(+i), forward exp(-i), as FFTW's
// plan_ifft!/plan_fft! (wrappers_fft.jl:42-62); the 1/N of plan_ifft! is folded into V.
#include "pw_internal.h"

namespace pw {

// Launch shapes: column kernels run 512 threads holding up to COL_ME elements each across
// a pass barrier; plane kernels run 1024 threads (16 waves on one CU, whose LDS the plane
// fills) holding up to PLANE_ME elements each.
#ifndef PW_COL_NT
#define PW_COL_NT 512
#endif
#ifndef PW_PLANE_NT
#define PW_PLANE_NT 512
#endif
#ifndef PW_PLANE_NT64
#define PW_PLANE_NT64 1024
#endif
// Launch shapes: workgroups of COL_NT / PLANE_NT threads; a lane holds at most FFT_ME
// elements of a line across one radix pass (FftDim::L lanes per line).
constexpr int COL_NT = PW_COL_NT;
constexpr int PLANE_NT = PW_PLANE_NT;  // FP32 planes (~41 KiB of LDS at 72^2): several workgroups per CU
constexpr int PLANE_NT64 = PW_PLANE_NT64;  // FP64 planes fill one CU's LDS: one wide workgroup
template <typename T>
struct PlaneNT {
  static constexpr int value = sizeof(T) == 8 ? PLANE_NT64 : PLANE_NT;
};
constexpr int FFT_ME = 12;

// ----------------------------------------------------------------------------- butterflies

template <typename T, int S, int R>
struct DFT;

template <typename T, int S>
struct DFT<T, S, 2> {
  __device__ static inline void run(cplx<T>* v) {
    cplx<T> a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
  }
};

template <typename T, int S>
struct DFT<T, S, 3> {
  __device__ static inline void run(cplx<T>* v) {
    const T c = (T)-0.5;
    const T d = (T)(S * 0.86602540378443864676);
    cplx<T> t1 = v[1] + v[2], t2 = v[1] - v[2];
    cplx<T> m = v[0] + scal(c, t1);
    cplx<T> q = mk(-d * t2.y, d * t2.x);  // i*d*t2
    v[0] = v[0] + t1;
    v[1] = m + q;
    v[2] = m - q;
  }
};

template <typename T, int S>
struct DFT<T, S, 4> {
  __device__ static inline void run(cplx<T>* v) {
    cplx<T> s02 = v[0] + v[2], d02 = v[0] - v[2];
    cplx<T> s13 = v[1] + v[3], d13 = mul_i<S>(v[1] - v[3]);
    v[0] = s02 + s13;
    v[2] = s02 - s13;
    v[1] = d02 + d13;
    v[3] = d02 - d13;
  }
};

template <typename T, int S>
struct DFT<T, S, 5> {
  __device__ static inline void run(cplx<T>* v) {
    const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
    const T s1 = (T)(S * 0.95105651629515357212), s2 = (T)(S * 0.58778525229247312917);
    cplx<T> t1 = v[1] + v[4], t2 = v[2] + v[3], t3 = v[1] - v[4], t4 = v[2] - v[3];
    cplx<T> a1 = v[0] + scal(c1, t1) + scal(c2, t2);
    cplx<T> a2 = v[0] + scal(c2, t1) + scal(c1, t2);
    cplx<T> b1 = scal(s1, t3) + scal(s2, t4);
    cplx<T> b2 = scal(s2, t3) - scal(s1, t4);
    cplx<T> ib1 = mk(-b1.y, b1.x), ib2 = mk(-b2.y, b2.x);
    v[0] = v[0] + t1 + t2;
    v[1] = a1 + ib1;
    v[4] = a1 - ib1;
    v[2] = a2 + ib2;
    v[3] = a2 - ib2;
  }
};

template <typename T, int S>
struct DFT<T, S, 8> {
  __device__ static inline void run(cplx<T>* v) {
    cplx<T> e[4] = {v[0], v[2], v[4], v[6]};
    cplx<T> o[4] = {v[1], v[3], v[5], v[7]};
    DFT<T, S, 4>::run(e);
    DFT<T, S, 4>::run(o);
    const T h = (T)0.70710678118654752440;
    // w^1 = (1 + S i)/sqrt2, w^2 = S i, w^3 = (-1 + S i)/sqrt2
    cplx<T> o1 = mk(h * (o[1].x - S * o[1].y), h * (o[1].y + S * o[1].x));
    cplx<T> o2 = mul_i<S>(o[2]);
    cplx<T> o3 = mk(h * (-o[3].x - S * o[3].y), h * (-o[3].y + S * o[3].x));
    v[0] = e[0] + o[0];
    v[4] = e[0] - o[0];
    v[1] = e[1] + o1;
    v[5] = e[1] - o1;
    v[2] = e[2] + o2;
    v[6] = e[2] - o2;
    v[3] = e[3] + o3;
    v[7] = e[3] - o3;
  }
};

template <typename T, int S>
struct DFT<T, S, 9> {
  __device__ static inline void run(cplx<T>* v) {
    // 9 = 3 x 3: inner DFT3 over p (r = 3p + q), twiddle w9^(q k1), outer DFT3 over q.
    cplx<T> y[3][3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      cplx<T> t[3] = {v[q], v[3 + q], v[6 + q]};
      DFT<T, S, 3>::run(t);
      y[q][0] = t[0];
      y[q][1] = t[1];
      y[q][2] = t[2];
    }
    const T c1 = (T)0.76604444311897803520, s1 = (T)(S * 0.64278760968653932632);  // w9^1
    const T c2 = (T)0.17364817766693034885, s2 = (T)(S * 0.98480775301220805936);  // w9^2
    const T c4 = (T)-0.93969262078590838405, s4 = (T)(S * 0.34202014332566873304); // w9^4
    y[1][1] = cmul(y[1][1], mk(c1, s1));
    y[1][2] = cmul(y[1][2], mk(c2, s2));
    y[2][1] = cmul(y[2][1], mk(c2, s2));
    y[2][2] = cmul(y[2][2], mk(c4, s4));
#pragma unroll
    for (int k1 = 0; k1 < 3; ++k1) {
      cplx<T> t[3] = {y[0][k1], y[1][k1], y[2][k1]};
      DFT<T, S, 3>::run(t);
      v[k1] = t[0];
      v[k1 + 3] = t[1];
      v[k1 + 6] = t[2];
    }
  }
};

// Wave-local Stockham FFT.  One line (length n, element r at lds[base + r*es]) is owned by
// a group of d.L lanes of one wavefront; a wavefront holds d.gpw groups.  Lane q of a group
// takes butterflies q, q + L, ... of each radix pass, so no workgroup barrier is needed
// between passes: the LDS operations of one wavefront complete in issue order, and the
// wavefront fence keeps the compiler from moving the next pass's reads above this pass's
// writes.  In-place: every lane reads all of its butterflies before any lane writes.
// Ns = product of the radices already applied; tw[t] = exp(+2 pi i t / n).
__device__ inline void wave_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDS pointers keep their address space across the (non-inlined) pass so that accesses
// compile to ds_read/ds_write rather than flat instructions.
#if defined(__HIP_DEVICE_COMPILE__)
#define LDS_AS __attribute__((address_space(3)))
#else
#define LDS_AS
#endif

template <typename T, int S, int R, int ME>
__device__ __attribute__((noinline)) void wave_pass(LDS_AS cplx<T>* lds, int base, int es, int n, int Ns,
                                                    const LDS_AS cplx<T>* tw, int L, int q, bool active,
                                                    const T* __restrict__ vpre, bool usev) {
  constexpr int MB = (ME / R) > 0 ? (ME / R) : 1;
  const int nbl = n / R;
  const int twstep = n / (Ns * R);
  cplx<T> v[MB][R];
  int dst[MB];
#pragma unroll
  for (int s = 0; s < MB; ++s) {
    const int j = q + s * L;
    dst[s] = -1;
    if (active && j < nbl) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        v[s][r] = lds[base + (j + r * nbl) * es];
        if (usev) v[s][r] = scal(vpre[j + r * nbl], v[s][r]);
      }
      const int jm = j % Ns;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const cplx<T> w = tw[jm * r * twstep];
        v[s][r] = S > 0 ? cmul(v[s][r], w) : cmulc(v[s][r], w);
      }
      DFT<T, S, R>::run(v[s]);
      dst[s] = base + ((j - jm) * R + jm) * es;
    }
  }
  wave_fence();
#pragma unroll
  for (int s = 0; s < MB; ++s) {
    if (dst[s] >= 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) lds[dst[s] + r * Ns * es] = v[s][r];
    }
  }
  wave_fence();
}

// Compile-time specialised line FFT: the same Stockham passes as wave_pass with n, the
// radices, Ns and the lane count known to the compiler, so index arithmetic folds to
// constants and the first pass (all twiddles = 1) does no twiddle multiply.
template <typename T, int S, int N, int NS, int R, int L>
__device__ inline void spass(LDS_AS cplx<T>* lds, int base, int es, const LDS_AS cplx<T>* tw, int q,
                             bool active, const T* __restrict__ vpre, bool usev) {
  constexpr int NBL = N / R;
  constexpr int MB = (NBL + L - 1) / L;
  constexpr int TWS = N / (NS * R);
  cplx<T> v[MB][R];
  int dst[MB];
#pragma unroll
  for (int s = 0; s < MB; ++s) {
    const int j = q + s * L;
    dst[s] = -1;
    if (active && (MB * L == NBL || j < NBL)) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        v[s][r] = lds[base + (j + r * NBL) * es];
        if (usev) v[s][r] = scal(vpre[j + r * NBL], v[s][r]);
      }
      const int jm = NS == 1 ? 0 : j % NS;
      if (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          const cplx<T> w = tw[jm * r * TWS];
          v[s][r] = S > 0 ? cmul(v[s][r], w) : cmulc(v[s][r], w);
        }
      }
      DFT<T, S, R>::run(v[s]);
      dst[s] = base + ((j - jm) * R + jm) * es;
    }
  }
  wave_fence();
#pragma unroll
  for (int s = 0; s < MB; ++s) {
    if (dst[s] >= 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) lds[dst[s] + r * NS * es] = v[s][r];
    }
  }
  wave_fence();
}

template <typename T, int S, int N, int R0, int R1, int R2>
__device__ __attribute__((noinline)) void line_static(LDS_AS cplx<T>* lds, int base, int es,
                                                      const LDS_AS cplx<T>* tw, int q, bool active,
                                                      const T* __restrict__ vpre, bool usev) {
  constexpr int L = lanes_for(N, R0, R1, R2, 12);
  spass<T, S, N, 1, R0, L>(lds, base, es, tw, q, active, vpre, usev);
  spass<T, S, N, R0, R1, L>(lds, base, es, tw, q, active, vpre, false);
  if constexpr (R2 > 1) spass<T, S, N, R0 * R1, R2, L>(lds, base, es, tw, q, active, vpre, false);
}

// Radix of pass p from the radices packed 4 bits each (every radix is at most 9): indexing
// the by-value FftDim argument with a run-time p would copy the struct into per-lane scratch
// memory at kernel entry (and the compiler turns selects between constant-index loads back
// into such an index).
__device__ inline int pass_radix(const FftDim& d, int p) {
  unsigned code = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) code |= (unsigned)d.radix[i] << (4 * i);
  return (int)((code >> (4 * p)) & 15u);
}

// One line through all radix passes.  vpre (length n, or null) multiplies the input
// element-wise before the first pass.
template <typename T, int S, int ME>
__device__ inline void wave_line_fft(cplx<T>* lds, int base, int es, const FftDim& d,
                                     const cplx<T>* __restrict__ tw, int q, bool active,
                                     const T* vpre) {
  {
    LDS_AS cplx<T>* l3 = (LDS_AS cplx<T>*)lds;
    const LDS_AS cplx<T>* t3 = (const LDS_AS cplx<T>*)tw;
    const bool uv = vpre != nullptr;
    switch (d.plan) {
      case 72: line_static<T, S, 72, 9, 8, 1>(l3, base, es, t3, q, active, vpre, uv); return;
      case 27: line_static<T, S, 27, 9, 3, 1>(l3, base, es, t3, q, active, vpre, uv); return;
      case 45: line_static<T, S, 45, 9, 5, 1>(l3, base, es, t3, q, active, vpre, uv); return;
      case 20: line_static<T, S, 20, 5, 4, 1>(l3, base, es, t3, q, active, vpre, uv); return;
      case 108: line_static<T, S, 108, 9, 4, 3>(l3, base, es, t3, q, active, vpre, uv); return;
      default: break;
    }
  }
  int Ns = 1;
  for (int p = 0; p < d.npass; ++p) {
    const int R = pass_radix(d, p);
    // global V row for the first pass only; the flag travels separately from the pointer
    const bool usev = (p == 0) && (vpre != nullptr);
    const T* vp = vpre;
    LDS_AS cplx<T>* l3 = (LDS_AS cplx<T>*)lds;
    const LDS_AS cplx<T>* t3 = (const LDS_AS cplx<T>*)tw;
    switch (R) {
      case 2: wave_pass<T, S, 2, ME>(l3, base, es, d.n, Ns, t3, d.L, q, active, vp, usev); break;
      case 3: wave_pass<T, S, 3, ME>(l3, base, es, d.n, Ns, t3, d.L, q, active, vp, usev); break;
      case 4: wave_pass<T, S, 4, ME>(l3, base, es, d.n, Ns, t3, d.L, q, active, vp, usev); break;
      case 5: wave_pass<T, S, 5, ME>(l3, base, es, d.n, Ns, t3, d.L, q, active, vp, usev); break;
      case 8: wave_pass<T, S, 8, ME>(l3, base, es, d.n, Ns, t3, d.L, q, active, vp, usev); break;
      default: wave_pass<T, S, 9, ME>(l3, base, es, d.n, Ns, t3, d.L, q, active, vp, usev); break;
    }
    Ns *= R;
  }
}

// One x-row of the local-potential product, row = F(V * F^-1(row)), for the two-pass
// compile-time plans N = R0 R1.  The inverse transform runs the radices R0 then R1 and the
// forward one R1 then R0.  The inverse's last pass leaves butterfly j holding the elements
// j + R0 r, exactly the inputs of the forward's first pass: the row stays in registers
// between the two transforms and V multiplies it there, so a row costs three LDS round
// trips instead of four.  vrow: V/N of the row (length N), read only by active lanes.
template <typename T, int N, int R0, int R1>
__device__ __attribute__((noinline)) void row_vloc_static(LDS_AS cplx<T>* lds, int base,
                                                          const LDS_AS cplx<T>* tw, int q, bool active,
                                                          const T* __restrict__ vrow) {
  constexpr int L = lanes_for(N, R0, R1, 1, 12);
  constexpr int NB = N / R1;   // butterflies of the radix-R1 passes (= R0)
  constexpr int MB = (NB + L - 1) / L;
  // inverse, radix R0 (NS = 1)
  spass<T, +1, N, 1, R0, L>(lds, base, 1, tw, q, active, nullptr, false);
  // inverse, radix R1 (NS = R0, the last pass): butterfly j reads and leaves j + NB r
  cplx<T> v[MB][R1];
  bool on[MB];
#pragma unroll
  for (int s = 0; s < MB; ++s) {
    const int j = q + s * L;
    on[s] = active && (MB * L == NB || j < NB);
    if (on[s]) {
#pragma unroll
      for (int r = 0; r < R1; ++r) v[s][r] = lds[base + j + r * NB];
#pragma unroll
      for (int r = 1; r < R1; ++r) v[s][r] = cmul(v[s][r], tw[j * r]);
      DFT<T, +1, R1>::run(v[s]);
      // times V/N, then the forward transform's first pass (radix R1, NS = 1) in registers
#pragma unroll
      for (int r = 0; r < R1; ++r) v[s][r] = scal(vrow[j + r * NB], v[s][r]);
      DFT<T, -1, R1>::run(v[s]);
    }
  }
  wave_fence();
#pragma unroll
  for (int s = 0; s < MB; ++s) {
    if (on[s]) {
      const int j = q + s * L;
#pragma unroll
      for (int r = 0; r < R1; ++r) lds[base + j * R1 + r] = v[s][r];
    }
  }
  wave_fence();
  // forward, radix R0 (NS = R1)
  spass<T, -1, N, R1, R0, L>(lds, base, 1, tw, q, active, nullptr, false);
}

// An x-row of the local-potential product through the fused two-pass routine where the
// row length has a compile-time plan of two passes, else as two line transforms.
template <typename T>
__device__ inline void wave_row_vloc(cplx<T>* lds, int base, const FftDim& d, const cplx<T>* __restrict__ tw, int q,
                                     bool active, const T* vrow) {
  LDS_AS cplx<T>* l3 = (LDS_AS cplx<T>*)lds;
  const LDS_AS cplx<T>* t3 = (const LDS_AS cplx<T>*)tw;
  switch (d.plan) {
    case 72: row_vloc_static<T, 72, 9, 8>(l3, base, t3, q, active, vrow); return;
    case 27: row_vloc_static<T, 27, 9, 3>(l3, base, t3, q, active, vrow); return;
    case 45: row_vloc_static<T, 45, 9, 5>(l3, base, t3, q, active, vrow); return;
    case 20: row_vloc_static<T, 20, 5, 4>(l3, base, t3, q, active, vrow); return;
    default: break;
  }
  wave_line_fft<T, +1, FFT_ME>(lds, base, 1, d, tw, q, active, nullptr);
  wave_line_fft<T, -1, FFT_ME>(lds, base, 1, d, tw, q, active, active ? vrow : nullptr);
}

// All `nlines` lines of a workgroup, distributed over its wavefronts; line t starts at
// lds[lbase ? lbase[t] : t * lstride].  Wavefronts run their lines independently.
template <typename T, int S, int ME>
__device__ inline void wave_lines(cplx<T>* lds, int nlines, const int* lbase, int lstride, int es,
                                  const FftDim& d, const cplx<T>* __restrict__ tw) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  const int g = lane / d.L, q = lane - g * d.L;
  for (int t0 = wave * d.gpw; t0 < nlines; t0 += nwave * d.gpw) {
    const int t = t0 + g;
    const bool active = g < d.gpw && t < nlines;
    const int base = active ? (lbase ? lbase[t] : t * lstride) : 0;
    wave_line_fft<T, S, ME>(lds, base, es, d, tw, q, active, nullptr);
  }
}

// copy a table into LDS (caller synchronises)
template <typename T>
__device__ inline void stage(T* dst, const T* __restrict__ src, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

// ----------------------------------------------------------------------------- kernels

// Column kernels run on a 1-D grid of nchunk * roundup(nbb, 8) workgroups.  Workgroups are
// dispatched round-robin over the 8 XCDs (w mod 8), so block w is given band
// (w mod 8) + 8 * ((w / 8) / nchunk) and column chunk (w / 8) mod nchunk: every chunk of a band
// then runs on one XCD and the scattered psi / Hpsi rows of that band share one L2 instead
// of being fetched into all eight.  Blocks past the last band return at once.
__device__ inline bool col_block(int ncol, int C, int nbb, int& c0, int& bl) {
  const int nchunk = (ncol + C - 1) / C;
  const int w = blockIdx.x, s = w >> 3;
  bl = (w & 7) + 8 * (s / nchunk);
  c0 = (s % nchunk) * C;
  return bl < nbb;
}

static inline int col_grid(int ncol, int C, int nbb) { return ((ncol + C - 1) / C) * ((nbb + 7) / 8 * 8); }

// 1. scatter + inverse z-FFT of the columns [c0, c0+C) of one band: p is the band's psi
// column, w its block of W (n3 x ncol).
template <typename T>
__device__ inline void col_ifft_body(const cplx<double>* __restrict__ p, cplx<T>* __restrict__ w,
                                     const int* __restrict__ ent_igw, const int* __restrict__ ent_pos,
                                     const int* __restrict__ col_ptr, int ncol, int C, int c0, FftDim dz,
                                     const cplx<T>* __restrict__ twz) {
  extern __shared__ char smem_raw[];
  cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
  const int n3 = dz.n, cs = odd_stride(n3);
  cplx<T>* tw = lds + C * cs;
  const int cn = min(C, ncol - c0);
  const int tid = threadIdx.x, nthr = blockDim.x;
  for (int i = tid; i < cn * cs; i += nthr) lds[i] = mk((T)0, (T)0);
  stage(tw, twz, n3);
  __syncthreads();
  const int e0 = col_ptr[c0], e1 = col_ptr[c0 + cn];
  const int off = c0 * cs;
  for (int e = e0 + tid; e < e1; e += nthr) lds[ent_pos[e] - off] = cvt<T>(p[ent_igw[e]]);
  __syncthreads();
  wave_lines<T, +1, FFT_ME>(lds, cn, nullptr, cs, 1, dz, tw);
  __syncthreads();
  for (int i = tid; i < cn * n3; i += nthr) {
    const int cl = i % cn, z = i / cn;
    w[(long)z * ncol + c0 + cl] = lds[cl * cs + z];
  }
}

template <typename T>
__global__ void __launch_bounds__(COL_NT)
k_col_ifft(const cplx<double>* __restrict__ psi, long ldpsi, int b0, cplx<T>* __restrict__ W,
           const int* __restrict__ ent_igw, const int* __restrict__ ent_pos,
           const int* __restrict__ col_ptr, int ncol, int C, int nbb, FftDim dz,
           const cplx<T>* __restrict__ twz) {
  int bl, c0;
  if (!col_block(ncol, C, nbb, c0, bl)) return;
  col_ifft_body<T>(psi + (long)(b0 + bl) * ldpsi, W + (long)bl * dz.n * ncol, ent_igw, ent_pos, col_ptr, ncol, C,
                   c0, dz, twz);
}

// 2. per (z-plane, band): inverse y-FFT on the active x lines, then per x-row one wavefront
// does inverse x-FFT, * V/N (read from L2 inside the first forward pass) and forward
// x-FFT, then forward y-FFT on the active x lines.  In place on the plane w (ncol entries).
template <typename T>
__device__ inline void plane_vloc_body(cplx<T>* __restrict__ w, const T* __restrict__ V, int z,
                                       const int* __restrict__ col_ij, int ncol, const int* __restrict__ xact,
                                       int nxact, FftDim dx, FftDim dy, const cplx<T>* __restrict__ twx,
                                       const cplx<T>* __restrict__ twy) {
  extern __shared__ char smem_raw[];
  cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
  const int n1 = dx.n, n2 = dy.n, np = n1 * n2, ps = odd_stride(n1);
  cplx<T>* tx = lds + ps * n2;
  cplx<T>* ty = tx + n1;
  int* xl = reinterpret_cast<int*>(ty + n2);
  const int tid = threadIdx.x, nthr = blockDim.x;
  for (int i = tid; i < ps * n2; i += nthr) lds[i] = mk((T)0, (T)0);
  stage(tx, twx, n1);
  stage(ty, twy, n2);
  stage(xl, xact, nxact);
  __syncthreads();
  for (int c = tid; c < ncol; c += nthr) lds[col_ij[c]] = w[c];
  __syncthreads();
  wave_lines<T, +1, FFT_ME>(lds, nxact, xl, 0, ps, dy, ty);
  __syncthreads();
  {
    const T* v = V + (long)z * np;
    const int lane = tid & 63, wave = tid >> 6, nwave = nthr >> 6;
    const int g = lane / dx.L, q = lane - g * dx.L;
    for (int t0 = wave * dx.gpw; t0 < n2; t0 += nwave * dx.gpw) {
      const int row = t0 + g;
      const bool active = g < dx.gpw && row < n2;
      const int base = active ? row * ps : 0;
      wave_row_vloc<T>(lds, base, dx, tx, q, active, v + (long)(active ? row : 0) * n1);
    }
  }
  __syncthreads();
  wave_lines<T, -1, FFT_ME>(lds, nxact, xl, 0, ps, dy, ty);
  __syncthreads();
  for (int c = tid; c < ncol; c += nthr) w[c] = lds[col_ij[c]];
}

template <typename T>
__global__ void __launch_bounds__(PlaneNT<T>::value)
k_plane_vloc(cplx<T>* __restrict__ W, const T* __restrict__ V, const int* __restrict__ col_ij,
             int ncol, const int* __restrict__ xact, int nxact, FftDim dx, FftDim dy,
             const cplx<T>* __restrict__ twx, const cplx<T>* __restrict__ twy, int n3) {
  const int z = blockIdx.x, bl = blockIdx.y;
  plane_vloc_body<T>(W + ((long)bl * n3 + z) * ncol, V, z, col_ij, ncol, xact, nxact, dx, dy, twx, twy);
}

// 2c. The plane pass for planes whose padded form does not fit in LDS (FP64 at 108^2):
// only the active x lines are kept, compressed (Yc: nxact lines of n2, stride ys; column c
// sits at slot col_cy[c] = a*ys + y for x = xact[a]).  The x-transform of each row runs in a
// per-wavefront row buffer of n1 entries: the row is expanded from Yc (zero at inactive x),
// transformed (inverse, * V/N, forward) and its active entries gathered back.  rl rows are
// in flight per wavefront (rl <= gpw, sized to the LDS left after Yc).
// With racc set the row is not transformed back: RM 0 adds wb |f(r)|^2 into racc (density),
// RM 1 writes wb Re f(r), RM 2 adds wb Re f(r) (G->R of a real field).
template <typename T, int RM = 0>
__device__ inline void plane_rows_c(cplx<T>* __restrict__ Yc, int ys, cplx<T>* __restrict__ rows,
                                    const int* __restrict__ xl, int nxact, const FftDim& dx, int n2,
                                    const cplx<T>* __restrict__ tx, const T* __restrict__ v, int rl,
                                    double* __restrict__ racc, double wb) {
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int lane = tid & 63, wave = tid >> 6, nwave = nthr >> 6;
  const int n1 = dx.n;
  const int g = lane / dx.L, q = lane - g * dx.L;
  for (int t0 = wave * rl; t0 < n2; t0 += nwave * rl) {
    const int row = t0 + g;
    const bool active = g < rl && row < n2;
    const int base = active ? (wave * rl + g) * n1 : 0;
    if (active)
      for (int i = q; i < n1; i += dx.L) rows[base + i] = mk((T)0, (T)0);
    wave_fence();
    if (active)
      for (int a = q; a < nxact; a += dx.L) rows[base + xl[a]] = Yc[a * ys + row];
    wave_fence();
    wave_line_fft<T, +1, FFT_ME>(rows, base, 1, dx, tx, q, active, nullptr);
    if (racc) {
      if (active)
        for (int i = q; i < n1; i += dx.L) {
          const double re = (double)rows[base + i].x, im = (double)rows[base + i].y;
          double& o = racc[(long)row * n1 + i];
          if (RM == 0)
            o += wb * (re * re + im * im);
          else if (RM == 1)
            o = wb * re;
          else
            o += wb * re;
        }
      wave_fence();
      continue;
    }
    wave_line_fft<T, -1, FFT_ME>(rows, base, 1, dx, tx, q, active, active ? v + (long)row * n1 : nullptr);
    if (active)
      for (int a = q; a < nxact; a += dx.L) Yc[a * ys + row] = rows[base + xl[a]];
    wave_fence();
  }
}

template <typename T>
__device__ inline void plane_vloc_c_body(cplx<T>* __restrict__ w, const T* __restrict__ V, int z,
                                         const int* __restrict__ col_cy, int ncol, const int* __restrict__ xact,
                                         int nxact, FftDim dx, FftDim dy, const cplx<T>* __restrict__ twx,
                                         const cplx<T>* __restrict__ twy, int rl) {
  extern __shared__ char smem_raw[];
  cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
  const int n1 = dx.n, n2 = dy.n, ys = odd_stride(n2);
  cplx<T>* Yc = lds;
  cplx<T>* tx = Yc + nxact * ys;
  cplx<T>* ty = tx + n1;
  cplx<T>* rows = ty + n2;
  int* xl = reinterpret_cast<int*>(rows + (blockDim.x >> 6) * rl * n1);
  const int tid = threadIdx.x, nthr = blockDim.x;
  for (int i = tid; i < nxact * ys; i += nthr) Yc[i] = mk((T)0, (T)0);
  stage(tx, twx, n1);
  stage(ty, twy, n2);
  stage(xl, xact, nxact);
  __syncthreads();
  for (int c = tid; c < ncol; c += nthr) Yc[col_cy[c]] = w[c];
  __syncthreads();
  wave_lines<T, +1, FFT_ME>(Yc, nxact, nullptr, ys, 1, dy, ty);
  __syncthreads();
  plane_rows_c<T>(Yc, ys, rows, xl, nxact, dx, n2, tx, V + (long)z * n1 * n2, rl, nullptr, 0.0);
  __syncthreads();
  wave_lines<T, -1, FFT_ME>(Yc, nxact, nullptr, ys, 1, dy, ty);
  __syncthreads();
  for (int c = tid; c < ncol; c += nthr) w[c] = Yc[col_cy[c]];
}

template <typename T>
__global__ void __launch_bounds__(PlaneNT<T>::value)
k_plane_vloc_c(cplx<T>* __restrict__ W, const T* __restrict__ V, const int* __restrict__ col_cy, int ncol,
               const int* __restrict__ xact, int nxact, FftDim dx, FftDim dy, const cplx<T>* __restrict__ twx,
               const cplx<T>* __restrict__ twy, int n3, int rl) {
  const int z = blockIdx.x, bl = blockIdx.y;
  plane_vloc_c_body<T>(W + ((long)bl * n3 + z) * ncol, V, z, col_cy, ncol, xact, nxact, dx, dy, twx, twy, rl);
}

// Density planes in the compressed form: sum_b w_b |psi_b(r)|^2 accumulated straight into
// the block's partial plane in global memory (each element is owned by one lane per band,
// and the bands of a group run one after another in the block).
template <typename T>
__global__ void __launch_bounds__(PlaneNT<T>::value)
k_plane_rhoe_c(const cplx<T>* __restrict__ W, const double* __restrict__ wts, int nb, int bpg,
               double* __restrict__ rpart, long N, const int* __restrict__ col_cy, int ncol,
               const int* __restrict__ xact, int nxact, FftDim dx, FftDim dy, const cplx<T>* __restrict__ twx,
               const cplx<T>* __restrict__ twy, int n3, int rl) {
  extern __shared__ char smem_raw[];
  cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
  const int n1 = dx.n, n2 = dy.n, np = n1 * n2, ys = odd_stride(n2);
  cplx<T>* Yc = lds;
  cplx<T>* tx = Yc + nxact * ys;
  cplx<T>* ty = tx + n1;
  cplx<T>* rows = ty + n2;
  int* xl = reinterpret_cast<int*>(rows + (blockDim.x >> 6) * rl * n1);
  const int z = blockIdx.x, g = blockIdx.y;
  const int tid = threadIdx.x, nthr = blockDim.x;
  double* out = rpart + (long)g * N + (long)z * np;
  stage(tx, twx, n1);
  stage(ty, twy, n2);
  stage(xl, xact, nxact);
  for (int i = tid; i < np; i += nthr) out[i] = 0.0;
  const int bend = min(nb, (g + 1) * bpg);
  for (int b = g * bpg; b < bend; ++b) {
    for (int i = tid; i < nxact * ys; i += nthr) Yc[i] = mk((T)0, (T)0);
    __syncthreads();
    const cplx<T>* w = W + ((long)b * n3 + z) * ncol;
    for (int c = tid; c < ncol; c += nthr) Yc[col_cy[c]] = w[c];
    __syncthreads();
    wave_lines<T, +1, FFT_ME>(Yc, nxact, nullptr, ys, 1, dy, ty);
    __syncthreads();
    plane_rows_c<T>(Yc, ys, rows, xl, nxact, dx, n2, tx, nullptr, rl, out, wts[b]);
    __syncthreads();
  }
}

// G->R of real fields on a sphere (FP64), after k_col_ifft: per (z-plane, field) the inverse
// y-FFT of the active x lines and the inverse x-FFT of every row, then
// out[f N + r] = scale Re f(r) (added to out when acc is set).
__global__ void __launch_bounds__(PlaneNT<double>::value)
k_plane_g2r(const cplx<double>* __restrict__ W, double* __restrict__ out, long N, double scale, int acc,
            const int* __restrict__ col_ij, int ncol, const int* __restrict__ xact, int nxact, FftDim dx, FftDim dy,
            const cplx<double>* __restrict__ twx, const cplx<double>* __restrict__ twy, int n3) {
  extern __shared__ char smem_raw[];
  cplx<double>* lds = reinterpret_cast<cplx<double>*>(smem_raw);
  const int n1 = dx.n, n2 = dy.n, np = n1 * n2, ps = odd_stride(n1);
  cplx<double>* tx = lds + ps * n2;
  cplx<double>* ty = tx + n1;
  int* xl = reinterpret_cast<int*>(ty + n2);
  const int z = blockIdx.x, f = blockIdx.y;
  const int tid = threadIdx.x, nthr = blockDim.x;
  for (int i = tid; i < ps * n2; i += nthr) lds[i] = mk(0.0, 0.0);
  stage(tx, twx, n1);
  stage(ty, twy, n2);
  stage(xl, xact, nxact);
  __syncthreads();
  const cplx<double>* w = W + ((long)f * n3 + z) * ncol;
  for (int c = tid; c < ncol; c += nthr) lds[col_ij[c]] = w[c];
  __syncthreads();
  wave_lines<double, +1, FFT_ME>(lds, nxact, xl, 0, ps, dy, ty);
  __syncthreads();
  wave_lines<double, +1, FFT_ME>(lds, n2, nullptr, ps, 1, dx, tx);
  __syncthreads();
  double* o = out + (long)f * N + (long)z * np;
  for (int i = tid; i < np; i += nthr) {
    const int j = i / n1, x = i - j * n1;
    const double v = scale * lds[x + ps * j].x;
    o[i] = acc ? o[i] + v : v;
  }
}

// The same for planes beyond LDS, in the compressed form of k_plane_vloc_c.
__global__ void __launch_bounds__(PlaneNT<double>::value)
k_plane_g2r_c(const cplx<double>* __restrict__ W, double* __restrict__ out, long N, double scale, int acc,
              const int* __restrict__ col_cy, int ncol, const int* __restrict__ xact, int nxact, FftDim dx,
              FftDim dy, const cplx<double>* __restrict__ twx, const cplx<double>* __restrict__ twy, int n3,
              int rl) {
  extern __shared__ char smem_raw[];
  cplx<double>* lds = reinterpret_cast<cplx<double>*>(smem_raw);
  const int n1 = dx.n, n2 = dy.n, np = n1 * n2, ys = odd_stride(n2);
  cplx<double>* Yc = lds;
  cplx<double>* tx = Yc + nxact * ys;
  cplx<double>* ty = tx + n1;
  cplx<double>* rows = ty + n2;
  int* xl = reinterpret_cast<int*>(rows + (blockDim.x >> 6) * rl * n1);
  const int z = blockIdx.x, f = blockIdx.y;
  const int tid = threadIdx.x, nthr = blockDim.x;
  for (int i = tid; i < nxact * ys; i += nthr) Yc[i] = mk(0.0, 0.0);
  stage(tx, twx, n1);
  stage(ty, twy, n2);
  stage(xl, xact, nxact);
  __syncthreads();
  const cplx<double>* w = W + ((long)f * n3 + z) * ncol;
  for (int c = tid; c < ncol; c += nthr) Yc[col_cy[c]] = w[c];
  __syncthreads();
  wave_lines<double, +1, FFT_ME>(Yc, nxact, nullptr, ys, 1, dy, ty);
  __syncthreads();
  double* o = out + (long)f * N + (long)z * np;
  if (acc)
    plane_rows_c<double, 2>(Yc, ys, rows, xl, nxact, dx, n2, tx, nullptr, rl, o, scale);
  else
    plane_rows_c<double, 1>(Yc, ys, rows, xl, nxact, dx, n2, tx, nullptr, rl, o, scale);
}

// 2g. Planes whose compressed form does not fit in LDS either (FP64 on the rho sphere at
// 108^2: 107 active x lines of 108): the compressed plane of each (z, band) lives in global
// memory (Yg: nxact lines of stride ys), and the plane pass is split in three launches.
// k_yfft_gp transforms groups of `glines` active x lines along y in LDS: with S = +1 the
// lines are scattered from W first (start of G->R), with S = -1 they are loaded from Yg and
// their sphere entries gathered back into W after the transform (end of R->G).  k_xrows_gp
// runs plane_rows_c on the rows of Yg (row buffers in LDS): x-inverse, * V/N, x-forward
// written back to Yg (RM < 0), or x-inverse and scale Re f(r) into out (RM 1 write, 2 add).
template <typename T, int S>
__global__ void __launch_bounds__(PlaneNT<T>::value)
k_yfft_gp(cplx<T>* __restrict__ W, cplx<T>* __restrict__ Yg, const int* __restrict__ col_cy, int ncol, int nxact,
          int glines, FftDim dy, const cplx<T>* __restrict__ twy, int n3) {
  extern __shared__ char smem_raw[];
  cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
  const int n2 = dy.n, ys = odd_stride(n2);
  const int z = blockIdx.x, b = blockIdx.y, a0 = blockIdx.z * glines;
  const int nl = min(glines, nxact - a0);
  if (nl <= 0) return;
  cplx<T>* ty = lds + glines * ys;
  const int tid = threadIdx.x, nthr = blockDim.x;
  cplx<T>* w = W + ((long)b * n3 + z) * ncol;
  cplx<T>* yg = Yg + ((long)b * n3 + z) * nxact * ys + (long)a0 * ys;
  const int s0 = a0 * ys, s1 = (a0 + nl) * ys;
  stage(ty, twy, n2);
  if (S > 0) {
    for (int i = tid; i < nl * ys; i += nthr) lds[i] = mk((T)0, (T)0);
    __syncthreads();
    for (int c = tid; c < ncol; c += nthr) {
      const int s = col_cy[c];
      if (s >= s0 && s < s1) lds[s - s0] = w[c];
    }
  } else {
    for (int i = tid; i < nl * ys; i += nthr) lds[i] = yg[i];
  }
  __syncthreads();
  wave_lines<T, S, FFT_ME>(lds, nl, nullptr, ys, 1, dy, ty);
  __syncthreads();
  if (S > 0) {
    for (int i = tid; i < nl * ys; i += nthr) yg[i] = lds[i];
  } else {
    for (int c = tid; c < ncol; c += nthr) {
      const int s = col_cy[c];
      if (s >= s0 && s < s1) w[c] = lds[s - s0];
    }
  }
}

template <typename T, int RM>
__global__ void __launch_bounds__(PlaneNT<T>::value)
k_xrows_gp(cplx<T>* __restrict__ Yg, const T* __restrict__ V, double* __restrict__ out, long N, double scale,
           const int* __restrict__ xact, int nxact, FftDim dx, int n2, const cplx<T>* __restrict__ twx, int n3,
           int rl) {
  extern __shared__ char smem_raw[];
  cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
  const int n1 = dx.n, ys = odd_stride(n2);
  cplx<T>* tx = lds;
  cplx<T>* rows = tx + n1;
  int* xl = reinterpret_cast<int*>(rows + (blockDim.x >> 6) * rl * n1);
  const int z = blockIdx.x, b = blockIdx.y;
  stage(tx, twx, n1);
  stage(xl, xact, nxact);
  __syncthreads();
  cplx<T>* yc = Yg + ((long)b * n3 + z) * nxact * ys;
  if constexpr (RM < 0)
    plane_rows_c<T>(yc, ys, rows, xl, nxact, dx, n2, tx, V + (long)z * n1 * n2, rl, nullptr, 0.0);
  else
    plane_rows_c<T, RM>(yc, ys, rows, xl, nxact, dx, n2, tx, nullptr, rl, out + (long)b * N + (long)z * n1 * n2,
                        scale);
}

// 3. forward z-FFT of the columns, gather the sphere, add kinetic, write Hpsi (one band:
// w its block of W, p / h its psi / Hpsi columns).
template <typename T>
__device__ inline void col_fft_body(const cplx<T>* __restrict__ w, const cplx<double>* __restrict__ p,
                                    cplx<double>* __restrict__ h, const int* __restrict__ ent_igw,
                                    const int* __restrict__ ent_pos, const int* __restrict__ col_ptr, int ncol,
                                    int C, int c0, FftDim dz, const cplx<T>* __restrict__ twz,
                                    const double* __restrict__ gk2, int kin, int acc) {
  extern __shared__ char smem_raw[];
  cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
  const int n3 = dz.n, cs = odd_stride(n3);
  cplx<T>* tw = lds + C * cs;
  const int cn = min(C, ncol - c0);
  const int tid = threadIdx.x, nthr = blockDim.x;
  for (int i = tid; i < cn * n3; i += nthr) {
    const int cl = i % cn, z = i / cn;
    lds[cl * cs + z] = w[(long)z * ncol + c0 + cl];
  }
  stage(tw, twz, n3);
  __syncthreads();
  wave_lines<T, -1, FFT_ME>(lds, cn, nullptr, cs, 1, dz, tw);
  __syncthreads();
  const int e0 = col_ptr[c0], e1 = col_ptr[c0 + cn];
  const int off = c0 * cs;
  for (int e = e0 + tid; e < e1; e += nthr) {
    const int g = ent_igw[e];
    cplx<double> r = cvt<double>(lds[ent_pos[e] - off]);
    if (kin) r = r + scal(0.5 * gk2[g], p[g]);
    if (acc) r = r + h[g];
    h[g] = r;
  }
}

template <typename T>
__global__ void __launch_bounds__(COL_NT)
k_col_fft(const cplx<T>* __restrict__ W, const cplx<double>* __restrict__ psi, long ldpsi,
          cplx<double>* __restrict__ hpsi, long ldh, int b0, const int* __restrict__ ent_igw,
          const int* __restrict__ ent_pos, const int* __restrict__ col_ptr, int ncol, int C, int nbb,
          FftDim dz, const cplx<T>* __restrict__ twz, const double* __restrict__ gk2, int kin,
          int acc) {
  int bl, c0;
  if (!col_block(ncol, C, nbb, c0, bl)) return;
  col_fft_body<T>(W + (long)bl * dz.n * ncol, psi + (long)(b0 + bl) * ldpsi, hpsi + (long)(b0 + bl) * ldh,
                  ent_igw, ent_pos, col_ptr, ncol, C, c0, dz, twz, gk2, kin, acc);
}

// ---- batched forms: several k-points (same band count nb) in one launch per stage.
// Column kernels: k-point kk owns workgroups [kb[kk].blk0, kb[kk+1].blk0), a multiple of 8,
// mapped to (band, chunk) as col_block does.  Plane kernel: blockIdx.y = kk * nb + band.
__device__ inline int batch_find(const KBatch* __restrict__ kb, int nk, int w) {
  int lo = 0, hi = nk - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (kb[mid].blk0 <= w) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__device__ inline bool batch_col_block(const KBatch& d, int C, int nb, int wl, int& c0, int& bl) {
  const int nchunk = (d.ncol + C - 1) / C;
  const int s = wl >> 3;
  bl = (wl & 7) + 8 * (s / nchunk);
  c0 = (s % nchunk) * C;
  return bl < nb;
}

template <typename T>
__global__ void __launch_bounds__(COL_NT)
k_col_ifft_batch(const KBatch* __restrict__ kb, int nk, int nb, cplx<T>* __restrict__ W, int C, FftDim dz,
                 const cplx<T>* __restrict__ twz) {
  const int w = blockIdx.x;
  const KBatch& d = kb[batch_find(kb, nk, w)];
  int bl, c0;
  if (!batch_col_block(d, C, nb, w - d.blk0, c0, bl)) return;
  col_ifft_body<T>(d.psi + (long)bl * d.ldpsi, W + d.woff + (long)bl * dz.n * d.ncol, d.ent_igw, d.ent_pos,
                   d.col_ptr, d.ncol, C, c0, dz, twz);
}

template <typename T>
__global__ void __launch_bounds__(PlaneNT<T>::value)
k_plane_vloc_batch(const KBatch* __restrict__ kb, int nb, cplx<T>* __restrict__ W, const T* __restrict__ V,
                   FftDim dx, FftDim dy, const cplx<T>* __restrict__ twx, const cplx<T>* __restrict__ twy,
                   int n3) {
  const int z = blockIdx.x, kk = blockIdx.y / nb, bl = blockIdx.y - kk * nb;
  const KBatch& d = kb[kk];
  plane_vloc_body<T>(W + d.woff + ((long)bl * n3 + z) * d.ncol, V, z, d.col_ij, d.ncol, d.xact, d.nxact, dx, dy,
                     twx, twy);
}

template <typename T>
__global__ void __launch_bounds__(COL_NT)
k_col_fft_batch(const KBatch* __restrict__ kb, int nk, int nb, const cplx<T>* __restrict__ W, int C, FftDim dz,
                const cplx<T>* __restrict__ twz, int kin, int acc) {
  const int w = blockIdx.x;
  const KBatch& d = kb[batch_find(kb, nk, w)];
  int bl, c0;
  if (!batch_col_block(d, C, nb, w - d.blk0, c0, bl)) return;
  col_fft_body<T>(W + d.woff + (long)bl * dz.n * d.ncol, d.psi + (long)bl * d.ldpsi, d.hpsi + (long)bl * d.ldh,
                  d.ent_igw, d.ent_pos, d.col_ptr, d.ncol, C, c0, dz, twz, d.gk2, kin, acc);
}

// rho planes: per (z-plane, band group): y/x inverse FFT of each band and
// sum_b w_b |psi_b(r)|^2 accumulated in registers (plane point tid + r * blockDim.x in
// register r, NR <= RHO_REGS per thread: the launch picks the smallest NR that covers the
// plane, so no accumulator registers go unused); one partial plane per group.
constexpr int RHO_REGS = 16;
template <typename T, int NR>
__global__ void __launch_bounds__(PlaneNT<T>::value)
k_plane_rhoe(const cplx<T>* __restrict__ W, const double* __restrict__ wts, int nb, int bpg,
             double* __restrict__ rpart, long N, const int* __restrict__ col_ij, int ncol,
             const int* __restrict__ xact, int nxact, FftDim dx, FftDim dy,
             const cplx<T>* __restrict__ twx, const cplx<T>* __restrict__ twy, int n3) {
  extern __shared__ char smem_raw[];
  cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
  const int n1 = dx.n, n2 = dy.n, np = n1 * n2, ps = odd_stride(n1);
  cplx<T>* tx = lds + ps * n2;
  cplx<T>* ty = tx + n1;
  int* xl = reinterpret_cast<int*>(ty + n2);
  const int z = blockIdx.x, g = blockIdx.y;
  const int tid = threadIdx.x, nthr = blockDim.x;
  stage(tx, twx, n1);
  stage(ty, twy, n2);
  stage(xl, xact, nxact);
  double acc[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) acc[r] = 0.0;
  const int bend = min(nb, (g + 1) * bpg);
  for (int b = g * bpg; b < bend; ++b) {
    for (int i = tid; i < ps * n2; i += nthr) lds[i] = mk((T)0, (T)0);
    __syncthreads();
    const cplx<T>* w = W + ((long)b * n3 + z) * ncol;
    for (int c = tid; c < ncol; c += nthr) lds[col_ij[c]] = w[c];
    __syncthreads();
    wave_lines<T, +1, FFT_ME>(lds, nxact, xl, 0, ps, dy, ty);
    __syncthreads();
    wave_lines<T, +1, FFT_ME>(lds, n2, nullptr, ps, 1, dx, tx);
    __syncthreads();
    const double wb = wts[b];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int i = tid + r * nthr;
      if (i < np) {
        const int j = i / n1, x = i - j * n1;
        const double re = (double)lds[x + ps * j].x, im = (double)lds[x + ps * j].y;
        acc[r] += wb * (re * re + im * im);
      }
    }
    __syncthreads();
  }
  double* out = rpart + (long)g * N + (long)z * np;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const int i = tid + r * nthr;
    if (i < np) out[i] = acc[r];
  }
}

__global__ void k_rho_reduce(double* __restrict__ rho, const double* __restrict__ rpart, int ng,
                             long N, double scale) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int g = 0; g < ng; ++g) s += rpart[(long)g * N + i];
    rho[i] += scale * s;
  }
}

__global__ void k_kinetic(const cplx<double>* __restrict__ psi, long ldpsi, cplx<double>* __restrict__ hpsi,
                          long ldh, const double* __restrict__ gk2, int ngw, int nb, int acc) {
  const long tot = (long)ngw * nb;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < tot; t += (long)gridDim.x * blockDim.x) {
    const int g = (int)(t % ngw), b = (int)(t / ngw);
    cplx<double> r = scal(0.5 * gk2[g], psi[(long)b * ldpsi + g]);
    if (acc) r = r + hpsi[(long)b * ldh + g];
    hpsi[(long)b * ldh + g] = r;
  }
}

__global__ void k_kprec(const cplx<double>* __restrict__ in, long ldin, cplx<double>* __restrict__ out,
                        long ldout, const double* __restrict__ gk2, int ngw, int nb) {
  const long tot = (long)ngw * nb;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < tot; t += (long)gridDim.x * blockDim.x) {
    const int g = (int)(t % ngw), b = (int)(t / ngw);
    out[(long)b * ldout + g] = scal(1.0 / (1.0 + gk2[g]), in[(long)b * ldin + g]);
  }
}

// ----------------------------------------------------------------------------- host drivers

static int cols_per_block(const pw_ctx* c, size_t esz) {
  // column tile <= 48 KiB of LDS, at least one full round of line groups
  const FftDim& d = c->dim[2];
  const int cap = (int)((48 * 1024) / (esz * odd_stride(d.n)));
  if (c->tune_col_chunk > 0) return std::max(1, std::min(c->tune_col_chunk, cap));
  int C = cap;
  const int round = d.gpw * (COL_NT / 64);
  if (C > round) C = (C / round) * round;
  if (C < 1) C = 1;
  return C;
}

static bool lanes_ok(const FftDim& d) { return d.L >= 1 && d.L <= 64 && d.gpw >= 1; }

static size_t plane_lds(const pw_ctx* c, size_t esz, int nxact) {
  return ((size_t)odd_stride(c->n1) * c->n2 + c->n1 + c->n2) * esz + (size_t)nxact * sizeof(int);
}
// Threads of a padded-plane workgroup (the kernels index by blockDim.x): wavefronts for one
// pass over the plane's lines, max(nxact y lines, n2 x rows) / lines per wavefront, when that
// is at most half of PlaneNT.  Small planes (20^2: one wavefront) then fit many workgroups
// per CU instead of one or two mostly idle ones; large planes keep PlaneNT.
template <typename T>
static int plane_threads(const pw_ctx* c, int nxact) {
  const FftDim& dx = c->dim[0];
  const FftDim& dy = c->dim[1];
  const int waves = std::max(1, std::max((nxact + dy.gpw - 1) / dy.gpw, (c->n2 + dx.gpw - 1) / dx.gpw));
  return 2 * 64 * waves <= PlaneNT<T>::value ? 64 * waves : PlaneNT<T>::value;
}

// Threads of a column workgroup holding C columns: wavefronts for one pass over its C lines
// when that is at most half of COL_NT, else COL_NT (the column kernels index by blockDim.x).
static int col_threads(const pw_ctx* c, int C) {
  const int waves = std::max(1, (C + c->dim[2].gpw - 1) / c->dim[2].gpw);
  return 2 * 64 * waves <= COL_NT ? 64 * waves : COL_NT;
}

static size_t col_lds(const pw_ctx* c, int C, size_t esz) { return ((size_t)C * odd_stride(c->n3) + c->n3) * esz; }

// Compressed plane (k_plane_vloc_c / k_plane_rhoe_c): rows in flight per wavefront that fit
// in LDS next to the compressed plane, and the LDS bytes; 0 rows when nothing fits.
static int plane_c_rows(const pw_ctx* c, size_t esz, int nxact, int nthreads, size_t* bytes) {
  const size_t fixed = ((size_t)nxact * odd_stride(c->n2) + c->n1 + c->n2) * esz + (size_t)nxact * sizeof(int);
  const int nwave = nthreads / 64;
  const size_t per = (size_t)nwave * c->n1 * esz;
  const size_t lim = 160 * 1024;
  int rl = fixed < lim ? (int)((lim - fixed) / per) : 0;
  rl = std::min(rl, c->dim[0].gpw);
  *bytes = fixed + (size_t)std::max(rl, 0) * per;
  return rl;
}

// Global-plane pass (k_yfft_gp / k_xrows_gp) for planes whose compressed form exceeds LDS:
// y lines per group (at most 64 KiB of lines) and the x pass's rows in flight per wavefront,
// with the LDS bytes of each launch.
struct PlaneGP {
  int glines = 0, rl = 0;
  size_t ybytes = 0, xbytes = 0;
};
static bool plane_gp(const pw_ctx* c, size_t esz, int nxact, int nthreads, PlaneGP* g) {
  const size_t line = (size_t)odd_stride(c->n2) * esz;
  g->glines = std::max(1, std::min(nxact, (int)((64 * 1024) / line)));
  g->ybytes = (size_t)g->glines * line + (size_t)c->n2 * esz;
  const size_t fixed = (size_t)c->n1 * esz + (size_t)nxact * sizeof(int);
  const size_t per = (size_t)(nthreads / 64) * c->n1 * esz;
  const size_t lim = 160 * 1024;
  int rl = fixed < lim ? (int)((lim - fixed) / per) : 0;
  g->rl = std::min(rl, c->dim[0].gpw);
  g->xbytes = fixed + (size_t)std::max(g->rl, 0) * per;
  return g->rl >= 1 && g->ybytes <= lim;
}

template <typename T>
static int vloc_run(pw_ctx* c, KData& k, const T* V, int nb, const cplx<double>* psi, long ldpsi,
                    cplx<double>* hpsi, long ldh, int kin, int acc, hipEvent_t join = nullptr) {
  const size_t esz = sizeof(cplx<T>);
  const int C = cols_per_block(c, esz);
  // the gather kernel's own chunk: in FP32 at most 32 columns, so fewer bands' chunks share
  // an XCD at once and Hpsi's scattered rows merge in L2 (Si64: 151 -> 87 MB written per
  // launch); W's layout does not depend on the chunk
  const int Cf = (sizeof(T) == 4 && c->tune_col_chunk <= 0) ? std::min(C, 32) : C;
  size_t plane_bytes = plane_lds(c, esz, k.nxact);
  int rl = 0;
  const bool compressed = plane_bytes > 160 * 1024 || c->plane_compressed != 0;
  PlaneGP gp;
  bool global_plane = false;
  if (compressed) {
    rl = plane_c_rows(c, esz, k.nxact, PlaneNT<T>::value, &plane_bytes);
    if (rl < 1 || c->plane_compressed == 2) {
      if (!plane_gp(c, esz, k.nxact, PlaneNT<T>::value, &gp))
        return set_error(c, PW_ERR_ARG, "z-plane does not fit in LDS for this precision");
      global_plane = true;
    }
  }
  if (!lanes_ok(c->dim[0]) || !lanes_ok(c->dim[1]) || !lanes_ok(c->dim[2]))
    return set_error(c, PW_ERR_ARG, "FFT line too long for one wavefront");
  int bb = std::max(1, std::min(nb, c->band_batch));
  // a no-allocation call takes as many bands per launch as its fixed work array holds
  if (c->no_alloc) bb = std::max(1, std::min<int>(bb, (int)(c->work_bytes / ((size_t)c->n3 * k.ncol * esz))));
  const size_t wbytes = (size_t)bb * c->n3 * k.ncol * esz;
  int rc = ensure(c, &c->d_work, &c->work_bytes, wbytes);
  if (rc) return rc;
  cplx<T>* W = reinterpret_cast<cplx<T>*>(c->d_work);
  cplx<T>* Yg = nullptr;
  if (global_plane) {
    rc = ensure(c, &c->d_yglob, &c->yglob_bytes, (size_t)bb * c->n3 * k.nxact * odd_stride(c->n2) * esz);
    if (rc) return rc;
    Yg = reinterpret_cast<cplx<T>*>(c->d_yglob);
  }
  const cplx<T>* tw[3];
  for (int d = 0; d < 3; ++d)
    tw[d] = reinterpret_cast<const cplx<T>*>(sizeof(T) == 8 ? c->d_tw64[d] : c->d_tw32[d]);
  for (int b0 = 0; b0 < nb; b0 += bb) {
    const int nbb = std::min(bb, nb - b0);
    dim3 g1(col_grid(k.ncol, C, nbb));
    // algorithmic HBM bytes: psi in (16 B/coef) + index maps, work array out
    const double wb = (double)nbb * c->n3 * k.ncol * esz;
    const double psib = (double)nbb * k.ngw * 16.0;
    {
      ProfScope ps(c, PW_PROF_COL_IFFT, psib + wb + k.ngw * 8.0);
      hipLaunchKernelGGL(k_col_ifft<T>, g1, dim3(COL_NT), col_lds(c, C, esz), c->stream, psi, ldpsi, b0,
                         W, k.d_ent_igw, k.d_ent_pos, k.d_col_ptr, k.ncol, C, nbb, c->dim[2], tw[2]);
    }
    {
      dim3 g2(c->n3, nbb);
      ProfScope ps(c, PW_PROF_PLANE_VLOC, 2.0 * wb + (double)c->N * sizeof(T));
      if (global_plane) {
        const dim3 gy(c->n3, nbb, (k.nxact + gp.glines - 1) / gp.glines);
        hipLaunchKernelGGL((k_yfft_gp<T, +1>), gy, dim3(PlaneNT<T>::value), gp.ybytes, c->stream, W, Yg, k.d_col_cy,
                           k.ncol, k.nxact, gp.glines, c->dim[1], tw[1], c->n3);
        hipLaunchKernelGGL((k_xrows_gp<T, -1>), g2, dim3(PlaneNT<T>::value), gp.xbytes, c->stream, Yg, V, nullptr,
                           0L, 0.0, k.d_xact, k.nxact, c->dim[0], c->n2, tw[0], c->n3, gp.rl);
        hipLaunchKernelGGL((k_yfft_gp<T, -1>), gy, dim3(PlaneNT<T>::value), gp.ybytes, c->stream, W, Yg, k.d_col_cy,
                           k.ncol, k.nxact, gp.glines, c->dim[1], tw[1], c->n3);
      } else if (compressed)
        hipLaunchKernelGGL(k_plane_vloc_c<T>, g2, dim3(PlaneNT<T>::value), plane_bytes, c->stream, W, V,
                           k.d_col_cy, k.ncol, k.d_xact, k.nxact, c->dim[0], c->dim[1], tw[0], tw[1], c->n3, rl);
      else
        hipLaunchKernelGGL(k_plane_vloc<T>, g2, dim3(plane_threads<T>(c, k.nxact)), plane_bytes, c->stream, W, V,
                           k.d_col_ij, k.ncol, k.d_xact, k.nxact, c->dim[0], c->dim[1], tw[0], tw[1], c->n3);
    }
    if (join && b0 == 0) {
      hipError_t e = hipStreamWaitEvent(c->stream, join, 0);
      if (e != hipSuccess) return hip_check(c, e, "stream join");
    }
    {
      ProfScope ps(c, PW_PROF_COL_FFT, wb + psib * ((kin ? 1 : 0) + (acc ? 1 : 0) + 1) + k.ngw * 16.0);
      hipLaunchKernelGGL(k_col_fft<T>, dim3(col_grid(k.ncol, Cf, nbb)), dim3(COL_NT), col_lds(c, Cf, esz), c->stream,
                         W, psi, ldpsi, hpsi, ldh, b0, k.d_ent_igw, k.d_ent_pos, k.d_col_ptr, k.ncol, Cf, nbb,
                         c->dim[2], tw[2], k.d_gk2, kin, acc);
    }
  }
  return hip_check(c, hipGetLastError(), "vloc kernels");
}

int vloc_apply(pw_ctx* c, int ik, int ispin, int nb, const cplx<double>* psi, long ldpsi,
               cplx<double>* hpsi, long ldh, unsigned flags, hipEvent_t join) {
  const int kin = (flags & PW_TERM_K) ? 1 : 0;
  const int acc = (flags & PW_ACCUMULATE) ? 1 : 0;
  if (flags & PW_MIXED)
    return vloc_run<float>(c, c->kd[ik], c->d_v32[ispin], nb, psi, ldpsi, hpsi, ldh, kin, acc, join);
  return vloc_run<double>(c, c->kd[ik], c->d_v64[ispin], nb, psi, ldpsi, hpsi, ldh, kin, acc, join);
}

int vloc_apply_k(pw_ctx* c, KData& k, const double* V64, int nb, const cplx<double>* psi, long ldpsi,
                 cplx<double>* out, long ldout) {
  return vloc_run<double>(c, k, V64, nb, psi, ldpsi, out, ldout, 0, 0);
}

template <typename T>
static int vloc_batch_run(pw_ctx* c, int nk, const int* iks, const T* V, int nb, const cplx<double>* const* psi,
                          const int64_t* ldpsi, cplx<double>* const* hpsi, const int64_t* ldh, int kin, int acc) {
  const size_t esz = sizeof(cplx<T>);
  int C = cols_per_block(c, esz);
  if (!lanes_ok(c->dim[0]) || !lanes_ok(c->dim[1]) || !lanes_ok(c->dim[2]))
    return set_error(c, PW_ERR_ARG, "FFT line too long for one wavefront");
  // no column chunk wider than the widest k-point's column set (small grids: one chunk per
  // band), so the LDS tile and the workgroup are sized to the columns actually there
  if (c->tune_col_chunk <= 0) {
    int maxcol = 1;
    for (int kk = 0; kk < nk; ++kk) maxcol = std::max(maxcol, c->kd[iks[kk]].ncol);
    C = std::min(C, maxcol);
  }
  const int cnt = col_threads(c, C);
  const int nb8 = (nb + 7) / 8 * 8;
  std::vector<KBatch> kb(nk);
  long woff = 0;
  long blk = 0;
  int maxx = 0;
  for (int kk = 0; kk < nk; ++kk) {
    const KData& k = c->kd[iks[kk]];
    KBatch d{};
    d.ent_igw = k.d_ent_igw;
    d.ent_pos = k.d_ent_pos;
    d.col_ptr = k.d_col_ptr;
    d.col_ij = k.d_col_ij;
    d.xact = k.d_xact;
    d.gk2 = k.d_gk2;
    d.psi = psi[kk];
    d.hpsi = hpsi[kk];
    d.ldpsi = ldpsi[kk];
    d.ldh = ldh[kk];
    d.woff = woff;
    d.ncol = k.ncol;
    d.nxact = k.nxact;
    d.blk0 = (int)blk;
    kb[kk] = d;
    woff += (long)nb * c->n3 * k.ncol;
    blk += (long)((k.ncol + C - 1) / C) * nb8;
    maxx = std::max(maxx, k.nxact);
  }
  if (blk > 0x7fffffffL) return set_error(c, PW_ERR_ARG, "op_H batch too large for one launch");
  const size_t plane_bytes = plane_lds(c, esz, maxx);
  if (plane_bytes > 160 * 1024)
    return set_error(c, PW_ERR_ARG, "z-plane does not fit in LDS for this precision");
  int rc = ensure(c, &c->d_work, &c->work_bytes, (size_t)woff * esz);
  if (rc) return rc;
  const void* dptr = nullptr;
  rc = desc_get(c, kb.data(), kb.size() * sizeof(KBatch), &dptr);
  if (rc) return rc;
  const KBatch* dkb = reinterpret_cast<const KBatch*>(dptr);
  cplx<T>* W = reinterpret_cast<cplx<T>*>(c->d_work);
  const cplx<T>* tw[3];
  for (int d = 0; d < 3; ++d)
    tw[d] = reinterpret_cast<const cplx<T>*>(sizeof(T) == 8 ? c->d_tw64[d] : c->d_tw32[d]);
  hipLaunchKernelGGL(k_col_ifft_batch<T>, dim3((unsigned)blk), dim3(cnt), col_lds(c, C, esz), c->stream, dkb,
                     nk, nb, W, C, c->dim[2], tw[2]);
  hipLaunchKernelGGL(k_plane_vloc_batch<T>, dim3(c->n3, nk * nb), dim3(plane_threads<T>(c, maxx)), plane_bytes, c->stream,
                     dkb, nb, W, V, c->dim[0], c->dim[1], tw[0], tw[1], c->n3);
  hipLaunchKernelGGL(k_col_fft_batch<T>, dim3((unsigned)blk), dim3(cnt), col_lds(c, C, esz), c->stream, dkb, nk,
                     nb, W, C, c->dim[2], tw[2], kin, acc);
  return hip_check(c, hipGetLastError(), "batched vloc kernels");
}

// ---------------------------------------------------------------- fused single-workgroup op_H
// For grids whose whole padded box fits one workgroup's LDS (20^3 in FP64: 134 KB), one
// workgroup applies op_H to one band of one k-point with no intermediate in HBM: scatter psi
// into the box, inverse z-FFT of the sphere's columns, inverse y-FFT of the lines at its
// active x, inverse x-FFT of every row, times V/N (read from L2) inside the forward x-FFT,
// forward y- and z-FFTs on the same pruned lines, then gather + 1/2 |G+k|^2 psi + the
// nonlocal term beta D (beta^H psi) (op_V_loc.jl:62-90, op_K.jl:44-75, op_V_Ps_nloc.jl:39-78).
// All bands of all k-points (uneven Ngw) go in one launch, one workgroup per (k-point, band).
struct FusedK {
  const cplx<double>* psi;
  cplx<double>* hpsi;
  long ldpsi, ldh;
  const int* gpos;       // psi row -> slot in the padded box
  const double* gk2;
  const int* col_ij;     // active columns (slot of z = 0)
  const int* xact;
  const void* beta;      // ngw x nbeta in the kernel's precision (or null)
  const void* betaD;
  int ngw, ncol, nxact, nbeta;
};


template <typename T>
__device__ inline void fused_rows_vloc(cplx<T>* lds, int nrows, int ps, const FftDim& dx, const cplx<T>* tx,
                                       const T* __restrict__ V) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwave = blockDim.x >> 6;
  const int n1 = dx.n;
  const int g = lane / dx.L, q = lane - g * dx.L;
  for (int t0 = wave * dx.gpw; t0 < nrows; t0 += nwave * dx.gpw) {
    const int row = t0 + g;
    const bool active = g < dx.gpw && row < nrows;
    const int base = active ? row * ps : 0;
    wave_row_vloc<T>(lds, base, dx, tx, q, active, V + (long)(active ? row : 0) * n1);
  }
}

template <typename T, int NT>
__global__ void __launch_bounds__(NT)
k_op_H_fused(const FusedK* __restrict__ kd, int nb, const T* __restrict__ V, FftDim dx, FftDim dy, FftDim dz,
             const cplx<T>* __restrict__ twx, const cplx<T>* __restrict__ twy, const cplx<T>* __restrict__ twz,
             int do_v, int do_k, int do_nl, int acc) {
  extern __shared__ char smem_raw[];
  const int kk = blockIdx.x / nb, b = blockIdx.x - kk * nb;
  const FusedK F = kd[kk];
  const int n1 = dx.n, n2 = dy.n, n3 = dz.n, ps = odd_stride(n1);
  const int box = ps * n2 * n3;
  cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* tx = lds + box;
  cplx<T>* ty = tx + n1;
  cplx<T>* tz = ty + n2;
  cplx<T>* Bn = tz + n3;                                   // beta^H psi (nbeta)
  int* lb = reinterpret_cast<int*>(Bn + (F.nbeta > 0 ? F.nbeta : 1));   // line bases
  const int tid = threadIdx.x, nthr = blockDim.x;
  const cplx<double>* p = F.psi + (long)b * F.ldpsi;
  cplx<double>* h = F.hpsi + (long)b * F.ldh;
  const int nyl = F.nxact * n3;
  if (do_v) {
    for (int i = tid; i < box; i += nthr) lds[i] = mk((T)0, (T)0);
    stage(tx, twx, n1);
    stage(ty, twy, n2);
    stage(tz, twz, n3);
    stage(lb, F.col_ij, F.ncol);
    __syncthreads();
    for (int g = tid; g < F.ngw; g += nthr) lds[F.gpos[g]] = cvt<T>(p[g]);
    __syncthreads();
    wave_lines<T, +1, FFT_ME>(lds, F.ncol, lb, 0, ps * n2, dz, tz);       // z-lines of the sphere's columns
    __syncthreads();
    for (int t = tid; t < nyl; t += nthr) {
      const int a = t % F.nxact, z = t / F.nxact;
      lb[t] = F.xact[a] + ps * n2 * z;
    }
    __syncthreads();
    wave_lines<T, +1, FFT_ME>(lds, nyl, lb, 0, ps, dy, ty);             // y-lines at the active x
    __syncthreads();
    fused_rows_vloc<T>(lds, n2 * n3, ps, dx, tx, V);                    // x-rows: inverse, * V/N, forward
    __syncthreads();
    wave_lines<T, -1, FFT_ME>(lds, nyl, lb, 0, ps, dy, ty);
    __syncthreads();
    stage(lb, F.col_ij, F.ncol);
    __syncthreads();
    wave_lines<T, -1, FFT_ME>(lds, F.ncol, lb, 0, ps * n2, dz, tz);
  }
  if (do_nl && F.nbeta > 0) {
    // B_j = sum_g conj(beta_gj) psi_g, one wavefront per projector
    const cplx<T>* be = reinterpret_cast<const cplx<T>*>(F.beta);
    const int lane = tid & 63, wave = tid >> 6, nwave = nthr >> 6;
    for (int j = wave; j < F.nbeta; j += nwave) {
      T sr = 0, si = 0;
      for (int g = lane; g < F.ngw; g += 64) {
        const cplx<T> bg = be[(long)j * F.ngw + g];
        const cplx<T> x = cvt<T>(p[g]);
        sr += bg.x * x.x + bg.y * x.y;
        si += bg.x * x.y - bg.y * x.x;
      }
      for (int o = 32; o > 0; o >>= 1) {
        sr += __shfl_xor(sr, o, 64);
        si += __shfl_xor(si, o, 64);
      }
      if (lane == 0) Bn[j] = mk(sr, si);
    }
  }
  __syncthreads();
  const cplx<T>* bd = reinterpret_cast<const cplx<T>*>(F.betaD);
  for (int g = tid; g < F.ngw; g += nthr) {
    const cplx<double> x = p[g];
    cplx<double> r = acc ? h[g] : mk(0.0, 0.0);
    if (do_v) r = r + cvt<double>(lds[F.gpos[g]]);
    if (do_k) r = r + scal(0.5 * F.gk2[g], x);
    if (do_nl && F.nbeta > 0) {
      cplx<T> s = mk((T)0, (T)0);
      for (int j = 0; j < F.nbeta; ++j) s = s + cmul(bd[(long)j * F.ngw + g], Bn[j]);
      r = r + cvt<double>(s);
    }
    h[g] = r;
  }
}

static size_t fused_lds(const pw_ctx* c, size_t esz, int maxcol, int maxyl, int nbeta) {
  const size_t box = (size_t)odd_stride(c->n1) * c->n2 * c->n3;
  return (box + c->n1 + c->n2 + c->n3 + std::max(nbeta, 1)) * esz + (size_t)std::max(maxcol, maxyl) * sizeof(int);
}

bool op_H_fused_fits(const pw_ctx* c, bool mixed) {
  const size_t esz = mixed ? sizeof(cplx<float>) : sizeof(cplx<double>);
  // worst case line tables: every column and every x active
  const size_t need = fused_lds(c, esz, c->n1 * c->n2, c->n1 * c->n3, c->nbeta);
  return need <= 160 * 1024 && lanes_ok(c->dim[0]) && lanes_ok(c->dim[1]) && lanes_ok(c->dim[2]);
}

template <typename T>
static int fused_run(pw_ctx* c, int nk, const int* iks, const T* V, int nb, const cplx<double>* const* psi,
                     const int64_t* ldpsi, cplx<double>* const* hpsi, const int64_t* ldh, unsigned flags) {
  const size_t esz = sizeof(cplx<T>);
  const unsigned terms = flags & PW_TERM_ALL;
  std::vector<FusedK> kb(nk);
  int maxcol = 1, maxyl = 1, nbeta = 0;
  for (int kk = 0; kk < nk; ++kk) {
    const KData& k = c->kd[iks[kk]];
    FusedK d{};
    d.psi = psi[kk];
    d.hpsi = hpsi[kk];
    d.ldpsi = ldpsi[kk];
    d.ldh = ldh[kk];
    d.gpos = k.d_gpos;
    d.gk2 = k.d_gk2;
    d.col_ij = k.d_col_ij;
    d.xact = k.d_xact;
    d.nbeta = (terms & PW_TERM_NLOC) ? k.nbeta : 0;
    d.beta = sizeof(T) == 8 ? k.d_beta64 : k.d_beta32;
    d.betaD = sizeof(T) == 8 ? k.d_betaD64 : k.d_betaD32;
    d.ngw = k.ngw;
    d.ncol = k.ncol;
    d.nxact = k.nxact;
    if (d.nbeta > 0 && (!d.beta || !d.betaD)) return set_error(c, PW_ERR_STATE, "op_H fused: projectors not set");
    kb[kk] = d;
    maxcol = std::max(maxcol, k.ncol);
    maxyl = std::max(maxyl, k.nxact * c->n3);
    nbeta = std::max(nbeta, d.nbeta);
  }
  const size_t lds = fused_lds(c, esz, maxcol, maxyl, nbeta);
  if (lds > 160 * 1024) return set_error(c, PW_ERR_ARG, "op_H fused: the grid does not fit in LDS");
  if ((long)nk * nb > 0x7fffffffL) return set_error(c, PW_ERR_ARG, "op_H fused: too many bands for one launch");
  const void* dptr = nullptr;
  int rc = desc_get(c, kb.data(), kb.size() * sizeof(FusedK), &dptr);
  if (rc) return rc;
  const cplx<T>* tw[3];
  for (int d = 0; d < 3; ++d)
    tw[d] = reinterpret_cast<const cplx<T>*>(sizeof(T) == 8 ? c->d_tw64[d] : c->d_tw32[d]);
  // Threads per workgroup: the workgroups that fit a CU's LDS together hold 1024 threads
  // (FP64 20^3: one 134 KB box per CU, 16 wavefronts share its line FFTs), unless the
  // tuning knob fixes the width.
  int nt = c->fused_nt;
  if (nt != 256 && nt != 512 && nt != 1024) {
    const int per_cu = std::max<int>(1, (int)((160 * 1024) / std::max<size_t>(lds, 1)));
    nt = per_cu >= 4 ? 256 : per_cu >= 2 ? 512 : 1024;
  }
  const dim3 grid((unsigned)(nk * nb));
  const FusedK* dk = reinterpret_cast<const FusedK*>(dptr);
  const int dv = (terms & PW_TERM_VLOC) ? 1 : 0, dkin = (terms & PW_TERM_K) ? 1 : 0;
  const int dnl = (terms & PW_TERM_NLOC) ? 1 : 0, acc = (flags & PW_ACCUMULATE) ? 1 : 0;
  if (nt == 1024)
    hipLaunchKernelGGL((k_op_H_fused<T, 1024>), grid, dim3(1024), lds, c->stream, dk, nb, V, c->dim[0], c->dim[1],
                       c->dim[2], tw[0], tw[1], tw[2], dv, dkin, dnl, acc);
  else if (nt == 512)
    hipLaunchKernelGGL((k_op_H_fused<T, 512>), grid, dim3(512), lds, c->stream, dk, nb, V, c->dim[0], c->dim[1],
                       c->dim[2], tw[0], tw[1], tw[2], dv, dkin, dnl, acc);
  else
    hipLaunchKernelGGL((k_op_H_fused<T, 256>), grid, dim3(256), lds, c->stream, dk, nb, V, c->dim[0], c->dim[1],
                       c->dim[2], tw[0], tw[1], tw[2], dv, dkin, dnl, acc);
  return hip_check(c, hipGetLastError(), "k_op_H_fused");
}

// ------------------------------------------------ fused op_H with line transforms in registers
// The same operator as k_op_H_fused for cubic grids N^3 with a two-pass register plan
// (N = R0 R1), built for occupancy: every 1-D transform is done by one lane on a line held in
// its registers (load N entries, two radix passes, store), and the box is kept in LDS only at
// the x values that carry data, as [z][a][y] (a: index of x among the k-point's active x,
// y stride 1, padded to an odd length).  The x-rows, which need all N x, are expanded in
// registers: zero at the inactive x, inverse transform, * V/N, forward transform, and only
// the active x are stored back.  The FP64 box of a 20^3 grid with 9 active x takes 60 KB
// instead of 134 KB, so two workgroups share a CU and hide each other's memory latency.

// X[k] = sum_n x[n] exp(S 2 pi i n k / N) of a line in registers, n = R1 n1 + n2,
// k = k1 + R0 k2: R1 transforms of length R0 over n1, twiddles exp(S 2 pi i n2 k1 / N), R0
// transforms of length R1 over n2.  tw[t] = exp(+2 pi i t / N).
template <typename T, int S, int N, int R0, int R1>
__device__ inline void fft_regs(cplx<T> (&v)[N], const LDS_AS cplx<T>* tw) {
  static_assert(R0 * R1 == N, "two-pass plan");
  cplx<T> y[R1][R0];
#pragma unroll
  for (int n2 = 0; n2 < R1; ++n2) {
    cplx<T> t[R0];
#pragma unroll
    for (int n1 = 0; n1 < R0; ++n1) t[n1] = v[R1 * n1 + n2];
    DFT<T, S, R0>::run(t);
#pragma unroll
    for (int k1 = 0; k1 < R0; ++k1) {
      if (n2 > 0 && k1 > 0) {
        const cplx<T> w = tw[n2 * k1];
        t[k1] = S > 0 ? cmul(t[k1], w) : cmulc(t[k1], w);
      }
      y[n2][k1] = t[k1];
    }
  }
#pragma unroll
  for (int k1 = 0; k1 < R0; ++k1) {
    cplx<T> t[R1];
#pragma unroll
    for (int n2 = 0; n2 < R1; ++n2) t[n2] = y[n2][k1];
    DFT<T, S, R1>::run(t);
#pragma unroll
    for (int k2 = 0; k2 < R1; ++k2) v[k1 + R0 * k2] = t[k2];
  }
}

// Lines [l0, nl) of the workgroup, one per thread: line l starts at lds[base(l)], stride es.
template <typename T, int S, int N, int R0, int R1, int NT, typename Base>
__device__ inline void reg_lines(LDS_AS cplx<T>* lds, int nl, int es, const LDS_AS cplx<T>* tw, Base base) {
  for (int l = threadIdx.x; l < nl; l += NT) {
    const int b0 = base(l);
    cplx<T> v[N];
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = lds[b0 + i * es];
    fft_regs<T, S, N, R0, R1>(v, tw);
#pragma unroll
    for (int i = 0; i < N; ++i) lds[b0 + i * es] = v[i];
  }
}

// One workgroup per (k-point, band).  LDS: box [N][nxa][YS] (sized for nxa_max), twiddles
// (N), per-wavefront partial sums of beta^H psi (NT/64 x nbeta_max), column bases (ncol_max
// ints), x -> a map (N ints).  Every global read that does not depend on the transforms is
// issued before them: the first GPT entries of the sphere per thread (their psi, box slot,
// kinetic, nonlocal and accumulated terms) stay in registers from the start to the final
// store, so the transforms run between one round of reads and one round of writes.
template <typename T, int N, int R0, int R1, int NT>
__global__ void __launch_bounds__(NT)
k_op_H_reg(const FusedK* __restrict__ kd, int nb, const T* __restrict__ V, const cplx<T>* __restrict__ twg,
           int nxa_max, int nbeta_max, int do_v, int do_k, int do_nl, int acc) {
  constexpr int YS = (N % 2 == 0) ? N + 1 : N;   // odd_stride(N): y stride of the box, and gpos's x stride
  constexpr int GPT = 2;                          // sphere entries per thread held in registers
  constexpr int NW = NT / 64;
  extern __shared__ char smem_raw[];
  const int kk = blockIdx.x / nb, b = blockIdx.x - kk * nb;
  const FusedK F = kd[kk];
  const int nxa = F.nxact;
  const int zs = nxa * YS;   // one z-slab of the box
  const int nbeta = do_nl ? F.nbeta : 0;
  LDS_AS cplx<T>* box = (LDS_AS cplx<T>*)smem_raw;
  LDS_AS cplx<T>* tw = box + (size_t)N * nxa_max * YS;
  cplx<T>* red = (cplx<T>*)(tw + N);
  int* cb = reinterpret_cast<int*>(red + (size_t)NW * (nbeta_max > 0 ? nbeta_max : 1));
  int* xa = cb + (F.ncol > 0 ? F.ncol : 1);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const cplx<double>* p = F.psi + (long)b * F.ldpsi;
  cplx<double>* h = F.hpsi + (long)b * F.ldh;
  const cplx<T>* be = reinterpret_cast<const cplx<T>*>(F.beta);
  const cplx<T>* bd = reinterpret_cast<const cplx<T>*>(F.betaD);

  // global reads of the set-up, issued together
  int gq[GPT];
  cplx<double> pv[GPT], hv[GPT];
#pragma unroll
  for (int i = 0; i < GPT; ++i) {
    const int g = tid + i * NT;
    const bool in = g < F.ngw;
    gq[i] = in ? F.gpos[g] : 0;
    pv[i] = in ? p[g] : mk(0.0, 0.0);
    hv[i] = (in && acc) ? h[g] : mk(0.0, 0.0);
  }
  const int xv = tid < nxa ? F.xact[tid] : 0;
  const cplx<T> twv = tid < N ? twg[tid] : mk((T)0, (T)0);
  // psi row g -> box slot: gpos = x + YS (y + N z)
  auto slot_of = [&](int q) {
    const int x = q % YS, r = q / YS;
    return (r / N) * zs + xa[x] * YS + (r % N);
  };
  if (do_v) {
    for (int i = tid; i < N * zs; i += NT) box[i] = mk((T)0, (T)0);
    if (tid < N) tw[tid] = twv;
  }
  // beta^H psi: each thread's entries times every projector, summed over each wavefront
  for (int j = 0; j < nbeta; ++j) {
    T sr = 0, si = 0;
#pragma unroll
    for (int i = 0; i < GPT; ++i) {
      const int g = tid + i * NT;
      if (g < F.ngw) {
        const cplx<T> bg = be[(long)j * F.ngw + g];
        const cplx<T> x = cvt<T>(pv[i]);
        sr += bg.x * x.x + bg.y * x.y;
        si += bg.x * x.y - bg.y * x.x;
      }
    }
    for (int g = tid + GPT * NT; g < F.ngw; g += NT) {
      const cplx<T> bg = be[(long)j * F.ngw + g];
      const cplx<T> x = cvt<T>(p[g]);
      sr += bg.x * x.x + bg.y * x.y;
      si += bg.x * x.y - bg.y * x.x;
    }
    for (int o = 32; o > 0; o >>= 1) {
      sr += __shfl_xor(sr, o, 64);
      si += __shfl_xor(si, o, 64);
    }
    if (lane == 0) red[wave * nbeta + j] = mk(sr, si);
  }
  __syncthreads();
  if (do_v && tid < nxa) xa[xv] = tid;
  // kinetic and nonlocal terms of the held entries (B_j summed from the wavefront partials)
#pragma unroll
  for (int i = 0; i < GPT; ++i) {
    const int g = tid + i * NT;
    if (g < F.ngw) {
      if (do_k) hv[i] = hv[i] + scal(0.5 * F.gk2[g], pv[i]);
      if (nbeta > 0) {
        cplx<T> s = mk((T)0, (T)0);
        for (int j = 0; j < nbeta; ++j) {
          cplx<T> B = red[j];
          for (int w = 1; w < NW; ++w) B = B + red[w * nbeta + j];
          s = s + cmul(bd[(long)j * F.ngw + g], B);
        }
        hv[i] = hv[i] + cvt<double>(s);
      }
    }
  }
  if (do_v) {
    __syncthreads();
    // column (i, j) -> (a, y) = (xa[i], j)
    for (int c = tid; c < F.ncol; c += NT) {
      const int cij = F.col_ij[c];
      cb[c] = xa[cij % YS] * YS + cij / YS;
    }
#pragma unroll
    for (int i = 0; i < GPT; ++i) {
      if (tid + i * NT < F.ngw) {
        gq[i] = slot_of(gq[i]);
        box[gq[i]] = cvt<T>(pv[i]);
      }
    }
    for (int g = tid + GPT * NT; g < F.ngw; g += NT) box[slot_of(F.gpos[g])] = cvt<T>(p[g]);
    __syncthreads();
    // inverse: z-lines of the sphere's columns, y-lines at the active x
    reg_lines<T, +1, N, R0, R1, NT>(box, F.ncol, zs, tw, [&](int l) { return cb[l]; });
    __syncthreads();
    reg_lines<T, +1, N, R0, R1, NT>(box, N * nxa, 1, tw, [&](int l) { return l * YS; });
    __syncthreads();
    // x-rows (y, z): expand the active x, inverse, * V/N, forward, keep the active x
    unsigned xmask = 0;
    for (int a = 0; a < nxa; ++a) xmask |= 1u << F.xact[a];
    for (int r = tid; r < N * N; r += NT) {
      const int z = r / N, y = r - z * N;
      const int b0 = z * zs + y;
      const T* vr = V + (long)r * N;
      cplx<T> v[N];
#pragma unroll
      for (int x = 0; x < N; ++x) {
        const int a = __popc(xmask & ((1u << x) - 1u));
        v[x] = ((xmask >> x) & 1u) ? box[b0 + a * YS] : mk((T)0, (T)0);
      }
      fft_regs<T, +1, N, R0, R1>(v, tw);
#pragma unroll
      for (int x = 0; x < N; ++x) v[x] = scal(vr[x], v[x]);
      fft_regs<T, -1, N, R0, R1>(v, tw);
#pragma unroll
      for (int x = 0; x < N; ++x) {
        const int a = __popc(xmask & ((1u << x) - 1u));
        if ((xmask >> x) & 1u) box[b0 + a * YS] = v[x];
      }
    }
    __syncthreads();
    // forward: y-lines, then z-lines
    reg_lines<T, -1, N, R0, R1, NT>(box, N * nxa, 1, tw, [&](int l) { return l * YS; });
    __syncthreads();
    reg_lines<T, -1, N, R0, R1, NT>(box, F.ncol, zs, tw, [&](int l) { return cb[l]; });
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < GPT; ++i) {
    const int g = tid + i * NT;
    if (g < F.ngw) h[g] = do_v ? hv[i] + cvt<double>(box[gq[i]]) : hv[i];
  }
  // entries beyond the held ones
  for (int g = tid + GPT * NT; g < F.ngw; g += NT) {
    const cplx<double> x = p[g];
    cplx<double> r = acc ? h[g] : mk(0.0, 0.0);
    if (do_v) r = r + cvt<double>(box[slot_of(F.gpos[g])]);
    if (do_k) r = r + scal(0.5 * F.gk2[g], x);
    if (nbeta > 0) {
      cplx<T> s = mk((T)0, (T)0);
      for (int j = 0; j < nbeta; ++j) {
        cplx<T> B = red[j];
        for (int w = 1; w < NW; ++w) B = B + red[w * nbeta + j];
        s = s + cmul(bd[(long)j * F.ngw + g], B);
      }
      r = r + cvt<double>(s);
    }
    h[g] = r;
  }
}

// The register-line kernel for a cubic 20^3 grid (radix plan 5 x 4); false where it does not apply.
static bool reg_plan(const pw_ctx* c) { return c->n1 == 20 && c->n2 == 20 && c->n3 == 20; }

template <typename T>
static int fused_reg_run(pw_ctx* c, int nk, const int* iks, const T* V, int nb, const cplx<double>* const* psi,
                         const int64_t* ldpsi, cplx<double>* const* hpsi, const int64_t* ldh, unsigned flags) {
  constexpr int N = 20, YS = 21;
  const unsigned terms = flags & PW_TERM_ALL;
  std::vector<FusedK> kb(nk);
  int nxa_max = 1, ncol_max = 1, nbeta_max = 0;
  for (int kk = 0; kk < nk; ++kk) {
    const KData& k = c->kd[iks[kk]];
    FusedK d{};
    d.psi = psi[kk];
    d.hpsi = hpsi[kk];
    d.ldpsi = ldpsi[kk];
    d.ldh = ldh[kk];
    d.gpos = k.d_gpos;
    d.gk2 = k.d_gk2;
    d.col_ij = k.d_col_ij;
    d.xact = k.d_xact;
    d.nbeta = (terms & PW_TERM_NLOC) ? k.nbeta : 0;
    d.beta = sizeof(T) == 8 ? k.d_beta64 : k.d_beta32;
    d.betaD = sizeof(T) == 8 ? k.d_betaD64 : k.d_betaD32;
    d.ngw = k.ngw;
    d.ncol = k.ncol;
    d.nxact = k.nxact;
    if (d.nbeta > 0 && (!d.beta || !d.betaD)) return set_error(c, PW_ERR_STATE, "op_H fused: projectors not set");
    if (k.nxact < 1 || k.nxact > N) return set_error(c, PW_ERR_STATE, "op_H fused: bad active-x count");
    kb[kk] = d;
    nxa_max = std::max(nxa_max, k.nxact);
    ncol_max = std::max(ncol_max, k.ncol);
    nbeta_max = std::max(nbeta_max, d.nbeta);
  }
  const int nw = (c->fused_nt == 512 ? 512 : 256) / 64;
  const size_t lds = ((size_t)N * nxa_max * YS + N + (size_t)nw * std::max(nbeta_max, 1)) * sizeof(cplx<T>) +
                     (size_t)(ncol_max + N) * sizeof(int);
  if (lds > 160 * 1024) return set_error(c, PW_ERR_ARG, "op_H fused: the compressed box does not fit in LDS");
  if ((long)nk * nb > 0x7fffffffL) return set_error(c, PW_ERR_ARG, "op_H fused: too many bands for one launch");
  const void* dptr = nullptr;
  int rc = desc_get(c, kb.data(), kb.size() * sizeof(FusedK), &dptr);
  if (rc) return rc;
  const cplx<T>* tw = reinterpret_cast<const cplx<T>*>(sizeof(T) == 8 ? c->d_tw64[0] : c->d_tw32[0]);
  const FusedK* dk = reinterpret_cast<const FusedK*>(dptr);
  const int dv = (terms & PW_TERM_VLOC) ? 1 : 0, dkin = (terms & PW_TERM_K) ? 1 : 0;
  const int dnl = (terms & PW_TERM_NLOC) ? 1 : 0, acc = (flags & PW_ACCUMULATE) ? 1 : 0;
  const dim3 grid((unsigned)(nk * nb));
  // 256 threads (the 400 x-rows in two rounds) unless the tuning knob asks for 512
  if (c->fused_nt == 512)
    hipLaunchKernelGGL((k_op_H_reg<T, N, 5, 4, 512>), grid, dim3(512), lds, c->stream, dk, nb, V, tw, nxa_max,
                       nbeta_max, dv, dkin, dnl, acc);
  else
    hipLaunchKernelGGL((k_op_H_reg<T, N, 5, 4, 256>), grid, dim3(256), lds, c->stream, dk, nb, V, tw, nxa_max,
                       nbeta_max, dv, dkin, dnl, acc);
  return hip_check(c, hipGetLastError(), "k_op_H_reg");
}

int op_H_fused_batch(pw_ctx* c, int nk, const int* iks, int ispin, int nb, const cplx<double>* const* psi,
                     const int64_t* ldpsi, cplx<double>* const* hpsi, const int64_t* ldh, unsigned flags) {
  if (!op_H_fused_fits(c, (flags & PW_MIXED) != 0)) return set_error(c, PW_ERR_ARG, "op_H fused: grid too large");
  const bool reg = c->fused_batch == 1 && reg_plan(c);
  if (flags & PW_MIXED) {
    if (reg) return fused_reg_run<float>(c, nk, iks, c->d_v32[ispin], nb, psi, ldpsi, hpsi, ldh, flags);
    return fused_run<float>(c, nk, iks, c->d_v32[ispin], nb, psi, ldpsi, hpsi, ldh, flags);
  }
  if (reg) return fused_reg_run<double>(c, nk, iks, c->d_v64[ispin], nb, psi, ldpsi, hpsi, ldh, flags);
  return fused_run<double>(c, nk, iks, c->d_v64[ispin], nb, psi, ldpsi, hpsi, ldh, flags);
}

int vloc_apply_batch(pw_ctx* c, int nk, const int* iks, int ispin, int nb, const cplx<double>* const* psi,
                     const int64_t* ldpsi, cplx<double>* const* hpsi, const int64_t* ldh, unsigned flags) {
  const int kin = (flags & PW_TERM_K) ? 1 : 0;
  const int acc = (flags & PW_ACCUMULATE) ? 1 : 0;
  if (flags & PW_MIXED)
    return vloc_batch_run<float>(c, nk, iks, c->d_v32[ispin], nb, psi, ldpsi, hpsi, ldh, kin, acc);
  return vloc_batch_run<double>(c, nk, iks, c->d_v64[ispin], nb, psi, ldpsi, hpsi, ldh, kin, acc);
}

int kinetic_apply(pw_ctx* c, int ik, int nb, const cplx<double>* psi, long ldpsi, cplx<double>* hpsi,
                  long ldh, unsigned flags) {
  KData& k = c->kd[ik];
  const long tot = (long)k.ngw * nb;
  const int blocks = (int)std::min<long>((tot + 255) / 256, 4096);
  hipLaunchKernelGGL(k_kinetic, dim3(blocks), dim3(256), 0, c->stream, psi, ldpsi, hpsi, ldh, k.d_gk2,
                     k.ngw, nb, (flags & PW_ACCUMULATE) ? 1 : 0);
  return hip_check(c, hipGetLastError(), "k_kinetic");
}

int kprec_apply(pw_ctx* c, int ik, int nb, const cplx<double>* in, long ldin, cplx<double>* out,
                long ldout) {
  KData& k = c->kd[ik];
  const long tot = (long)k.ngw * nb;
  const int blocks = (int)std::min<long>((tot + 255) / 256, 4096);
  hipLaunchKernelGGL(k_kprec, dim3(blocks), dim3(256), 0, c->stream, in, ldin, out, ldout, k.d_gk2, k.ngw,
                     nb);
  return hip_check(c, hipGetLastError(), "k_kprec");
}

template <typename T>
static int rhoe_run(pw_ctx* c, int ik, int nb, const cplx<double>* psi, long ldpsi, const double* d_w) {
  KData& k = c->kd[ik];
  const size_t esz = sizeof(cplx<T>);
  const int C = cols_per_block(c, esz);
  const int np = c->n1 * c->n2;
  size_t plane_bytes = plane_lds(c, esz, k.nxact);
  int rl = 0;
  // the padded-plane kernel keeps each thread's share of the plane (<= RHO_REGS points) in
  // registers; planes beyond that take the compressed kernel, which accumulates in HBM
  const bool compressed = plane_bytes > 160 * 1024 || c->plane_compressed != 0 ||
                          np > RHO_REGS * PlaneNT<T>::value;
  if (compressed) {
    rl = plane_c_rows(c, esz, k.nxact, PlaneNT<T>::value, &plane_bytes);
    if (rl < 1) return set_error(c, PW_ERR_ARG, "z-plane too large for calc_rhoe kernel");
  }
  if (!lanes_ok(c->dim[0]) || !lanes_ok(c->dim[1]) || !lanes_ok(c->dim[2]))
    return set_error(c, PW_ERR_ARG, "FFT line too long for one wavefront");
  const int bb = std::max(1, std::min(nb, c->band_batch));
  int rc = ensure(c, &c->d_work, &c->work_bytes, (size_t)bb * c->n3 * k.ncol * esz);
  if (rc) return rc;
  // band groups: about one round of resident plane workgroups (3 per CU on 256 CUs) over the
  // n3 planes; each group's partial plane is reduced afterwards in a fixed order
  int ngroup = std::max(1, std::min(bb, (768 + c->n3 - 1) / c->n3));
  rc = ensure(c, &c->d_rpart, &c->rpart_bytes, (size_t)ngroup * c->N * sizeof(double));
  if (rc) return rc;
  cplx<T>* W = reinterpret_cast<cplx<T>*>(c->d_work);
  const cplx<T>* tw[3];
  for (int d = 0; d < 3; ++d)
    tw[d] = reinterpret_cast<const cplx<T>*>(sizeof(T) == 8 ? c->d_tw64[d] : c->d_tw32[d]);
  for (int b0 = 0; b0 < nb; b0 += bb) {
    const int nbb = std::min(bb, nb - b0);
    const int ng = std::max(1, std::min(nbb, ngroup));
    const int bpg = (nbb + ng - 1) / ng;
    const int ngu = (nbb + bpg - 1) / bpg;
    hipLaunchKernelGGL(k_col_ifft<T>, dim3(col_grid(k.ncol, C, nbb)), dim3(COL_NT), col_lds(c, C, esz),
                       c->stream, psi, ldpsi, b0, W, k.d_ent_igw, k.d_ent_pos, k.d_col_ptr, k.ncol, C, nbb,
                       c->dim[2], tw[2]);
    ProfScope ps(c, PW_PROF_PLANE_RHOE, (double)nbb * c->n3 * k.ncol * esz + (double)ngu * c->N * 8.0);
    if (compressed)
      hipLaunchKernelGGL(k_plane_rhoe_c<T>, dim3(c->n3, ngu), dim3(PlaneNT<T>::value), plane_bytes, c->stream, W,
                         d_w + b0, nbb, bpg, reinterpret_cast<double*>(c->d_rpart), c->N, k.d_col_cy, k.ncol,
                         k.d_xact, k.nxact, c->dim[0], c->dim[1], tw[0], tw[1], c->n3, rl);
    else {
      // the plane's points over the threads, at most RHO_REGS each (np <= RHO_REGS * PlaneNT here)
      int nt = plane_threads<T>(c, k.nxact);
      while ((np + nt - 1) / nt > RHO_REGS) nt += 64;
      const int nr = (np + nt - 1) / nt;
      auto kern = nr <= 4 ? k_plane_rhoe<T, 4> : nr <= 6 ? k_plane_rhoe<T, 6> : nr <= 8 ? k_plane_rhoe<T, 8>
                : nr <= 12 ? k_plane_rhoe<T, 12> : k_plane_rhoe<T, RHO_REGS>;
      hipLaunchKernelGGL(kern, dim3(c->n3, ngu), dim3(nt), plane_bytes, c->stream, W, d_w + b0, nbb,
                         bpg, reinterpret_cast<double*>(c->d_rpart), c->N, k.d_col_ij, k.ncol, k.d_xact, k.nxact,
                         c->dim[0], c->dim[1], tw[0], tw[1], c->n3);
    }
    hipLaunchKernelGGL(k_rho_reduce, dim3(1024), dim3(256), 0, c->stream, c->d_rho,
                       reinterpret_cast<const double*>(c->d_rpart), ngu, c->N, 1.0 / c->omega);
  }
  return hip_check(c, hipGetLastError(), "rhoe kernels");
}

int g2r_apply(pw_ctx* c, KData& k, int nf, const cplx<double>* fG, long ldf, double* out, double scale,
              int acc) {
  const size_t esz = sizeof(cplx<double>);
  const int C = cols_per_block(c, esz);
  size_t plane_bytes = plane_lds(c, esz, k.nxact);
  int rl = 0;
  const bool compressed = plane_bytes > 160 * 1024 || c->plane_compressed != 0;
  PlaneGP gp;
  bool global_plane = false;
  if (compressed) {
    rl = plane_c_rows(c, esz, k.nxact, PlaneNT<double>::value, &plane_bytes);
    if (rl < 1 || c->plane_compressed == 2) {
      if (!plane_gp(c, esz, k.nxact, PlaneNT<double>::value, &gp))
        return set_error(c, PW_ERR_ARG, "z-plane too large for the G->R kernel");
      global_plane = true;
    }
  }
  if (!lanes_ok(c->dim[0]) || !lanes_ok(c->dim[1]) || !lanes_ok(c->dim[2]))
    return set_error(c, PW_ERR_ARG, "FFT line too long for one wavefront");
  const int bb = std::max(1, std::min(nf, c->band_batch));
  int rc = ensure(c, &c->d_work, &c->work_bytes, (size_t)bb * c->n3 * k.ncol * esz);
  if (rc) return rc;
  cplx<double>* W = reinterpret_cast<cplx<double>*>(c->d_work);
  cplx<double>* Yg = nullptr;
  if (global_plane) {
    rc = ensure(c, &c->d_yglob, &c->yglob_bytes, (size_t)bb * c->n3 * k.nxact * odd_stride(c->n2) * esz);
    if (rc) return rc;
    Yg = reinterpret_cast<cplx<double>*>(c->d_yglob);
  }
  const cplx<double>* tw[3];
  for (int d = 0; d < 3; ++d) tw[d] = reinterpret_cast<const cplx<double>*>(c->d_tw64[d]);
  for (int b0 = 0; b0 < nf; b0 += bb) {
    const int nbb = std::min(bb, nf - b0);
    hipLaunchKernelGGL(k_col_ifft<double>, dim3(col_grid(k.ncol, C, nbb)), dim3(COL_NT), col_lds(c, C, esz),
                       c->stream, fG, ldf, b0, W, k.d_ent_igw, k.d_ent_pos, k.d_col_ptr, k.ncol, C, nbb,
                       c->dim[2], tw[2]);
    double* o = out + (long)b0 * c->N;
    if (global_plane) {
      const dim3 gy(c->n3, nbb, (k.nxact + gp.glines - 1) / gp.glines);
      hipLaunchKernelGGL((k_yfft_gp<double, +1>), gy, dim3(PlaneNT<double>::value), gp.ybytes, c->stream, W, Yg,
                         k.d_col_cy, k.ncol, k.nxact, gp.glines, c->dim[1], tw[1], c->n3);
      if (acc)
        hipLaunchKernelGGL((k_xrows_gp<double, 2>), dim3(c->n3, nbb), dim3(PlaneNT<double>::value), gp.xbytes,
                           c->stream, Yg, nullptr, o, c->N, scale, k.d_xact, k.nxact, c->dim[0], c->n2, tw[0],
                           c->n3, gp.rl);
      else
        hipLaunchKernelGGL((k_xrows_gp<double, 1>), dim3(c->n3, nbb), dim3(PlaneNT<double>::value), gp.xbytes,
                           c->stream, Yg, nullptr, o, c->N, scale, k.d_xact, k.nxact, c->dim[0], c->n2, tw[0],
                           c->n3, gp.rl);
    } else if (compressed)
      hipLaunchKernelGGL(k_plane_g2r_c, dim3(c->n3, nbb), dim3(PlaneNT<double>::value), plane_bytes, c->stream, W,
                         o, c->N, scale, acc, k.d_col_cy, k.ncol, k.d_xact, k.nxact, c->dim[0], c->dim[1], tw[0],
                         tw[1], c->n3, rl);
    else
      hipLaunchKernelGGL(k_plane_g2r, dim3(c->n3, nbb), dim3(plane_threads<double>(c, k.nxact)), plane_bytes, c->stream, W, o,
                         c->N, scale, acc, k.d_col_ij, k.ncol, k.d_xact, k.nxact, c->dim[0], c->dim[1], tw[0],
                         tw[1], c->n3);
  }
  return hip_check(c, hipGetLastError(), "G->R kernels");
}

int rhoe_accum(pw_ctx* c, int ik, int nb, const cplx<double>* psi, long ldpsi, const double* d_w,
               unsigned flags) {
  if (flags & PW_MIXED) return rhoe_run<float>(c, ik, nb, psi, ldpsi, d_w);
  return rhoe_run<double>(c, ik, nb, psi, ldpsi, d_w);
}

}  // namespace pw
