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
#define PW_PLANE_NT 1024
#endif
constexpr int COL_NT = PW_COL_NT;
constexpr int COL_ME = 12;
constexpr int PLANE_NT = PW_PLANE_NT;
constexpr int PLANE_ME = 12;

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

// One Stockham pass over `nlines` lines of length n held in LDS.  Element r of line L is
// at lds[base(L) + r*es], base(L) = lbase ? lbase[L] : L*lstride.  Ns = product of the
// radices already applied.  tw[t] = exp(+2 pi i t / n).
template <typename T, int S, int R, int ME>
__device__ void fft_pass(cplx<T>* lds, int nlines, const int* lbase, int lstride, int es, int n,
                         int Ns, const cplx<T>* __restrict__ tw) {
  constexpr int MB = (ME / R) > 0 ? (ME / R) : 1;
  const int nthr = blockDim.x, tid = threadIdx.x;
  const int nbl = n / R;
  const int nbf = nlines * nbl;
  const int twstep = n / (Ns * R);
  cplx<T> v[MB][R];
  int dst[MB];
#pragma unroll
  for (int s = 0; s < MB; ++s) {
    const int b = tid + s * nthr;
    dst[s] = -1;
    if (b < nbf) {
      const int line = b / nbl, j = b - line * nbl;
      const int base = lbase ? lbase[line] : line * lstride;
#pragma unroll
      for (int r = 0; r < R; ++r) v[s][r] = lds[base + (j + r * nbl) * es];
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
  __syncthreads();
#pragma unroll
  for (int s = 0; s < MB; ++s) {
    if (dst[s] >= 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) lds[dst[s] + r * Ns * es] = v[s][r];
    }
  }
  __syncthreads();
}

template <typename T, int S, int ME>
__device__ void fft_lines(cplx<T>* lds, int nlines, const int* lbase, int lstride, int es,
                          const FftDim& d, const cplx<T>* __restrict__ tw) {
  int Ns = 1;
  for (int p = 0; p < d.npass; ++p) {
    const int R = d.radix[p];
    switch (R) {
      case 2: fft_pass<T, S, 2, ME>(lds, nlines, lbase, lstride, es, d.n, Ns, tw); break;
      case 3: fft_pass<T, S, 3, ME>(lds, nlines, lbase, lstride, es, d.n, Ns, tw); break;
      case 4: fft_pass<T, S, 4, ME>(lds, nlines, lbase, lstride, es, d.n, Ns, tw); break;
      case 5: fft_pass<T, S, 5, ME>(lds, nlines, lbase, lstride, es, d.n, Ns, tw); break;
      case 8: fft_pass<T, S, 8, ME>(lds, nlines, lbase, lstride, es, d.n, Ns, tw); break;
      default: fft_pass<T, S, 9, ME>(lds, nlines, lbase, lstride, es, d.n, Ns, tw); break;
    }
    Ns *= R;
  }
}

// copy a twiddle table into LDS (caller synchronises)
template <typename T>
__device__ inline void stage_tw(cplx<T>* dst, const cplx<T>* __restrict__ src, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

// ----------------------------------------------------------------------------- kernels

// 1. scatter + inverse z-FFT of the columns [c0, c0+C) of band (b0 + blockIdx.y).
template <typename T>
__global__ void __launch_bounds__(COL_NT)
k_col_ifft(const cplx<double>* __restrict__ psi, long ldpsi, int b0, cplx<T>* __restrict__ W,
           const int* __restrict__ ent_igw, const int* __restrict__ ent_pos,
           const int* __restrict__ col_ptr, int ncol, int C, FftDim dz,
           const cplx<T>* __restrict__ twz) {
  extern __shared__ char smem_raw[];
  cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
  const int n3 = dz.n, cs = odd_stride(n3);
  cplx<T>* tw = lds + C * cs;
  const int bl = blockIdx.y;
  const int c0 = blockIdx.x * C;
  const int cn = min(C, ncol - c0);
  const int tid = threadIdx.x, nthr = blockDim.x;
  for (int i = tid; i < cn * cs; i += nthr) lds[i] = mk((T)0, (T)0);
  stage_tw(tw, twz, n3);
  __syncthreads();
  const int e0 = col_ptr[c0], e1 = col_ptr[c0 + cn];
  const cplx<double>* p = psi + (long)(b0 + bl) * ldpsi;
  const int off = c0 * cs;
  for (int e = e0 + tid; e < e1; e += nthr) lds[ent_pos[e] - off] = cvt<T>(p[ent_igw[e]]);
  __syncthreads();
  fft_lines<T, +1, COL_ME>(lds, cn, nullptr, cs, 1, dz, tw);
  cplx<T>* w = W + (long)bl * n3 * ncol;
  for (int i = tid; i < cn * n3; i += nthr) {
    const int cl = i % cn, z = i / cn;
    w[(long)z * ncol + c0 + cl] = lds[cl * cs + z];
  }
}

// 2. per (z-plane, band): y/x inverse FFT, * V/N, x/y forward FFT.  In place on W.
template <typename T>
__global__ void __launch_bounds__(PLANE_NT)
k_plane_vloc(cplx<T>* __restrict__ W, const T* __restrict__ V, const int* __restrict__ col_ij,
             int ncol, const int* __restrict__ xact, int nxact, FftDim dx, FftDim dy,
             const cplx<T>* __restrict__ twx, const cplx<T>* __restrict__ twy, int n3) {
  extern __shared__ char smem_raw[];
  cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
  const int n1 = dx.n, n2 = dy.n, np = n1 * n2, ps = odd_stride(n1);
  cplx<T>* tx = lds + ps * n2;
  cplx<T>* ty = tx + n1;
  const int z = blockIdx.x, bl = blockIdx.y;
  const int tid = threadIdx.x, nthr = blockDim.x;
  for (int i = tid; i < ps * n2; i += nthr) lds[i] = mk((T)0, (T)0);
  stage_tw(tx, twx, n1);
  stage_tw(ty, twy, n2);
  __syncthreads();
  cplx<T>* w = W + ((long)bl * n3 + z) * ncol;
  for (int c = tid; c < ncol; c += nthr) lds[col_ij[c]] = w[c];
  __syncthreads();
  fft_lines<T, +1, PLANE_ME>(lds, nxact, xact, 0, ps, dy, ty);   // y lines at active x
  fft_lines<T, +1, PLANE_ME>(lds, n2, nullptr, ps, 1, dx, tx);   // x rows
  const T* v = V + (long)z * np;
  for (int i = tid; i < np; i += nthr) {
    const int j = i / n1, x = i - j * n1;
    lds[x + ps * j] = scal(v[i], lds[x + ps * j]);
  }
  __syncthreads();
  fft_lines<T, -1, PLANE_ME>(lds, n2, nullptr, ps, 1, dx, tx);
  fft_lines<T, -1, PLANE_ME>(lds, nxact, xact, 0, ps, dy, ty);
  for (int c = tid; c < ncol; c += nthr) w[c] = lds[col_ij[c]];
}

// 3. forward z-FFT of the columns, gather the sphere, add kinetic, write Hpsi.
template <typename T>
__global__ void __launch_bounds__(COL_NT)
k_col_fft(const cplx<T>* __restrict__ W, const cplx<double>* __restrict__ psi, long ldpsi,
          cplx<double>* __restrict__ hpsi, long ldh, int b0, const int* __restrict__ ent_igw,
          const int* __restrict__ ent_pos, const int* __restrict__ col_ptr, int ncol, int C,
          FftDim dz, const cplx<T>* __restrict__ twz, const double* __restrict__ gk2, int kin,
          int acc) {
  extern __shared__ char smem_raw[];
  cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
  const int n3 = dz.n, cs = odd_stride(n3);
  cplx<T>* tw = lds + C * cs;
  const int bl = blockIdx.y;
  const int c0 = blockIdx.x * C;
  const int cn = min(C, ncol - c0);
  const int tid = threadIdx.x, nthr = blockDim.x;
  const cplx<T>* w = W + (long)bl * n3 * ncol;
  for (int i = tid; i < cn * n3; i += nthr) {
    const int cl = i % cn, z = i / cn;
    lds[cl * cs + z] = w[(long)z * ncol + c0 + cl];
  }
  stage_tw(tw, twz, n3);
  __syncthreads();
  fft_lines<T, -1, COL_ME>(lds, cn, nullptr, cs, 1, dz, tw);
  const int e0 = col_ptr[c0], e1 = col_ptr[c0 + cn];
  const int off = c0 * cs;
  const cplx<double>* p = psi + (long)(b0 + bl) * ldpsi;
  cplx<double>* h = hpsi + (long)(b0 + bl) * ldh;
  for (int e = e0 + tid; e < e1; e += nthr) {
    const int g = ent_igw[e];
    cplx<double> r = cvt<double>(lds[ent_pos[e] - off]);
    if (kin) r = r + scal(0.5 * gk2[g], p[g]);
    if (acc) r = r + h[g];
    h[g] = r;
  }
}

// rho planes: per (z-plane, band group): y/x inverse FFT of each band and
// sum_b w_b |psi_b(r)|^2 accumulated in an LDS plane of doubles; one partial plane per group.
template <typename T>
__global__ void __launch_bounds__(PLANE_NT)
k_plane_rhoe(const cplx<T>* __restrict__ W, const double* __restrict__ wts, int nb, int bpg,
             double* __restrict__ rpart, long N, const int* __restrict__ col_ij, int ncol,
             const int* __restrict__ xact, int nxact, FftDim dx, FftDim dy,
             const cplx<T>* __restrict__ twx, const cplx<T>* __restrict__ twy, int n3) {
  extern __shared__ char smem_raw[];
  cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
  const int n1 = dx.n, n2 = dy.n, np = n1 * n2, ps = odd_stride(n1);
  cplx<T>* tx = lds + ps * n2;
  cplx<T>* ty = tx + n1;
  double* acc = reinterpret_cast<double*>(ty + n2);
  const int z = blockIdx.x, g = blockIdx.y;
  const int tid = threadIdx.x, nthr = blockDim.x;
  stage_tw(tx, twx, n1);
  stage_tw(ty, twy, n2);
  for (int i = tid; i < np; i += nthr) acc[i] = 0.0;
  const int bend = min(nb, (g + 1) * bpg);
  for (int b = g * bpg; b < bend; ++b) {
    for (int i = tid; i < ps * n2; i += nthr) lds[i] = mk((T)0, (T)0);
    __syncthreads();
    const cplx<T>* w = W + ((long)b * n3 + z) * ncol;
    for (int c = tid; c < ncol; c += nthr) lds[col_ij[c]] = w[c];
    __syncthreads();
    fft_lines<T, +1, PLANE_ME>(lds, nxact, xact, 0, ps, dy, ty);
    fft_lines<T, +1, PLANE_ME>(lds, n2, nullptr, ps, 1, dx, tx);
    const double wb = wts[b];
    for (int i = tid; i < np; i += nthr) {
      const int j = i / n1, x = i - j * n1;
      const double re = (double)lds[x + ps * j].x, im = (double)lds[x + ps * j].y;
      acc[i] += wb * (re * re + im * im);
    }
    __syncthreads();
  }
  double* out = rpart + (long)g * N + (long)z * np;
  for (int i = tid; i < np; i += nthr) out[i] = acc[i];
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

static int per_thread(int me, int R) { return (me / R) > 0 ? (me / R) : 1; }

static int cols_per_block(const pw_ctx* c, size_t esz) {
  // keep the column tile at <= 48 KiB of LDS and within one pass's register budget:
  // C * (n3 / R) butterflies must fit per_thread(COL_ME, R) per thread for every radix R.
  const FftDim& d = c->dim[2];
  int C = (int)((48 * 1024) / (esz * odd_stride(d.n)));
  for (int p = 0; p < d.npass; ++p) {
    const int R = d.radix[p];
    const int cap = per_thread(COL_ME, R) * COL_NT / (d.n / R);
    if (C > cap) C = cap;
  }
  if (C < 1) C = 1;
  return C;
}

static bool pass_fits(const FftDim& d, int nlines, int nt, int me) {
  for (int p = 0; p < d.npass; ++p) {
    const int R = d.radix[p];
    if ((long)nlines * (d.n / R) > (long)per_thread(me, R) * nt) return false;
  }
  return true;
}

static size_t plane_lds(const pw_ctx* c, size_t esz) {
  return ((size_t)odd_stride(c->n1) * c->n2 + c->n1 + c->n2) * esz;
}
static size_t col_lds(const pw_ctx* c, int C, size_t esz) { return ((size_t)C * odd_stride(c->n3) + c->n3) * esz; }

template <typename T>
static int vloc_run(pw_ctx* c, int ik, int ispin, int nb, const cplx<double>* psi, long ldpsi,
                    cplx<double>* hpsi, long ldh, int kin, int acc) {
  KData& k = c->kd[ik];
  const size_t esz = sizeof(cplx<T>);
  const int C = cols_per_block(c, esz);
  const int nchunk = (k.ncol + C - 1) / C;
  const size_t plane_bytes = plane_lds(c, esz);
  if (plane_bytes > 160 * 1024)
    return set_error(c, PW_ERR_ARG, "z-plane does not fit in LDS for this precision");
  if (!pass_fits(c->dim[2], C, COL_NT, COL_ME) || !pass_fits(c->dim[0], c->n2, PLANE_NT, PLANE_ME) ||
      !pass_fits(c->dim[1], k.nxact, PLANE_NT, PLANE_ME))
    return set_error(c, PW_ERR_ARG, "FFT plane too large for one workgroup");
  const int bb = std::max(1, std::min(nb, c->band_batch));
  const size_t wbytes = (size_t)bb * c->n3 * k.ncol * esz;
  int rc = ensure(c, &c->d_work, &c->work_bytes, wbytes);
  if (rc) return rc;
  cplx<T>* W = reinterpret_cast<cplx<T>*>(c->d_work);
  const cplx<T>* tw[3];
  for (int d = 0; d < 3; ++d)
    tw[d] = reinterpret_cast<const cplx<T>*>(sizeof(T) == 8 ? c->d_tw64[d] : c->d_tw32[d]);
  const T* V = sizeof(T) == 8 ? reinterpret_cast<const T*>(c->d_v64[ispin])
                              : reinterpret_cast<const T*>(c->d_v32[ispin]);
  for (int b0 = 0; b0 < nb; b0 += bb) {
    const int nbb = std::min(bb, nb - b0);
    dim3 g1(nchunk, nbb);
    // algorithmic HBM bytes: psi in (16 B/coef) + index maps, work array out
    const double wb = (double)nbb * c->n3 * k.ncol * esz;
    const double psib = (double)nbb * k.ngw * 16.0;
    {
      ProfScope ps(c, PW_PROF_COL_IFFT, psib + wb + k.ngw * 8.0);
      hipLaunchKernelGGL(k_col_ifft<T>, g1, dim3(COL_NT), col_lds(c, C, esz), c->stream, psi, ldpsi, b0,
                         W, k.d_ent_igw, k.d_ent_pos, k.d_col_ptr, k.ncol, C, c->dim[2], tw[2]);
    }
    dim3 g2(c->n3, nbb);
    {
      ProfScope ps(c, PW_PROF_PLANE_VLOC, 2.0 * wb + (double)c->N * sizeof(T));
      hipLaunchKernelGGL(k_plane_vloc<T>, g2, dim3(PLANE_NT), plane_bytes, c->stream, W, V, k.d_col_ij,
                         k.ncol, k.d_xact, k.nxact, c->dim[0], c->dim[1], tw[0], tw[1], c->n3);
    }
    {
      ProfScope ps(c, PW_PROF_COL_FFT, wb + psib * ((kin ? 1 : 0) + (acc ? 1 : 0) + 1) + k.ngw * 16.0);
      hipLaunchKernelGGL(k_col_fft<T>, g1, dim3(COL_NT), col_lds(c, C, esz), c->stream, W, psi, ldpsi,
                         hpsi, ldh, b0, k.d_ent_igw, k.d_ent_pos, k.d_col_ptr, k.ncol, C, c->dim[2],
                         tw[2], k.d_gk2, kin, acc);
    }
  }
  return hip_check(c, hipGetLastError(), "vloc kernels");
}

int vloc_apply(pw_ctx* c, int ik, int ispin, int nb, const cplx<double>* psi, long ldpsi,
               cplx<double>* hpsi, long ldh, unsigned flags) {
  const int kin = (flags & PW_TERM_K) ? 1 : 0;
  const int acc = (flags & PW_ACCUMULATE) ? 1 : 0;
  if (flags & PW_MIXED) return vloc_run<float>(c, ik, ispin, nb, psi, ldpsi, hpsi, ldh, kin, acc);
  return vloc_run<double>(c, ik, ispin, nb, psi, ldpsi, hpsi, ldh, kin, acc);
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
  const int nchunk = (k.ncol + C - 1) / C;
  const int np = c->n1 * c->n2;
  const size_t plane_bytes = plane_lds(c, esz) + (size_t)np * sizeof(double);
  if (plane_bytes > 160 * 1024)
    return set_error(c, PW_ERR_ARG, "z-plane too large for calc_rhoe kernel");
  if (!pass_fits(c->dim[2], C, COL_NT, COL_ME) || !pass_fits(c->dim[0], c->n2, PLANE_NT, PLANE_ME) ||
      !pass_fits(c->dim[1], k.nxact, PLANE_NT, PLANE_ME))
    return set_error(c, PW_ERR_ARG, "FFT plane too large for one workgroup");
  const int bb = std::max(1, std::min(nb, c->band_batch));
  int rc = ensure(c, &c->d_work, &c->work_bytes, (size_t)bb * c->n3 * k.ncol * esz);
  if (rc) return rc;
  int ngroup = std::max(1, std::min(bb, (512 + c->n3 - 1) / c->n3));
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
    hipLaunchKernelGGL(k_col_ifft<T>, dim3(nchunk, nbb), dim3(COL_NT), col_lds(c, C, esz), c->stream,
                       psi, ldpsi, b0, W, k.d_ent_igw, k.d_ent_pos, k.d_col_ptr, k.ncol, C, c->dim[2],
                       tw[2]);
    ProfScope ps(c, PW_PROF_PLANE_RHOE, (double)nbb * c->n3 * k.ncol * esz + (double)ngu * c->N * 8.0);
    hipLaunchKernelGGL(k_plane_rhoe<T>, dim3(c->n3, ngu), dim3(PLANE_NT), plane_bytes, c->stream, W,
                       d_w + b0, nbb, bpg, reinterpret_cast<double*>(c->d_rpart), c->N, k.d_col_ij, k.ncol,
                       k.d_xact, k.nxact, c->dim[0], c->dim[1], tw[0], tw[1], c->n3);
    hipLaunchKernelGGL(k_rho_reduce, dim3(1024), dim3(256), 0, c->stream, c->d_rho,
                       reinterpret_cast<const double*>(c->d_rpart), ngu, c->N, 1.0 / c->omega);
  }
  return hip_check(c, hipGetLastError(), "rhoe kernels");
}

int rhoe_accum(pw_ctx* c, int ik, int nb, const cplx<double>* psi, long ldpsi, const double* d_w,
               unsigned flags) {
  if (flags & PW_MIXED) return rhoe_run<float>(c, ik, nb, psi, ldpsi, d_w);
  return rhoe_run<double>(c, ik, nb, psi, ldpsi, d_w);
}

}  // namespace pw
