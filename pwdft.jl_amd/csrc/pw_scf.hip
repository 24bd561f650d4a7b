// The density side of an SCF iteration on the device, so that rho and V stay in HBM between
// the eigensolver sweeps: update! (Hamiltonian.jl:510-545) with Poisson_solve!
// (Poisson_solve.jl:22-39), Slater + VWN (XC_funcs/XC_x_slater.jl, XC_funcs/XC_c_vwn.jl) or
// PBE (GGA_PBE_internal.jl:29-55, 119-166 with XC_funcs/XC_x_pbe.jl, XC_c_pw.jl,
// XC_c_pbe.jl), V_Total = V_Ps_loc + V_H + V_xc, the local energy terms of calc_energies
// (calc_energies.jl:245-333), simple density mixing with the eps() clamp
// (KS_solve_SCF.jl:240-281) and symmetrize_rhoe! (RhoeSymmetrizer.jl:111-236).  FP64 only.
//
// Fields on the rho sphere move between the grid and G space through the sphere-pruned FFT
// passes of op_V_loc: R->G is that pipeline run with psi = delta_{G,0} and V = f, whose
// forward pass returns sum_r f(r) e^{-i G.r} (R_to_G, unnormalised); G->R is its inverse
// column pass followed by an inverse plane pass that writes Re f(r) (g2r_apply).  Grid
// components outside the rho sphere are kept wherever the reference keeps them:
//   Poisson: V_H = rho + G_to_R[d] with d(G) = (4 pi / G^2 - 1) rho(G), d(0) = -rho(0) on the
//            sphere -- the reference's ctmp, whose entries outside the sphere stay unscaled;
//   symmetrisation: rho += G_to_R[rho_sym(G) - rho(G)] on the sphere.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "pw_internal.h"

namespace pw {

constexpr int SNT = 256;     // threads per block of the grid kernels
constexpr int SRED = 512;    // blocks of the energy reductions

// grid scratch slots (N doubles each)
enum { SG_GRAD = 0, SG_H = 3, SG_V1 = 6, SG_EPS = 7, SG_VH = 8, SG_VXC = 9, SG_VTOT = 10, SG_COUNT = 11 };
// sphere scratch slots (ng complex each)
enum { SS_DELTA = 0, SS_RHOG = 1, SS_T = 2, SS_HG = 5, SS_DIV = 8, SS_COUNT = 9 };

static inline int grid_blocks(long n) { return (int)std::min<long>((n + SNT - 1) / SNT, 8192); }

__global__ void k_zero_d(double* __restrict__ p, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = 0.0;
}

__global__ void k_copy_d(const double* __restrict__ a, double* __restrict__ b, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) b[i] = a[i];
}

__global__ void k_delta0(cplx<double>* __restrict__ d, int n, int g0) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    d[i] = mk(i == g0 ? 1.0 : 0.0, 0.0);
}

// ---- exchange-correlation, one grid point -----------------------------------------------

// Slater exchange + VWN5 correlation, spin-unpolarised; zero at rho <= 1e-10, the density
// threshold the reference gives Libxc (LDA_VWN.jl:175-186).
__device__ inline void vwn_point(double rho, double& eps, double& v) {
  if (!(rho > 1e-10)) {
    eps = 0.0;
    v = 0.0;
    return;
  }
  const double rs = 0.6203504908994 / cbrt(rho);
  const double fa = -0.687247939924714 * 2.0 / 3.0;
  const double ex = fa / rs, vx = 4.0 / 3.0 * fa / rs;
  const double a = 0.0310907, b = 3.72744, c = 12.9352, x0 = -0.10498;
  const double q = sqrt(4.0 * c - b * b);
  const double f1 = 2.0 * b / q, f2 = b * x0 / (x0 * x0 + b * x0 + c), f3 = 2.0 * (2.0 * x0 + b) / q;
  const double r12 = sqrt(rs);
  const double fx = rs + b * r12 + c;
  const double qx = atan(q / (2.0 * r12 + b));
  const double ec = a * (log(rs / fx) + f1 * qx - f2 * (log((r12 - x0) * (r12 - x0) / fx) + f3 * qx));
  const double tx = 2.0 * r12 + b;
  const double tt = tx * tx + q * q;
  const double vc = ec - r12 * a / 6.0 *
                             (2.0 / r12 - tx / fx - 4.0 * b / tt -
                              f2 * (2.0 / (r12 - x0) - tx / fx - 4.0 * (2.0 * x0 + b) / tt));
  eps = ex + ec;
  v = vx + vc;
}

// Perdew-Wang 92 correlation of the uniform gas (XC_funcs/XC_c_pw.jl), r > 0.
__device__ inline void pw92(double r, double& ec, double& vc) {
  const double rs = 0.6203504908994 / cbrt(r);
  const double a = 0.031091, a1 = 0.21370, b1 = 7.5957, b2 = 3.5876, b3 = 1.6382, b4 = 0.49294;
  const double q = sqrt(rs);
  const double om = 2.0 * a * (b1 * q + b2 * rs + b3 * rs * q + b4 * rs * rs);
  const double dom = 2.0 * a * (0.5 * b1 * q + b2 * rs + 1.5 * b3 * rs * q + 2.0 * b4 * rs * rs);
  const double lg = log(1.0 + 1.0 / om);
  ec = -2.0 * a * (1.0 + a1 * rs) * lg;
  vc = -2.0 * a * (1.0 + 2.0 / 3.0 * a1 * rs) * lg - 2.0 / 3.0 * a * (1.0 + a1 * rs) * dom / (om * (om + 1.0));
}

// PBE at one point from rho and |grad rho|^2: eps_xc, the local part v1 of V_xc and vsig,
// with V_xc = v1 - div(vsig grad rho).  Thresholds as the reference's: 1e-10 for the LDA
// parts, rho > 1e-6 and |grad rho|^2 > 1e-10 for the gradient corrections.
__device__ inline void pbe_point(double rho, double g2, double& eps, double& v1, double& vsig) {
  double ex = 0.0, vx = 0.0, ec = 0.0, vc = 0.0;
  if (rho > 1e-10) {
    const double rs = 0.6203504908994 / cbrt(rho);
    ex = -0.687247939924714 * (2.0 / 3.0) / rs;
    vx = 4.0 / 3.0 * ex;
    pw92(rho, ec, vc);
  }
  double sx = 0.0, v1x = 0.0, v2x = 0.0, sc = 0.0, v1c = 0.0, v2c = 0.0;
  const bool gga = rho > 1e-6 && g2 > 1e-10;
  if (gga) {
    // exchange enhancement (XC_x_pbe.jl)
    const double kf = 3.093667726280136 * cbrt(rho);
    const double ag = sqrt(g2);
    const double dsg = 0.5 / kf;
    const double s = ag * dsg / rho;
    const double kap = 0.804, mu = 0.2195149727645171;
    const double f2 = 1.0 + s * s * mu / kap;
    const double fx = kap - kap / f2;
    const double exu = -0.75 / M_PI * kf;
    const double dfx = 2.0 * mu * s / (f2 * f2);
    sx = exu * fx * rho;
    v1x = exu * fx + exu / 3.0 * fx + exu * dfx * (-4.0 / 3.0 * s);
    v2x = exu * dfx * dsg / ag;
    // correlation gradient term (XC_c_pbe.jl)
    const double ga = 0.0310906908696548950, be = 0.06672455060314922;
    double ecg, vcg;
    pw92(rho, ecg, vcg);
    const double ks = 1.128379167095513 * sqrt(1.919158292677513 / (0.6203504908994 / cbrt(rho)));
    const double t = ag / (2.0 * ks * rho);
    const double ee = exp(-ecg / ga);
    const double af = be / ga / (ee - 1.0);
    const double bf = ee * (vcg - ecg);
    const double y = af * t * t;
    const double den = 1.0 + y + y * y;
    const double xy = (1.0 + y) / den;
    const double qy = y * y * (2.0 + y) / (den * den);
    const double s1 = 1.0 + be / ga * t * t * xy;
    const double h0 = ga * log(s1);
    sc = rho * h0;
    v1c = h0 + be * t * t / s1 * (-7.0 / 3.0 * xy - qy * (af * bf / be - 7.0 / 3.0));
    v2c = be / (2.0 * ks * ks * rho) * (xy - qy) / s1;
  }
  eps = ex + ec + (gga ? (sx + sc) / rho : 0.0);
  v1 = vx + vc + v1x + v1c;
  vsig = v2x + v2c;
}

__global__ void k_vwn(const double* __restrict__ rho, long N, double* __restrict__ eps, double* __restrict__ v) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
    double e, x;
    vwn_point(rho[i], e, x);
    eps[i] = e;
    v[i] = x;
  }
}

// grad: 3 fields of N (d rho / dx_alpha); h_alpha = vsig grad_alpha, v1 into v.
__global__ void k_pbe(const double* __restrict__ rho, const double* __restrict__ grad, long N,
                      double* __restrict__ eps, double* __restrict__ v, double* __restrict__ h) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
    const double gx = grad[i], gy = grad[N + i], gz = grad[2 * N + i];
    double e, v1, vs;
    pbe_point(rho[i], gx * gx + gy * gy + gz * gz, e, v1, vs);
    eps[i] = e;
    v[i] = v1;
    h[i] = vs * gx;
    h[N + i] = vs * gy;
    h[2 * N + i] = vs * gz;
  }
}

// ---- sphere coefficients -------------------------------------------------------------------

// Hartree: out = (4 pi / G^2 - 1) rho(G), out(G = 0) = -rho(0) (see the file comment).
__global__ void k_poisson_sph(const cplx<double>* __restrict__ rhoG, const double* __restrict__ G2, int ng, int g0,
                              cplx<double>* __restrict__ out) {
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < ng; g += gridDim.x * blockDim.x) {
    const cplx<double> r = rhoG[g];
    out[g] = g == g0 ? mk(-r.x, -r.y) : scal(4.0 * M_PI / G2[g] - 1.0, r);
  }
}

// op_nabla (PWGrid.jl:351-372): out[alpha ng + g] = i G_alpha rho(G).
__global__ void k_grad_sph(const cplx<double>* __restrict__ rhoG, const double* __restrict__ G, int ng,
                           cplx<double>* __restrict__ out) {
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < ng; g += gridDim.x * blockDim.x) {
    const cplx<double> r = rhoG[g];
    for (int a = 0; a < 3; ++a) {
      const double q = G[3 * g + a];
      out[(long)a * ng + g] = mk(-q * r.y, q * r.x);
    }
  }
}

// op_nabla_dot (PWGrid.jl:377-398): out = i (G . h(G)), h given as 3 sphere fields.
__global__ void k_div_sph(const cplx<double>* __restrict__ hG, const double* __restrict__ G, int ng,
                          cplx<double>* __restrict__ out) {
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < ng; g += gridDim.x * blockDim.x) {
    double sr = 0.0, si = 0.0;
    for (int a = 0; a < 3; ++a) {
      const double q = G[3 * g + a];
      const cplx<double> h = hG[(long)a * ng + g];
      sr += q * h.x;
      si += q * h.y;
    }
    out[g] = mk(-si, sr);
  }
}

// symmetrize_rhoe! on the sphere: out(G) = 1/nsym sum_s rho(M_s m) e^{2 pi i t_s . (M_s m)} - rho(G),
// m the integer crystal coordinates of G.  pw_set_symmetry has checked that every M_s m
// lies in the lookup box and in the sphere.
__global__ void k_sym_sph(const cplx<double>* __restrict__ rhoG, const int* __restrict__ gm, int ng,
                          const int* __restrict__ M, const double* __restrict__ t, int nsym,
                          const int* __restrict__ tab, int lo0, int lo1, int lo2, int sp0, int sp1,
                          cplx<double>* __restrict__ out) {
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < ng; g += gridDim.x * blockDim.x) {
    const int m0 = gm[3 * g], m1 = gm[3 * g + 1], m2 = gm[3 * g + 2];
    double ar = 0.0, ai = 0.0;
    for (int s = 0; s < nsym; ++s) {
      const int* R = M + 9 * s;
      const int a = R[0] * m0 + R[1] * m1 + R[2] * m2;
      const int b = R[3] * m0 + R[4] * m1 + R[5] * m2;
      const int cc = R[6] * m0 + R[7] * m1 + R[8] * m2;
      const int j = tab[(a - lo0) + sp0 * ((b - lo1) + sp1 * (cc - lo2))];
      const double ph = t[3 * s] * a + t[3 * s + 1] * b + t[3 * s + 2] * cc;
      double sn, cs;
      sincospi(2.0 * ph, &sn, &cs);
      const cplx<double> v = rhoG[j];
      ar += v.x * cs - v.y * sn;
      ai += v.x * sn + v.y * cs;
    }
    const cplx<double> r = rhoG[g];
    out[g] = mk(ar / nsym - r.x, ai / nsym - r.y);
  }
}

// ---- grid combinations -------------------------------------------------------------------

// V_Total = V_Ps_loc + V_H + V_xc; with v64 set also the op_H potential V / N (FP64 and FP32).
__global__ void k_vtot(const double* __restrict__ vps, const double* __restrict__ vh, const double* __restrict__ vxc,
                       double* __restrict__ vtot, double* __restrict__ v64, float* __restrict__ v32, long N,
                       double s) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
    const double v = vps[i] + vh[i] + vxc[i];
    vtot[i] = v;
    if (v64) {
      const double w = v * s;
      v64[i] = w;
      v32[i] = (float)w;
    }
  }
}

// rho_out = beta rho_new + (1 - beta) rho_in, clamped below at eps(Float64).
__global__ void k_mix(const double* __restrict__ rn, const double* __restrict__ ri, double* __restrict__ ro, long N,
                      double beta) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
    const double r = beta * rn[i] + (1.0 - beta) * ri[i];
    ro[i] = r > DBL_EPSILON ? r : DBL_EPSILON;
  }
}

// part[3 b + j] = block b's share of sum_r rho (vps, vh, eps)_j, fixed order per block.
__global__ void __launch_bounds__(SNT) k_dot3(const double* __restrict__ rho, const double* __restrict__ a,
                                              const double* __restrict__ b, const double* __restrict__ e, long N,
                                              double* __restrict__ part) {
  __shared__ double sh[3][SNT / 64];
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (long i = blockIdx.x * (long)SNT + threadIdx.x; i < N; i += (long)gridDim.x * SNT) {
    const double r = rho[i];
    s0 += a[i] * r;
    s1 += b[i] * r;
    s2 += e[i] * r;
  }
  for (int off = 32; off > 0; off >>= 1) {
    s0 += __shfl_xor(s0, off, 64);
    s1 += __shfl_xor(s1, off, 64);
    s2 += __shfl_xor(s2, off, 64);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sh[0][w] = s0;
    sh[1][w] = s1;
    sh[2][w] = s2;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double t = 0.0;
    for (int i = 0; i < SNT / 64; ++i) t += sh[threadIdx.x][i];
    part[3 * blockIdx.x + threadIdx.x] = t;
  }
}

// ---- vector primitives of the density mixers (mix_*.jl) -------------------------------------

constexpr int VDOT_MAX = 64;   // dot products per launch
constexpr int VLIN_MAX = 32;   // vectors of one linear combination
constexpr int VDOT_BLK = 128;  // blocks per dot product

struct VecPairs {
  const double* a[VDOT_MAX];
  const double* b[VDOT_MAX];
};
struct VecLin {
  const double* x[VLIN_MAX];
  double c[VLIN_MAX];
};

// part[blk * npair + q] = block blk's share of a_q . b_q (fixed order).
__global__ void __launch_bounds__(SNT) k_vdots(VecPairs v, long n, int npair, double* __restrict__ part) {
  __shared__ double sh[SNT / 64];
  const int q = blockIdx.y;
  const double* __restrict__ a = v.a[q];
  const double* __restrict__ b = v.b[q];
  double s = 0.0;
  for (long i = blockIdx.x * (long)SNT + threadIdx.x; i < n; i += (long)gridDim.x * SNT) s += a[i] * b[i];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < SNT / 64; ++i) t += sh[i];
    part[(long)blockIdx.x * npair + q] = t;
  }
}

// y = ay y + sum_j c_j x_j (y is not read when ay == 0; y may be one of the x_j).
__global__ void k_vlin(VecLin v, int m, double ay, double* y, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double t = ay != 0.0 ? ay * y[i] : 0.0;
    for (int j = 0; j < m; ++j) t += v.c[j] * v.x[j][i];
    y[i] = t;
  }
}

// A[:, j] *= s[j] for a column-major complex block (rows x ncols, leading dimension ld).
__global__ void k_scale_cols(cplx<double>* __restrict__ A, long ld, int rows, int ncols, const double* __restrict__ s) {
  const long tot = (long)rows * ncols;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < tot; t += (long)gridDim.x * blockDim.x) {
    const int j = (int)(t / rows), i = (int)(t - (long)j * rows);
    cplx<double>& a = A[(long)j * ld + i];
    a = scal(s[j], a);
  }
}

__global__ void k_vclamp(double* __restrict__ y, long n, double lo) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    if (y[i] < lo) y[i] = lo;
}

// mix_adaptive! (mix_adaptive.jl:34-54): per point, beta grows by beta0 (capped at betamax)
// while the residual keeps its sign and is halved towards beta0 when it flips.
__global__ void k_mix_adaptive(double* __restrict__ mu, const double* __restrict__ nu, double* __restrict__ beta,
                               double* __restrict__ f, long n, double beta0, double betamax) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double t = nu[i] - mu[i];
    double b = beta[i];
    if (t * f[i] >= 0.0) {
      b = b + beta0;
      if (b > betamax) b = betamax;
    } else {
      b = 0.5 * (b + beta0);
    }
    beta[i] = b;
    f[i] = t;
    mu[i] = b * nu[i] + (1.0 - b) * mu[i];
  }
}

// ---- drivers -------------------------------------------------------------------------------

void free_scf(pw_ctx* c) {
  hipFree(c->d_vps);
  hipFree(c->d_sgrid);
  hipFree(c->d_ssph);
  hipFree(c->d_spart);
  hipFree(c->d_symM);
  hipFree(c->d_symt);
  hipFree(c->d_gm);
  hipFree(c->d_gtab);
  c->d_vps = nullptr;
  c->d_sgrid = c->d_ssph = c->d_spart = nullptr;
  c->sgrid_bytes = c->ssph_bytes = c->spart_bytes = 0;
  c->d_symM = nullptr;
  c->d_symt = nullptr;
  c->d_gm = c->d_gtab = nullptr;
  c->nsym = 0;
  c->sym_ng = 0;
}

int zero_fill(pw_ctx* c, double* out, long n) {
  if (n <= 0) return PW_OK;
  hipLaunchKernelGGL(k_zero_d, dim3(grid_blocks(n)), dim3(SNT), 0, c->stream, out, n);
  return hip_check(c, hipGetLastError(), "k_zero_d");
}

static int check_sphere(pw_ctx* c, const char* who) {
  if (c->N == 0) return set_error(c, PW_ERR_STATE, std::string(who) + ": pw_set_grid has not been called");
  if (c->krho.ngw == 0 || c->krho_g0 < 0)
    return set_error(c, PW_ERR_STATE, std::string(who) + ": pw_set_rho_sphere has not been called");
  return PW_OK;
}

// Sphere scratch with the delta_{G,0} column in slot SS_DELTA.
static int sphere_scratch(pw_ctx* c, cplx<double>** S) {
  const int ng = c->krho.ngw;
  int rc = ensure(c, &c->d_ssph, &c->ssph_bytes, (size_t)SS_COUNT * ng * sizeof(cplx<double>));
  if (rc) return rc;
  *S = reinterpret_cast<cplx<double>*>(c->d_ssph);
  hipLaunchKernelGGL(k_delta0, dim3((ng + 255) / 256), dim3(256), 0, c->stream, *S + (long)SS_DELTA * ng, ng,
                     c->krho_g0);
  return hip_check(c, hipGetLastError(), "k_delta0");
}

// fG = sum_r f(r) e^{-i G.r} on the rho sphere.
static int r2g(pw_ctx* c, const cplx<double>* delta, const double* f, cplx<double>* fG) {
  const int ng = c->krho.ngw;
  return vloc_apply_k(c, c->krho, f, 1, delta, ng, fG, ng);
}

static int ensure_vloc(pw_ctx* c, int ispin) {
  if (c->nspin <= ispin) {
    c->d_v64.resize(ispin + 1, nullptr);
    c->d_v32.resize(ispin + 1, nullptr);
    c->nspin = ispin + 1;
  }
  if (!c->d_v64[ispin]) {
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&c->d_v64[ispin]), c->N * sizeof(double));
    if (e != hipSuccess) return set_error(c, PW_ERR_NOMEM, "hipMalloc V_loc");
  }
  if (!c->d_v32[ispin]) {
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&c->d_v32[ispin]), c->N * sizeof(float));
    if (e != hipSuccess) return set_error(c, PW_ERR_NOMEM, "hipMalloc V_loc");
  }
  return PW_OK;
}

static int update_potential(pw_ctx* c, int ispin, int xc, int install, const double* rho, double* VH, double* Vxc,
                            double* Vtot, double* E) {
  int rc = check_sphere(c, "update_potential");
  if (rc) return rc;
  if (!c->d_vps) return set_error(c, PW_ERR_STATE, "update_potential: pw_set_vps_loc has not been called");
  if (!rho || ispin < 0 || ispin > 3 || (xc != PW_XC_VWN && xc != PW_XC_PBE))
    return set_error(c, PW_ERR_ARG, "update_potential: bad arguments");
  if (c->capturing) return set_error(c, PW_ERR_STATE, "update_potential: not available during graph capture");
  KData& k = c->krho;
  const long N = c->N;
  const int ng = k.ngw;
  rc = ensure(c, &c->d_sgrid, &c->sgrid_bytes, (size_t)SG_COUNT * N * sizeof(double));
  if (!rc) rc = ensure(c, &c->d_spart, &c->spart_bytes, 3 * (size_t)SRED * sizeof(double));
  if (rc) return rc;
  cplx<double>* S = nullptr;
  rc = sphere_scratch(c, &S);
  if (rc) return rc;
  if (install) {
    rc = ensure_vloc(c, ispin);
    if (rc) return rc;
  }
  double* Gd = reinterpret_cast<double*>(c->d_sgrid);
  double* grad = Gd + (long)SG_GRAD * N;
  double* h = Gd + (long)SG_H * N;
  double* eps = Gd + (long)SG_EPS * N;
  if (!VH) VH = Gd + (long)SG_VH * N;
  if (!Vxc) Vxc = Gd + (long)SG_VXC * N;
  if (!Vtot) Vtot = Gd + (long)SG_VTOT * N;
  const cplx<double>* delta = S + (long)SS_DELTA * ng;
  cplx<double>* rhoG = S + (long)SS_RHOG * ng;
  cplx<double>* T = S + (long)SS_T * ng;
  cplx<double>* hG = S + (long)SS_HG * ng;
  cplx<double>* dv = S + (long)SS_DIV * ng;
  const int gb = grid_blocks(N), sb = (ng + 255) / 256;
  const double invN = 1.0 / (double)N;

  rc = r2g(c, delta, rho, rhoG);
  if (rc) return rc;
  // Hartree
  hipLaunchKernelGGL(k_poisson_sph, dim3(sb), dim3(256), 0, c->stream, rhoG, k.d_gk2, ng, c->krho_g0, T);
  hipLaunchKernelGGL(k_copy_d, dim3(gb), dim3(SNT), 0, c->stream, rho, VH, N);
  rc = hip_check(c, hipGetLastError(), "Poisson kernels");
  if (!rc) rc = g2r_apply(c, k, 1, T, ng, VH, invN, 1);
  if (rc) return rc;
  // exchange-correlation
  if (xc == PW_XC_VWN) {
    hipLaunchKernelGGL(k_vwn, dim3(gb), dim3(SNT), 0, c->stream, rho, N, eps, Vxc);
    rc = hip_check(c, hipGetLastError(), "k_vwn");
    if (rc) return rc;
  } else {
    hipLaunchKernelGGL(k_grad_sph, dim3(sb), dim3(256), 0, c->stream, rhoG, k.d_gk, ng, T);
    rc = hip_check(c, hipGetLastError(), "k_grad_sph");
    if (!rc) rc = g2r_apply(c, k, 3, T, ng, grad, invN, 0);
    if (rc) return rc;
    hipLaunchKernelGGL(k_pbe, dim3(gb), dim3(SNT), 0, c->stream, rho, grad, N, eps, Vxc, h);
    rc = hip_check(c, hipGetLastError(), "k_pbe");
    for (int a = 0; a < 3 && !rc; ++a) rc = r2g(c, delta, h + (long)a * N, hG + (long)a * ng);
    if (rc) return rc;
    hipLaunchKernelGGL(k_div_sph, dim3(sb), dim3(256), 0, c->stream, hG, k.d_gk, ng, dv);
    rc = hip_check(c, hipGetLastError(), "k_div_sph");
    if (!rc) rc = g2r_apply(c, k, 1, dv, ng, Vxc, -invN, 1);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_vtot, dim3(gb), dim3(SNT), 0, c->stream, c->d_vps, VH, Vxc, Vtot,
                     install ? c->d_v64[ispin] : nullptr, install ? c->d_v32[ispin] : nullptr, N, invN);
  rc = hip_check(c, hipGetLastError(), "k_vtot");
  if (rc || !E) return rc;
  double* part = reinterpret_cast<double*>(c->d_spart);
  hipLaunchKernelGGL(k_dot3, dim3(SRED), dim3(SNT), 0, c->stream, rho, c->d_vps, VH, eps, N, part);
  rc = hip_check(c, hipGetLastError(), "k_dot3");
  if (rc) return rc;
  std::vector<double> hp(3 * SRED);
  hipError_t e = hipMemcpyAsync(hp.data(), part, hp.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_check(c, e, "energy download");
  double s[3] = {0.0, 0.0, 0.0};
  for (int b = 0; b < SRED; ++b)
    for (int j = 0; j < 3; ++j) s[j] += hp[3 * b + j];
  const double dV = c->omega / (double)N;
  E[0] = s[0] * dV;
  E[1] = 0.5 * s[1] * dV;
  E[2] = s[2] * dV;
  return PW_OK;
}

static int symmetrize(pw_ctx* c, double* rho) {
  int rc = check_sphere(c, "symmetrize_rhoe");
  if (rc) return rc;
  if (!rho) return set_error(c, PW_ERR_ARG, "symmetrize_rhoe: null rho");
  if (c->nsym <= 1) return PW_OK;
  if (c->sym_ng != c->krho.ngw)
    return set_error(c, PW_ERR_STATE, "symmetrize_rhoe: pw_set_symmetry must follow pw_set_rho_sphere");
  KData& k = c->krho;
  const int ng = k.ngw;
  cplx<double>* S = nullptr;
  rc = sphere_scratch(c, &S);
  if (rc) return rc;
  cplx<double>* rhoG = S + (long)SS_RHOG * ng;
  cplx<double>* dv = S + (long)SS_DIV * ng;
  rc = r2g(c, S + (long)SS_DELTA * ng, rho, rhoG);
  if (rc) return rc;
  hipLaunchKernelGGL(k_sym_sph, dim3((ng + 255) / 256), dim3(256), 0, c->stream, rhoG, c->d_gm, ng, c->d_symM,
                     c->d_symt, c->nsym, c->d_gtab, c->gtab_lo[0], c->gtab_lo[1], c->gtab_lo[2], c->gtab_span[0],
                     c->gtab_span[1], dv);
  rc = hip_check(c, hipGetLastError(), "k_sym_sph");
  if (rc) return rc;
  return g2r_apply(c, k, 1, dv, ng, rho, 1.0 / (double)c->N, 1);
}

}  // namespace pw

using namespace pw;

#define CTX_CHECK(c) \
  if (!(c)) return PW_ERR_ARG

extern "C" {

int pw_set_vps_loc(pw_ctx* c, const double* Vps) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (c->N == 0) return set_error(c, PW_ERR_STATE, "pw_set_grid must precede pw_set_vps_loc");
  if (!Vps) return set_error(c, PW_ERR_ARG, "pw_set_vps_loc: null V");
  return upload(c, &c->d_vps, Vps, (size_t)c->N);
}

int pw_set_symmetry(pw_ctx* c, int nsym, const int* M, const double* t, int ng, const int* gm) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  int rc = check_sphere(c, "pw_set_symmetry");
  if (rc) return rc;
  if (nsym < 1 || !M || !t || !gm || ng != c->krho.ngw)
    return set_error(c, PW_ERR_ARG, "pw_set_symmetry: bad arguments (ng must be the rho-sphere size)");
  int lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    lo[d] = hi[d] = gm[d];
    for (int g = 1; g < ng; ++g) {
      lo[d] = std::min(lo[d], gm[3 * g + d]);
      hi[d] = std::max(hi[d], gm[3 * g + d]);
    }
  }
  const long sp0 = hi[0] - lo[0] + 1, sp1 = hi[1] - lo[1] + 1, sp2 = hi[2] - lo[2] + 1;
  if (sp0 * sp1 * sp2 > (1L << 30)) return set_error(c, PW_ERR_ARG, "pw_set_symmetry: G box too large");
  std::vector<int> tab((size_t)(sp0 * sp1 * sp2), -1);
  auto at = [&](long a, long b, long q) -> long {
    if (a < lo[0] || a > hi[0] || b < lo[1] || b > hi[1] || q < lo[2] || q > hi[2]) return -1;
    return (a - lo[0]) + sp0 * ((b - lo[1]) + sp1 * (q - lo[2]));
  };
  for (int g = 0; g < ng; ++g) {
    const long p = at(gm[3 * g], gm[3 * g + 1], gm[3 * g + 2]);
    if (tab[p] >= 0) return set_error(c, PW_ERR_ARG, "pw_set_symmetry: repeated G-vector in the sphere");
    tab[p] = g;
  }
  for (int s = 0; s < nsym; ++s) {
    const int* R = M + 9 * s;
    for (int g = 0; g < ng; ++g) {
      const long m0 = gm[3 * g], m1 = gm[3 * g + 1], m2 = gm[3 * g + 2];
      const long p = at(R[0] * m0 + R[1] * m1 + R[2] * m2, R[3] * m0 + R[4] * m1 + R[5] * m2,
                        R[6] * m0 + R[7] * m1 + R[8] * m2);
      if (p < 0 || tab[p] < 0)
        return set_error(c, PW_ERR_ARG, "pw_set_symmetry: the G sphere is not closed under the operations");
    }
  }
  c->nsym = 0;
  rc = upload(c, &c->d_symM, M, 9 * (size_t)nsym);
  if (!rc) rc = upload(c, &c->d_symt, t, 3 * (size_t)nsym);
  if (!rc) rc = upload(c, &c->d_gm, gm, 3 * (size_t)ng);
  if (!rc) rc = upload(c, &c->d_gtab, tab.data(), tab.size());
  if (rc) return rc;
  for (int d = 0; d < 3; ++d) c->gtab_lo[d] = lo[d];
  c->gtab_span[0] = (int)sp0;
  c->gtab_span[1] = (int)sp1;
  c->gtab_span[2] = (int)sp2;
  c->nsym = nsym;
  c->sym_ng = ng;
  return PW_OK;
}

int pw_update_potential_dev(pw_ctx* c, int ispin, int xc, int install, const double* rho, double* VH, double* Vxc,
                            double* Vtot, double* E) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  return update_potential(c, ispin, xc, install, rho, VH, Vxc, Vtot, E);
}

int pw_set_vloc_parts_dev(pw_ctx* c, int ispin, const double* VH, const double* Vxc, double* Vtot) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (c->N == 0) return set_error(c, PW_ERR_STATE, "pw_set_grid has not been called");
  if (!c->d_vps) return set_error(c, PW_ERR_STATE, "pw_set_vloc_parts_dev: V_Ps_loc has not been set");
  if (ispin < 0 || ispin > 3 || !VH || !Vxc) return set_error(c, PW_ERR_ARG, "pw_set_vloc_parts_dev: bad arguments");
  int rc = ensure_vloc(c, ispin);
  if (rc) return rc;
  if (!Vtot) {
    rc = ensure(c, &c->d_sgrid, &c->sgrid_bytes, (size_t)SG_COUNT * c->N * sizeof(double));
    if (rc) return rc;
    Vtot = reinterpret_cast<double*>(c->d_sgrid) + (long)SG_VTOT * c->N;
  }
  hipLaunchKernelGGL(k_vtot, dim3(grid_blocks(c->N)), dim3(SNT), 0, c->stream, c->d_vps, VH, Vxc, Vtot, c->d_v64[ispin],
                     c->d_v32[ispin], c->N, 1.0 / (double)c->N);
  return hip_check(c, hipGetLastError(), "k_vtot");
}

int pw_mix_rhoe_dev(pw_ctx* c, double beta, const double* rho_new, const double* rho_in, double* rho_out) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (c->N == 0) return set_error(c, PW_ERR_STATE, "pw_set_grid has not been called");
  if (!rho_new || !rho_in || !rho_out) return set_error(c, PW_ERR_ARG, "pw_mix_rhoe_dev: null pointer");
  hipLaunchKernelGGL(k_mix, dim3(grid_blocks(c->N)), dim3(SNT), 0, c->stream, rho_new, rho_in, rho_out, c->N, beta);
  return hip_check(c, hipGetLastError(), "k_mix");
}

int pw_symmetrize_rhoe_dev(pw_ctx* c, double* rho) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  return symmetrize(c, rho);
}

int pw_vec_dots_dev(pw_ctx* c, int64_t n, int npair, const double* const* a, const double* const* b, double* out) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (n < 0 || npair < 0 || npair > VDOT_MAX || (npair > 0 && (!a || !b || !out)))
    return set_error(c, PW_ERR_ARG, "pw_vec_dots_dev: bad arguments (at most 64 pairs)");
  if (npair == 0) return PW_OK;
  VecPairs v{};
  for (int q = 0; q < npair; ++q) {
    if (!a[q] || !b[q]) return set_error(c, PW_ERR_ARG, "pw_vec_dots_dev: null vector");
    v.a[q] = a[q];
    v.b[q] = b[q];
  }
  int rc = ensure(c, &c->d_spart, &c->spart_bytes,
                  std::max<size_t>(3 * (size_t)SRED, (size_t)VDOT_BLK * npair) * sizeof(double));
  if (rc) return rc;
  double* part = reinterpret_cast<double*>(c->d_spart);
  hipLaunchKernelGGL(k_vdots, dim3(VDOT_BLK, npair), dim3(SNT), 0, c->stream, v, (long)n, npair, part);
  rc = hip_check(c, hipGetLastError(), "k_vdots");
  if (rc) return rc;
  std::vector<double> hp((size_t)VDOT_BLK * npair);
  hipError_t e = hipMemcpyAsync(hp.data(), part, hp.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_check(c, e, "dot products download");
  for (int q = 0; q < npair; ++q) {
    double t = 0.0;
    for (int b = 0; b < VDOT_BLK; ++b) t += hp[(size_t)b * npair + q];
    out[q] = t;
  }
  return PW_OK;
}

int pw_vec_lincomb_dev(pw_ctx* c, int64_t n, int m, const double* coef, const double* const* x, double ay, double* y) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (n < 0 || m < 0 || m > VLIN_MAX || !y || (m > 0 && (!coef || !x)))
    return set_error(c, PW_ERR_ARG, "pw_vec_lincomb_dev: bad arguments (at most 32 vectors)");
  if (n == 0) return PW_OK;
  VecLin v{};
  for (int j = 0; j < m; ++j) {
    if (!x[j]) return set_error(c, PW_ERR_ARG, "pw_vec_lincomb_dev: null vector");
    v.x[j] = x[j];
    v.c[j] = coef[j];
  }
  hipLaunchKernelGGL(k_vlin, dim3(grid_blocks(n)), dim3(SNT), 0, c->stream, v, m, ay, y, (long)n);
  return hip_check(c, hipGetLastError(), "k_vlin");
}

int pw_scale_cols_dev(pw_ctx* c, int rows, int ncols, const double* s, void* A, int64_t ld) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (rows < 0 || ncols < 0 || ld < rows || (ncols > 0 && (!s || !A)))
    return set_error(c, PW_ERR_ARG, "pw_scale_cols_dev: bad arguments");
  if (rows == 0 || ncols == 0) return PW_OK;
  int rc = ensure(c, &c->d_spart, &c->spart_bytes,
                  std::max<size_t>(3 * (size_t)SRED, (size_t)VDOT_BLK * VDOT_MAX) * sizeof(double) +
                      (size_t)ncols * sizeof(double));
  if (rc) return rc;
  double* ds = reinterpret_cast<double*>(c->d_spart) + std::max<size_t>(3 * (size_t)SRED, (size_t)VDOT_BLK * VDOT_MAX);
  hipError_t e = hipMemcpyAsync(ds, s, ncols * sizeof(double), hipMemcpyHostToDevice, c->stream);
  if (e != hipSuccess) return hip_check(c, e, "column scales upload");
  hipLaunchKernelGGL(k_scale_cols, dim3(grid_blocks((long)rows * ncols)), dim3(SNT), 0, c->stream,
                     reinterpret_cast<cplx<double>*>(A), (long)ld, rows, ncols, ds);
  return hip_check(c, hipGetLastError(), "k_scale_cols");
}

int pw_vec_clamp_dev(pw_ctx* c, int64_t n, double lo, double* y) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (n < 0 || !y) return set_error(c, PW_ERR_ARG, "pw_vec_clamp_dev: bad arguments");
  if (n == 0) return PW_OK;
  hipLaunchKernelGGL(k_vclamp, dim3(grid_blocks(n)), dim3(SNT), 0, c->stream, y, (long)n, lo);
  return hip_check(c, hipGetLastError(), "k_vclamp");
}

int pw_mix_adaptive_dev(pw_ctx* c, int64_t n, double beta0, double betamax, double* mu, const double* nu,
                        double* beta, double* f) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (n < 0 || !mu || !nu || !beta || !f) return set_error(c, PW_ERR_ARG, "pw_mix_adaptive_dev: bad arguments");
  if (n == 0) return PW_OK;
  hipLaunchKernelGGL(k_mix_adaptive, dim3(grid_blocks(n)), dim3(SNT), 0, c->stream, mu, nu, beta, f, (long)n, beta0,
                     betamax);
  return hip_check(c, hipGetLastError(), "k_mix_adaptive");
}

int pw_rhoe_get_dev(pw_ctx* c, double* rho) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (!c->d_rho || !rho) return set_error(c, PW_ERR_STATE, "rho not available");
  hipLaunchKernelGGL(k_copy_d, dim3(grid_blocks(c->N)), dim3(SNT), 0, c->stream, c->d_rho, rho, c->N);
  return hip_check(c, hipGetLastError(), "k_copy_d");
}

}  // extern "C"
