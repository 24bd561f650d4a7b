// Per-geometry set-up on the device, for relaxation / MD rebuilds without host work:
//   the GTH projectors betaNL[ik] of PsPotNL (PsPotNL.jl:16-77):
//     beta(g, j) = (-i)^l Y_lm(g) p_{l,iprj}(|g|) exp(-i g.tau_a),  g = G + k,
//   with Ylm_real (Ylm_real.jl:5-107) and eval_proj_G (PsPot_GTH.jl:338-391), plus the
//   products beta D and the FP32 copies the op_H kernels read;
//   V_Ps_loc (Hamiltonian.jl:119-141, calc_strfact.jl:2-16):
//     V(r) = Re sum_G [sum_a V_sp(a)(|G|) exp(-i G.tau_a) / Omega] e^{i G.r} on the rho sphere.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "pw_internal.h"

namespace pw {

// Real spherical harmonics in the QE convention (Ylm_real.jl:5-107).
__device__ inline double ylm_real(int l, int m, double x, double y, double z) {
  const double SMALL = 1.0e-9;
  const double rmod = sqrt(x * x + y * y + z * z);
  const double cost = rmod < SMALL ? 0.0 : z / rmod;
  double phi;
  if (x > SMALL)
    phi = atan(y / x);
  else if (x < -SMALL)
    phi = atan(y / x) + M_PI;
  else
    phi = y >= 0.0 ? M_PI / 2 : -M_PI / 2;
  const double sint = sqrt(fmax(0.0, 1.0 - cost * cost));
  switch (l) {
    case 0:
      return 0.5 * sqrt(1.0 / M_PI);
    case 1: {
      const double f = 0.5 * sqrt(3.0 / M_PI);
      if (m == -1) return f * sint * sin(phi);
      if (m == 0) return f * cost;
      return f * sint * cos(phi);
    }
    case 2: {
      const double a = sqrt(15.0 / 16.0 / M_PI), b = sqrt(15.0 / 4.0 / M_PI);
      if (m == -2) return a * sint * sint * sin(2.0 * phi);
      if (m == -1) return b * cost * sint * sin(phi);
      if (m == 0) return 0.25 * sqrt(5.0 / M_PI) * (3.0 * cost * cost - 1.0);
      if (m == 1) return b * cost * sint * cos(phi);
      return a * sint * sint * cos(2.0 * phi);
    }
    default: {
      const double a = 0.25 * sqrt(35.0 / 2.0 / M_PI), b = 0.25 * sqrt(105.0 / M_PI),
                   d = 0.25 * sqrt(21.0 / 2.0 / M_PI);
      if (m == -3) return a * sint * sint * sint * sin(3.0 * phi);
      if (m == -2) return b * sint * sint * cost * sin(2.0 * phi);
      if (m == -1) return d * sint * (5.0 * cost * cost - 1.0) * sin(phi);
      if (m == 0) return 0.25 * sqrt(7.0 / M_PI) * (5.0 * cost * cost * cost - 3.0 * cost);
      if (m == 1) return d * sint * (5.0 * cost * cost - 1.0) * cos(phi);
      if (m == 2) return b * sint * sint * cost * cos(2.0 * phi);
      return a * sint * sint * sint * cos(3.0 * phi);
    }
  }
}

// GTH projector in G space (PsPot_GTH.jl:338-391), iproj 1-based.
__device__ inline double proj_g(int l, int iproj, double gm, double rc, double omega) {
  const double x = (gm * rc) * (gm * rc);
  const double e = exp(-0.5 * x);
  double v = 0.0;
  if (l == 0) {
    if (iproj == 1) v = e;
    else if (iproj == 2) v = 2.0 / sqrt(15.0) * e * (3.0 - x);
    else v = (4.0 / 3.0) / sqrt(105.0) * e * (15.0 - 10.0 * x + x * x);
  } else if (l == 1) {
    if (iproj == 1) v = (1.0 / sqrt(3.0)) * e * gm;
    else if (iproj == 2) v = (2.0 / sqrt(105.0)) * e * gm * (5.0 - x);
    else v = (4.0 / 3.0) / sqrt(1155.0) * e * gm * (35.0 - 14.0 * x + x * x);
  } else if (l == 2) {
    if (iproj == 1) v = (1.0 / sqrt(15.0)) * e * gm * gm;
    else v = (2.0 / 3.0) / sqrt(105.0) * e * gm * gm * (7.0 - x);
  } else {
    v = gm * gm * gm * e / sqrt(105.0);
  }
  const double pre = 4.0 * pow(M_PI, 1.25) * sqrt(pow(2.0, l + 1) * pow(rc, 2 * l + 3) / omega);
  return pre * v;
}

// beta[j * ngw + g] for descriptor j = (atom, l, m, iprj).
__global__ void k_build_beta(const double* __restrict__ gk, int ngw, const int4* __restrict__ desc, int nbeta,
                             const double* __restrict__ tau, const int* __restrict__ species,
                             const double* __restrict__ rc, double omega, cplx<double>* __restrict__ beta) {
  const long tot = (long)ngw * nbeta;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < tot; t += (long)gridDim.x * blockDim.x) {
    const int j = (int)(t / ngw), g = (int)(t - (long)j * ngw);
    const int4 d = desc[j];
    const int a = d.x, l = d.y, m = d.z, ip = d.w;
    const double gx = gk[3 * g], gy = gk[3 * g + 1], gz = gk[3 * g + 2];
    const double gm = sqrt(gx * gx + gy * gy + gz * gz);
    const double r = ylm_real(l, m, gx, gy, gz) * proj_g(l, ip, gm, rc[4 * species[a] + l], omega);
    double sn, cs;
    sincos(tau[3 * a] * gx + tau[3 * a + 1] * gy + tau[3 * a + 2] * gz, &sn, &cs);
    // (-i)^l (cos - i sin)
    const cplx<double> sf = mk(r * cs, -r * sn);
    cplx<double> v;
    switch (l & 3) {
      case 0: v = sf; break;
      case 1: v = mk(sf.y, -sf.x); break;   // -i sf
      case 2: v = mk(-sf.x, -sf.y); break;  // -sf
      default: v = mk(-sf.y, sf.x); break;  // i sf
    }
    beta[t] = v;
  }
}

// bD[:, j] = sum_i D[i, j] beta[:, i] over the non-zeros of column j of D.
__global__ void k_beta_d(const cplx<double>* __restrict__ beta, int ngw, int nbeta, const int* __restrict__ cptr,
                         const int* __restrict__ rows, const double* __restrict__ vals, cplx<double>* __restrict__ bD) {
  const long tot = (long)ngw * nbeta;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < tot; t += (long)gridDim.x * blockDim.x) {
    const int j = (int)(t / ngw), g = (int)(t - (long)j * ngw);
    cplx<double> s = mk(0.0, 0.0);
    for (int q = cptr[j]; q < cptr[j + 1]; ++q) s = s + scal(vals[q], beta[(long)rows[q] * ngw + g]);
    bD[t] = s;
  }
}

__global__ void k_to_c32(const cplx<double>* __restrict__ in, cplx<float>* __restrict__ out, long n) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
    out[t] = cvt<float>(in[t]);
}

// coef[g] = sum_a V_sp(a)(g) exp(-i G.tau_a) / Omega on the rho sphere.
__global__ void k_vps_coef(const double* __restrict__ G, int ng, const double* __restrict__ vg,
                           const double* __restrict__ tau, const int* __restrict__ species, int natoms, double omega,
                           cplx<double>* __restrict__ coef) {
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < ng; g += gridDim.x * blockDim.x) {
    const double gx = G[3 * g], gy = G[3 * g + 1], gz = G[3 * g + 2];
    double sr = 0.0, si = 0.0;
    for (int a = 0; a < natoms; ++a) {
      double sn, cs;
      sincos(gx * tau[3 * a] + gy * tau[3 * a + 1] + gz * tau[3 * a + 2], &sn, &cs);
      const double v = vg[(size_t)species[a] * ng + g];
      sr += v * cs;
      si -= v * sn;
    }
    coef[g] = mk(sr / omega, si / omega);
  }
}

static int realloc_dev(pw_ctx* c, void** p, size_t bytes) {
  c->graph_gen++;
  hipFree(*p);
  *p = nullptr;
  if (hipMalloc(p, std::max<size_t>(bytes, 16)) != hipSuccess)
    return set_error(c, PW_ERR_NOMEM, "hipMalloc projectors");
  return PW_OK;
}

static inline int blocks_for(long n) { return (int)std::min<long>((n + 255) / 256, 16384); }

}  // namespace pw

using namespace pw;

#define CTX_CHECK(c) \
  if (!(c)) return PW_ERR_ARG

extern "C" {

int pw_build_projectors_dev(pw_ctx* c, int ik, int nbeta, const int* desc, int nsp, const double* rc) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (c->N == 0) return set_error(c, PW_ERR_STATE, "pw_set_grid has not been called");
  if (ik < 0 || ik >= (int)c->kd.size() || c->kd[ik].ngw == 0)
    return set_error(c, PW_ERR_STATE, "k-point " + std::to_string(ik) + " is not set up");
  if (nbeta != c->nbeta) return set_error(c, PW_ERR_ARG, "pw_build_projectors_dev: nbeta differs from pw_set_dij");
  if (nbeta > 0 && (!desc || nsp <= 0 || !rc)) return set_error(c, PW_ERR_ARG, "pw_build_projectors_dev: bad arguments");
  KData& k = c->kd[ik];
  k.nbeta = nbeta;
  if (nbeta == 0) return PW_OK;
  if (c->natoms == 0 || !c->d_tau) return set_error(c, PW_ERR_STATE, "pw_set_atoms has not been called");
  if (c->species_max >= nsp) return set_error(c, PW_ERR_ARG, "pw_build_projectors_dev: species index beyond nsp");
  std::vector<int4> dv(nbeta);
  for (int j = 0; j < nbeta; ++j) {
    const int a = desc[4 * j], l = desc[4 * j + 1], m = desc[4 * j + 2], ip = desc[4 * j + 3];
    const int nmax = l <= 1 ? 3 : (l == 2 ? 2 : 1);
    if (a < 0 || a >= c->natoms || l < 0 || l > 3 || m < -l || m > l || ip < 1 || ip > nmax)
      return set_error(c, PW_ERR_ARG, "pw_build_projectors_dev: projector " + std::to_string(j) + " out of range");
    dv[j] = make_int4(a, l, m, ip);
  }
  int rc_ = ensure_dcsr(c);
  if (rc_) return rc_;
  int4* d_desc = nullptr;
  double* d_rc = nullptr;
  rc_ = upload(c, reinterpret_cast<int**>(&d_desc), reinterpret_cast<const int*>(dv.data()), 4 * (size_t)nbeta);
  if (!rc_) rc_ = upload(c, &d_rc, rc, 4 * (size_t)nsp);
  const size_t n = (size_t)k.ngw * nbeta;
  if (!rc_) rc_ = realloc_dev(c, &k.d_beta64, n * sizeof(cplx<double>));
  if (!rc_) rc_ = realloc_dev(c, &k.d_betaD64, n * sizeof(cplx<double>));
  if (!rc_) rc_ = realloc_dev(c, &k.d_beta32, n * sizeof(cplx<float>));
  if (!rc_) rc_ = realloc_dev(c, &k.d_betaD32, n * sizeof(cplx<float>));
  if (rc_) {
    hipFree(d_desc);
    hipFree(d_rc);
    return rc_;
  }
  cplx<double>* b = reinterpret_cast<cplx<double>*>(k.d_beta64);
  cplx<double>* bd = reinterpret_cast<cplx<double>*>(k.d_betaD64);
  const int nb = blocks_for((long)n);
  hipLaunchKernelGGL(k_build_beta, dim3(nb), dim3(256), 0, c->stream, k.d_gk, k.ngw, d_desc, nbeta, c->d_tau,
                     c->d_species, d_rc, c->omega, b);
  hipLaunchKernelGGL(k_beta_d, dim3(nb), dim3(256), 0, c->stream, b, k.ngw, nbeta, c->d_dcol_ptr, c->d_drow,
                     c->d_dval, bd);
  hipLaunchKernelGGL(k_to_c32, dim3(nb), dim3(256), 0, c->stream, b, reinterpret_cast<cplx<float>*>(k.d_beta32),
                     (long)n);
  hipLaunchKernelGGL(k_to_c32, dim3(nb), dim3(256), 0, c->stream, bd, reinterpret_cast<cplx<float>*>(k.d_betaD32),
                     (long)n);
  rc_ = hip_check(c, hipGetLastError(), "projector kernels");
  hipError_t e = hipStreamSynchronize(c->stream);
  hipFree(d_desc);
  hipFree(d_rc);
  if (rc_) return rc_;
  return hip_check(c, e, "projector build");
}

int pw_build_vps_loc_dev(pw_ctx* c, double* Vps) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (c->N == 0) return set_error(c, PW_ERR_STATE, "pw_set_grid has not been called");
  KData& k = c->krho;
  if (k.ngw == 0 || c->krho_g0 < 0 || !c->d_vg)
    return set_error(c, PW_ERR_STATE, "pw_set_rho_sphere has not been called");
  if (c->natoms == 0 || !c->d_tau) return set_error(c, PW_ERR_STATE, "pw_set_atoms has not been called");
  if (c->species_max >= c->nsp_loc)
    return set_error(c, PW_ERR_ARG, "V_Ps_loc: a species index exceeds the species of pw_set_rho_sphere");
  const int ng = k.ngw;
  int rc = ensure(c, &c->d_fgrid, &c->fgrid_bytes, (size_t)ng * sizeof(cplx<double>));
  if (rc) return rc;
  if (!c->d_vps) {
    if (hipMalloc(reinterpret_cast<void**>(&c->d_vps), c->N * sizeof(double)) != hipSuccess)
      return set_error(c, PW_ERR_NOMEM, "hipMalloc V_Ps_loc");
  }
  cplx<double>* coef = reinterpret_cast<cplx<double>*>(c->d_fgrid);
  hipLaunchKernelGGL(k_vps_coef, dim3((ng + 255) / 256), dim3(256), 0, c->stream, k.d_gk, ng, c->d_vg, c->d_tau,
                     c->d_species, c->natoms, c->omega, coef);
  rc = hip_check(c, hipGetLastError(), "k_vps_coef");
  if (!rc) rc = g2r_apply(c, k, 1, coef, ng, c->d_vps, 1.0, 0);
  if (rc) return rc;
  hipError_t e = hipSuccess;
  if (Vps) e = hipMemcpyAsync(Vps, c->d_vps, c->N * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  return hip_check(c, e, "V_Ps_loc build");
}

}  // extern "C"
