// Forces and energy gradients of the pseudopotential energies: the adjoint of E_Ps_loc with
// respect to V_sp(G) and the atomic positions (through the structure factor), and the GTH
// nonlocal force (through the projectors' exp(-i (G+k).tau) phase).
//
// rho(G) for the local terms comes from the sphere-pruned FFT pipeline of op_V_loc run on
// the rho sphere with psi = delta_{G,0}: IFFT[delta] = 1 everywhere, so the forward pass
// returns FFT[rho / N] on the sphere, exactly RhoeG / Npoints of calc_forces_Ps_loc.jl:37-38.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "pw_internal.h"

namespace pw {

constexpr int FNT = 256;  // threads per block of the reductions below (4 wavefronts)

// Sum over the block; every thread gets the result.  Deterministic order.
__device__ inline double block_sum(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[w] = v;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += sh[i];
  return t;
}

__global__ void k_scale_grid(const double* __restrict__ in, double* __restrict__ out, long n, double s) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    out[i] = in[i] * s;
}

__global__ void k_delta(cplx<double>* __restrict__ d, int n, int g0) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    d[i] = mk(i == g0 ? 1.0 : 0.0, 0.0);
}

// dE_Ps_loc/dtau_a = sum_G Re(-i G V_sp(G) exp(-i G.tau_a) conj(rho(G)))
//                  = -sum_G G V_sp(G) (cos(G.tau) Im rho + sin(G.tau) Re rho).
// One block per atom; the G = 0 term vanishes with G.
__global__ void __launch_bounds__(FNT) k_loc_tau(const cplx<double>* __restrict__ rhoG,
                                                 const double* __restrict__ G, int ng,
                                                 const double* __restrict__ vg,
                                                 const double* __restrict__ tau,
                                                 const int* __restrict__ species, double* __restrict__ dtau) {
  __shared__ double sh[FNT / 64];
  const int a = blockIdx.x;
  const double tx = tau[3 * a], ty = tau[3 * a + 1], tz = tau[3 * a + 2];
  const double* v = vg + (size_t)species[a] * ng;
  double fx = 0.0, fy = 0.0, fz = 0.0;
  for (int g = threadIdx.x; g < ng; g += FNT) {
    const double gx = G[3 * g], gy = G[3 * g + 1], gz = G[3 * g + 2];
    double s, cs;
    sincos(gx * tx + gy * ty + gz * tz, &s, &cs);
    const cplx<double> r = rhoG[g];
    const double t = v[g] * (cs * r.y + s * r.x);
    fx -= gx * t;
    fy -= gy * t;
    fz -= gz * t;
  }
  fx = block_sum(fx, sh);
  fy = block_sum(fy, sh);
  fz = block_sum(fz, sh);
  if (threadIdx.x == 0) {
    dtau[3 * a] = fx;
    dtau[3 * a + 1] = fy;
    dtau[3 * a + 2] = fz;
  }
}

// dE_Ps_loc/dV_sp(G) = Re(S_sp(G) conj(rho(G))), S_sp(G) = sum_{a in sp} exp(-i G.tau_a);
// part[block] = partial sums of E_Ps_loc = sum_{G,sp} V_sp(G) dE/dV_sp(G).
__global__ void __launch_bounds__(FNT) k_loc_vg(const cplx<double>* __restrict__ rhoG,
                                                const double* __restrict__ G, int ng,
                                                const double* __restrict__ vg, int nsp,
                                                const double* __restrict__ tau,
                                                const int* __restrict__ species, int natoms,
                                                double* __restrict__ dvg, double* __restrict__ part) {
  __shared__ double sh[FNT / 64];
  const int g = blockIdx.x * FNT + threadIdx.x;
  double e = 0.0;
  if (g < ng) {
    const double gx = G[3 * g], gy = G[3 * g + 1], gz = G[3 * g + 2];
    const cplx<double> r = rhoG[g];
    for (int sp = 0; sp < nsp; ++sp) {
      double sr = 0.0, si = 0.0;
      for (int a = 0; a < natoms; ++a) {
        if (species[a] != sp) continue;
        double s, cs;
        sincos(gx * tau[3 * a] + gy * tau[3 * a + 1] + gz * tau[3 * a + 2], &s, &cs);
        sr += cs;
        si -= s;
      }
      const double d = sr * r.x + si * r.y;
      if (dvg) dvg[g + (size_t)ng * sp] = d;
      e += vg[g + (size_t)ng * sp] * d;
    }
  }
  e = block_sum(e, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = e;
}

// out[:, alpha * nbeta + j] = (G+k)_alpha * beta[:, j]  (the i is applied in k_nloc_reduce)
__global__ void k_gbeta3(const cplx<double>* __restrict__ beta, const double* __restrict__ gk, int ngw,
                         int nbeta, cplx<double>* __restrict__ out) {
  const long per = (long)ngw * nbeta;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < 3 * per; t += (long)gridDim.x * blockDim.x) {
    const int alpha = (int)(t / per);
    const long r = t - alpha * per;
    const int g = (int)(r % ngw);
    out[t] = scal(gk[3 * g + alpha], beta[r]);
  }
}

// One block per projector j: DB_j = sum_i D[i, j] B[i, st] over the non-zeros of column j,
//   part[4j]       = sum_st w_st Re(conj(DB_j) B[j, st])               (E_Ps_nloc)
//   part[4j+1+al]  = sum_st w_st Im(conj(DB_j) X[al*nbeta + j, st])   (half the force)
// with B = beta^H psi (nbeta x nb) and X = ((G+k)_al beta)^H psi (3 nbeta x nb), so that
// Im(conj(DB) X) = Re(conj(DB) <i (G+k) beta|psi>).
__global__ void __launch_bounds__(FNT) k_nloc_reduce(const cplx<double>* __restrict__ B,
                                                     const cplx<double>* __restrict__ X, int nbeta, int nb,
                                                     const int* __restrict__ cptr, const int* __restrict__ rows,
                                                     const double* __restrict__ vals,
                                                     const double* __restrict__ w, double* __restrict__ part) {
  __shared__ double sh[FNT / 64];
  const int j = blockIdx.x;
  const int q0 = cptr[j], q1 = cptr[j + 1];
  double e = 0.0, f0 = 0.0, f1 = 0.0, f2 = 0.0;
  for (int st = threadIdx.x; st < nb; st += FNT) {
    const cplx<double>* Bs = B + (size_t)nbeta * st;
    cplx<double> db = mk(0.0, 0.0);
    for (int q = q0; q < q1; ++q) db = db + scal(vals[q], Bs[rows[q]]);
    const double ws = w[st];
    const cplx<double> b = Bs[j];
    e += ws * (db.x * b.x + db.y * b.y);
    const cplx<double>* Xs = X + (size_t)3 * nbeta * st + j;
    const cplx<double> x0 = Xs[0], x1 = Xs[nbeta], x2 = Xs[2 * nbeta];
    f0 += ws * (db.x * x0.y - db.y * x0.x);
    f1 += ws * (db.x * x1.y - db.y * x1.x);
    f2 += ws * (db.x * x2.y - db.y * x2.x);
  }
  e = block_sum(e, sh);
  f0 = block_sum(f0, sh);
  f1 = block_sum(f1, sh);
  f2 = block_sum(f2, sh);
  if (threadIdx.x == 0) {
    part[4 * j] = e;
    part[4 * j + 1] = f0;
    part[4 * j + 2] = f1;
    part[4 * j + 3] = f2;
  }
}

// calc_E_kin (calc_energies.jl:342-375) of one k-point: part[st] = w_st / 2 sum_g |G+k|^2 |psi_g,st|^2,
// one block per band.
__global__ void __launch_bounds__(FNT) k_ekin(const cplx<double>* __restrict__ psi, long ld,
                                              const double* __restrict__ gk2, int ngw, const double* __restrict__ w,
                                              double* __restrict__ part) {
  __shared__ double sh[FNT / 64];
  const int st = blockIdx.x;
  const cplx<double>* p = psi + (long)st * ld;
  double e = 0.0;
  for (int g = threadIdx.x; g < ngw; g += FNT) {
    const cplx<double> v = p[g];
    e += gk2[g] * (v.x * v.x + v.y * v.y);
  }
  e = block_sum(e, sh);
  if (threadIdx.x == 0) part[st] = 0.5 * w[st] * e;
}

// calc_E_Ps_nloc (calc_energies.jl:7-58, GTH) of one k-point: part[st] = w_st sum_j
// Re(conj((D B)_j) B_j) with B = beta^H psi (nbeta x nb); one thread per band.
__global__ void k_enl(const cplx<double>* __restrict__ B, int nbeta, int nb, const int* __restrict__ cptr,
                      const int* __restrict__ rows, const double* __restrict__ vals, const double* __restrict__ w,
                      double* __restrict__ part) {
  const int st = blockIdx.x * blockDim.x + threadIdx.x;
  if (st >= nb) return;
  const cplx<double>* Bs = B + (size_t)nbeta * st;
  double e = 0.0;
  for (int j = 0; j < nbeta; ++j) {
    cplx<double> db = mk(0.0, 0.0);
    for (int q = cptr[j]; q < cptr[j + 1]; ++q) db = db + scal(vals[q], Bs[rows[q]]);
    e += db.x * Bs[j].x + db.y * Bs[j].y;
  }
  part[st] = w[st] * e;
}

void free_forces(pw_ctx* c) {
  free_k(c->krho);
  c->krho_g0 = -1;
  hipFree(c->d_vg);
  hipFree(c->d_tau);
  hipFree(c->d_species);
  hipFree(c->d_beta_atom);
  hipFree(c->d_dcol_ptr);
  hipFree(c->d_drow);
  hipFree(c->d_dval);
  hipFree(c->d_fbeta);
  hipFree(c->d_fmat);
  hipFree(c->d_fgrid);
  hipFree(c->d_fpart);
  c->d_vg = nullptr;
  c->d_tau = nullptr;
  c->d_species = nullptr;
  c->d_beta_atom = nullptr;
  c->d_dcol_ptr = c->d_drow = nullptr;
  c->d_dval = nullptr;
  c->d_fbeta = c->d_fmat = c->d_fgrid = c->d_fpart = nullptr;
  c->fbeta_bytes = c->fmat_bytes = c->fgrid_bytes = c->fpart_bytes = 0;
  c->dcsr_valid = false;
}

// Non-zeros of each column of D, on the device.
int ensure_dcsr(pw_ctx* c) {
  if (c->dcsr_valid) return PW_OK;
  const int nb = c->nbeta;
  std::vector<int> ptr(nb + 1, 0), rows;
  std::vector<double> vals;
  for (int j = 0; j < nb; ++j) {
    for (int i = 0; i < nb; ++i) {
      const double d = c->h_D[(size_t)j * nb + i];
      if (d != 0.0) {
        rows.push_back(i);
        vals.push_back(d);
      }
    }
    ptr[j + 1] = (int)rows.size();
  }
  int rc = upload(c, &c->d_dcol_ptr, ptr.data(), ptr.size());
  if (!rc) rc = upload(c, &c->d_drow, rows.data(), rows.size());
  if (!rc) rc = upload(c, &c->d_dval, vals.data(), vals.size());
  if (!rc) c->dcsr_valid = true;
  return rc;
}

static int loc_grad(pw_ctx* c, const double* rho, double* E, double* dE_dvg, double* dE_dtau) {
  if (c->N == 0) return set_error(c, PW_ERR_STATE, "pw_set_grid has not been called");
  KData& k = c->krho;
  if (k.ngw == 0 || c->krho_g0 < 0 || !c->d_vg)
    return set_error(c, PW_ERR_STATE, "pw_set_rho_sphere has not been called");
  if (c->natoms == 0) return set_error(c, PW_ERR_STATE, "pw_set_atoms has not been called");
  if (c->species_max >= c->nsp_loc)
    return set_error(c, PW_ERR_ARG, "E_Ps_loc: a species index exceeds the species of pw_set_rho_sphere");
  const int ng = k.ngw;
  // grid: rho / N (N doubles), then delta_G0 and rho(G) (ng complex each)
  const size_t gbytes = ((size_t)c->N * sizeof(double) + 255) / 256 * 256;
  int rc = ensure(c, &c->d_fgrid, &c->fgrid_bytes, gbytes + 2 * (size_t)ng * sizeof(cplx<double>));
  if (rc) return rc;
  double* rhoN = reinterpret_cast<double*>(c->d_fgrid);
  cplx<double>* delta = reinterpret_cast<cplx<double>*>(reinterpret_cast<char*>(c->d_fgrid) + gbytes);
  cplx<double>* rhoG = delta + ng;
  const int nblk = (ng + FNT - 1) / FNT;
  // partial sums: dE/dvg (ng * nsp), E partials (nblk), dE/dtau (3 natoms)
  const size_t nv = (size_t)ng * c->nsp_loc;
  rc = ensure(c, &c->d_fpart, &c->fpart_bytes, (nv + nblk + 3 * (size_t)c->natoms) * sizeof(double));
  if (rc) return rc;
  double* dvg = reinterpret_cast<double*>(c->d_fpart);
  double* epart = dvg + nv;
  double* dtau = epart + nblk;
  const double s = 1.0 / (double)c->N;
  const int gb = (int)std::min<long>((c->N + 255) / 256, 4096);
  hipError_t e = hipSuccess;
  if (rho) {
    e = hipMemcpyAsync(rhoN, rho, c->N * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) return hip_check(c, e, "rho upload");
    hipLaunchKernelGGL(k_scale_grid, dim3(gb), dim3(256), 0, c->stream, rhoN, rhoN, c->N, s);
  } else {
    hipLaunchKernelGGL(k_scale_grid, dim3(gb), dim3(256), 0, c->stream, c->d_rho, rhoN, c->N, s);
  }
  hipLaunchKernelGGL(k_delta, dim3((ng + 255) / 256), dim3(256), 0, c->stream, delta, ng, c->krho_g0);
  rc = hip_check(c, hipGetLastError(), "rho(G) set-up kernels");
  if (rc) return rc;
  rc = vloc_apply_k(c, k, rhoN, 1, delta, ng, rhoG, ng);
  if (rc) return rc;
  if (dE_dtau) {
    hipLaunchKernelGGL(k_loc_tau, dim3(c->natoms), dim3(FNT), 0, c->stream, rhoG, k.d_gk, ng, c->d_vg, c->d_tau,
                       c->d_species, dtau);
    rc = hip_check(c, hipGetLastError(), "k_loc_tau");
    if (rc) return rc;
    e = hipMemcpyAsync(dE_dtau, dtau, 3 * (size_t)c->natoms * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e != hipSuccess) return hip_check(c, e, "dE/dtau download");
  }
  std::vector<double> hp;
  if (E || dE_dvg) {
    hipLaunchKernelGGL(k_loc_vg, dim3(nblk), dim3(FNT), 0, c->stream, rhoG, k.d_gk, ng, c->d_vg, c->nsp_loc,
                       c->d_tau, c->d_species, c->natoms, dvg, epart);
    rc = hip_check(c, hipGetLastError(), "k_loc_vg");
    if (rc) return rc;
    hp.resize(nblk);
    e = hipMemcpyAsync(hp.data(), epart, nblk * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && dE_dvg)
      e = hipMemcpyAsync(dE_dvg, dvg, nv * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e != hipSuccess) return hip_check(c, e, "dE/dvg download");
  }
  e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_check(c, e, "E_Ps_loc gradient");
  if (E) {
    double t = 0.0;
    for (double v : hp) t += v;
    *E = t;
  }
  return PW_OK;
}

static int nloc_forces(pw_ctx* c, int ik, int nb, const cplx<double>* psi, long ldpsi, const double* w,
                       double* F, double* E) {
  if (c->N == 0) return set_error(c, PW_ERR_STATE, "pw_set_grid has not been called");
  if (ik < 0 || ik >= (int)c->kd.size() || c->kd[ik].ngw == 0)
    return set_error(c, PW_ERR_STATE, "k-point " + std::to_string(ik) + " is not set up");
  KData& k = c->kd[ik];
  if (nb < 0 || ldpsi < k.ngw || (nb > 0 && (!psi || !w)))
    return set_error(c, PW_ERR_ARG, "calc_forces_Ps_nloc: bad arguments");
  if (k.nbeta == 0 || nb == 0) return PW_OK;
  if (c->natoms == 0 || !c->d_beta_atom || c->nbeta_atoms != k.nbeta)
    return set_error(c, PW_ERR_STATE, "pw_set_atoms with beta_atom for every projector has not been called");
  const int nbeta = k.nbeta;
  int rc = ensure_dcsr(c);
  if (rc) return rc;
  rc = ensure(c, &c->d_fbeta, &c->fbeta_bytes, 3 * (size_t)k.ngw * nbeta * sizeof(cplx<double>));
  if (rc) return rc;
  const size_t mB = (size_t)nbeta * nb, mX = 3 * mB;
  rc = ensure(c, &c->d_fmat, &c->fmat_bytes, (mB + mX) * sizeof(cplx<double>) + (size_t)nb * sizeof(double));
  if (rc) return rc;
  rc = ensure(c, &c->d_fpart, &c->fpart_bytes, 4 * (size_t)nbeta * sizeof(double));
  if (rc) return rc;
  cplx<double>* gb = reinterpret_cast<cplx<double>*>(c->d_fbeta);
  cplx<double>* Bm = reinterpret_cast<cplx<double>*>(c->d_fmat);
  cplx<double>* Xm = Bm + mB;
  double* dw = reinterpret_cast<double*>(Xm + mX);
  double* part = reinterpret_cast<double*>(c->d_fpart);
  const cplx<double>* beta = reinterpret_cast<const cplx<double>*>(k.d_beta64);
  hipError_t e = hipMemcpyAsync(dw, w, nb * sizeof(double), hipMemcpyHostToDevice, c->stream);
  if (e != hipSuccess) return hip_check(c, e, "band weights upload");
  const long tot = 3L * k.ngw * nbeta;
  hipLaunchKernelGGL(k_gbeta3, dim3((int)std::min<long>((tot + 255) / 256, 8192)), dim3(256), 0, c->stream, beta,
                     k.d_gk, k.ngw, nbeta, gb);
  rc = hip_check(c, hipGetLastError(), "k_gbeta3");
  if (rc) return rc;
  const cplx<double> one = mk(1.0, 0.0), zero = mk(0.0, 0.0);
  rc = zgemm(c, 1, nbeta, nb, k.ngw, one, beta, k.ngw, psi, ldpsi, zero, Bm, nbeta, 0);
  if (rc) return rc;
  rc = zgemm(c, 1, 3 * nbeta, nb, k.ngw, one, gb, k.ngw, psi, ldpsi, zero, Xm, 3 * nbeta, 0);
  if (rc) return rc;
  hipLaunchKernelGGL(k_nloc_reduce, dim3(nbeta), dim3(FNT), 0, c->stream, Bm, Xm, nbeta, nb, c->d_dcol_ptr,
                     c->d_drow, c->d_dval, dw, part);
  rc = hip_check(c, hipGetLastError(), "k_nloc_reduce");
  if (rc) return rc;
  std::vector<double> hp(4 * (size_t)nbeta);
  std::vector<int> atom(nbeta);
  e = hipMemcpyAsync(hp.data(), part, hp.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(atom.data(), c->d_beta_atom, nbeta * sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_check(c, e, "nonlocal force download");
  double et = 0.0;
  for (int j = 0; j < nbeta; ++j) {
    et += hp[4 * j];
    if (F) {
      const int a = atom[j];
      for (int x = 0; x < 3; ++x) F[3 * a + x] += 2.0 * hp[4 * j + 1 + x];
    }
  }
  if (E) *E += et;
  return PW_OK;
}

// E[0] += E_kin and E[1] += E_Ps_nloc contributions of k-point ik with band weights w.
static int psi_energies(pw_ctx* c, int ik, int nb, const cplx<double>* psi, long ldpsi, const double* w, double* E) {
  if (c->N == 0) return set_error(c, PW_ERR_STATE, "pw_set_grid has not been called");
  if (ik < 0 || ik >= (int)c->kd.size() || c->kd[ik].ngw == 0)
    return set_error(c, PW_ERR_STATE, "k-point " + std::to_string(ik) + " is not set up");
  KData& k = c->kd[ik];
  if (nb < 0 || ldpsi < k.ngw || !E || (nb > 0 && (!psi || !w)))
    return set_error(c, PW_ERR_ARG, "calc_energies: bad arguments");
  if (nb == 0) return PW_OK;
  const int nbeta = k.nbeta;
  const size_t mB = (size_t)nbeta * nb;
  int rc = ensure(c, &c->d_fmat, &c->fmat_bytes, mB * sizeof(cplx<double>) + (size_t)nb * sizeof(double));
  if (!rc) rc = ensure(c, &c->d_fpart, &c->fpart_bytes, 2 * (size_t)nb * sizeof(double));
  if (!rc && nbeta > 0) rc = ensure_dcsr(c);
  if (rc) return rc;
  cplx<double>* Bm = reinterpret_cast<cplx<double>*>(c->d_fmat);
  double* dw = reinterpret_cast<double*>(Bm + mB);
  double* part = reinterpret_cast<double*>(c->d_fpart);
  hipError_t e = hipMemcpyAsync(dw, w, nb * sizeof(double), hipMemcpyHostToDevice, c->stream);
  if (e != hipSuccess) return hip_check(c, e, "band weights upload");
  hipLaunchKernelGGL(k_ekin, dim3(nb), dim3(FNT), 0, c->stream, psi, ldpsi, k.d_gk2, k.ngw, dw, part);
  rc = hip_check(c, hipGetLastError(), "k_ekin");
  if (rc) return rc;
  if (nbeta > 0) {
    const cplx<double> one = mk(1.0, 0.0), zero = mk(0.0, 0.0);
    rc = zgemm(c, 1, nbeta, nb, k.ngw, one, reinterpret_cast<const cplx<double>*>(k.d_beta64), k.ngw, psi, ldpsi,
               zero, Bm, nbeta, 0);
    if (rc) return rc;
    hipLaunchKernelGGL(k_enl, dim3((nb + 63) / 64), dim3(64), 0, c->stream, Bm, nbeta, nb, c->d_dcol_ptr, c->d_drow,
                       c->d_dval, dw, part + nb);
    rc = hip_check(c, hipGetLastError(), "k_enl");
    if (rc) return rc;
  }
  std::vector<double> hp(2 * (size_t)nb, 0.0);
  e = hipMemcpyAsync(hp.data(), part, (nbeta > 0 ? 2 : 1) * (size_t)nb * sizeof(double), hipMemcpyDeviceToHost,
                     c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_check(c, e, "energy download");
  double ek = 0.0, en = 0.0;
  for (int st = 0; st < nb; ++st) {
    ek += hp[st];
    en += hp[nb + st];
  }
  E[0] += ek;
  E[1] += en;
  return PW_OK;
}

}  // namespace pw

using namespace pw;

#define CTX_CHECK(c) \
  if (!(c)) return PW_ERR_ARG

extern "C" {

int pw_set_atoms(pw_ctx* c, int natoms, const double* tau, const int* species, int nbeta,
                 const int* beta_atom) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (natoms <= 0 || !tau || !species || nbeta < 0 || (nbeta > 0 && !beta_atom))
    return set_error(c, PW_ERR_ARG, "pw_set_atoms: bad arguments");
  for (int a = 0; a < natoms; ++a)
    if (species[a] < 0) return set_error(c, PW_ERR_ARG, "pw_set_atoms: negative species index");
  for (int j = 0; j < nbeta; ++j)
    if (beta_atom[j] < 0 || beta_atom[j] >= natoms)
      return set_error(c, PW_ERR_ARG, "pw_set_atoms: beta_atom out of range");
  int rc = upload(c, &c->d_tau, tau, 3 * (size_t)natoms);
  if (!rc) rc = upload(c, &c->d_species, species, natoms);
  if (rc) return rc;
  if (nbeta > 0) {
    rc = upload(c, &c->d_beta_atom, beta_atom, nbeta);
    if (rc) return rc;
  } else {
    hipFree(c->d_beta_atom);
    c->d_beta_atom = nullptr;
  }
  c->natoms = natoms;
  c->species_max = *std::max_element(species, species + natoms);
  c->nbeta_atoms = nbeta;
  return PW_OK;
}

int pw_set_rho_sphere(pw_ctx* c, int ng, const int64_t* idx, int base, const double* G, int nsp,
                      const double* vg) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (c->N == 0) return set_error(c, PW_ERR_STATE, "pw_set_grid must precede pw_set_rho_sphere");
  if (ng <= 0 || !idx || !G || nsp <= 0 || !vg) return set_error(c, PW_ERR_ARG, "pw_set_rho_sphere: bad arguments");
  int g0 = -1;
  for (int g = 0; g < ng; ++g)
    if (idx[g] - base == 0) {
      g0 = g;
      break;
    }
  if (g0 < 0) return set_error(c, PW_ERR_ARG, "pw_set_rho_sphere: G = 0 is not in the sphere");
  int rc = build_kdata(c, c->krho, ng, idx, base, G, "pw_set_rho_sphere");
  if (rc) return rc;
  c->krho_g0 = g0;
  rc = upload(c, &c->d_vg, vg, (size_t)ng * nsp);
  if (rc) return rc;
  c->nsp_loc = nsp;
  return PW_OK;
}

int pw_E_Ps_loc_grad(pw_ctx* c, const double* rho, double* E, double* dE_dvg, double* dE_dtau) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  return loc_grad(c, rho, E, dE_dvg, dE_dtau);
}

int pw_calc_forces_Ps_loc(pw_ctx* c, const double* rho, double* F) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (!F) return set_error(c, PW_ERR_ARG, "pw_calc_forces_Ps_loc: null F");
  int rc = loc_grad(c, rho, nullptr, nullptr, F);
  if (rc) return rc;
  for (int i = 0; i < 3 * c->natoms; ++i) F[i] = -F[i];
  return PW_OK;
}

int pw_calc_forces_Ps_nloc_dev(pw_ctx* c, int ik, int nb, const void* psi, int64_t ldpsi, const double* w,
                               double* F, double* E) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  return nloc_forces(c, ik, nb, reinterpret_cast<const cplx<double>*>(psi), ldpsi, w, F, E);
}

int pw_energies_psi_dev(pw_ctx* c, int ik, int nb, const void* psi, int64_t ldpsi, const double* w, double* E) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  return psi_energies(c, ik, nb, reinterpret_cast<const cplx<double>*>(psi), ldpsi, w, E);
}

int pw_calc_forces_Ps_nloc(pw_ctx* c, int ik, int nb, const void* psi, int64_t ldpsi, const double* w,
                           double* F, double* E) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (nb <= 0) return nb == 0 ? PW_OK : set_error(c, PW_ERR_ARG, "calc_forces_Ps_nloc: nb < 0");
  if (!psi) return set_error(c, PW_ERR_ARG, "calc_forces_Ps_nloc: null psi");
  const size_t bp = (size_t)ldpsi * nb * 16;
  int rc = ensure(c, &c->d_tmp, &c->tmp_bytes, bp);
  if (rc) return rc;
  hipError_t e = hipMemcpyAsync(c->d_tmp, psi, bp, hipMemcpyHostToDevice, c->stream);
  if (e != hipSuccess) return hip_check(c, e, "psi upload");
  return nloc_forces(c, ik, nb, reinterpret_cast<const cplx<double>*>(c->d_tmp), ldpsi, w, F, E);
}

}  // extern "C"
