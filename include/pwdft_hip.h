/*
 * pwdft_hip.h -- C ABI of libpwdft_hip.so, the MI355X (gfx950) implementation of
 * PWDFT.jl's Kohn-Sham Hamiltonian application, density accumulation and the dense
 * products of its block eigensolvers.
 *
 * Conventions (identical to the reference's Julia data):
 *  - complex numbers are interleaved (re, im) double pairs, i.e. Julia ComplexF64 /
 *    C double _Complex / hipDoubleComplex;
 *  - wave functions are column-major Ngw x Nstates with an explicit leading dimension;
 *  - real-space grid arrays are flat column-major (x fastest), ip = i + n1*(j + n2*k);
 *  - index maps are passed as int64 with an explicit base (1 for Julia's 1-based maps);
 *  - k-point and spin indices are 0-based here (the Julia shim subtracts one).
 *
 * Every entry point returns PW_OK (0) or a negative PW_ERR_* code; the message of the
 * last failure is available from pw_last_error().  No C++ exception crosses the ABI.
 * One context owns one HIP stream and its device buffers; a context is not thread
 * safe (the reference's Hamiltonian cursor Ham.ik/Ham.ispin is not re-entrant either).
 * Entry points taking host pointers are synchronous with respect to those buffers.
 * Entry points with the _dev suffix take device pointers (from pw_malloc) and are
 * asynchronous on the context stream; pw_synchronize() waits for them.
 */
#ifndef PWDFT_HIP_H
#define PWDFT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pw_ctx pw_ctx;

enum {
  PW_OK = 0,
  PW_ERR_ARG = -1,          /* invalid argument (ArgumentError / @assert in the reference) */
  PW_ERR_HIP = -2,          /* HIP runtime failure */
  PW_ERR_STATE = -3,        /* grid / k-point / potential not set up */
  PW_ERR_NOT_POSDEF = -4,   /* reserved for Cholesky failures (PosDefException) */
  PW_ERR_NOMEM = -5         /* device allocation failed */
};

/* op_H term mask and options (flags argument) */
enum {
  PW_TERM_K = 1u << 0,      /* op_K            (op_K.jl:44-75)          */
  PW_TERM_VLOC = 1u << 1,   /* op_V_loc        (op_V_loc.jl:33-92)      */
  PW_TERM_NLOC = 1u << 2,   /* op_V_Ps_nloc    (op_V_Ps_nloc.jl:39-78)  */
  PW_TERM_ALL = 7u,         /* op_H            (op_H.jl:4-27)           */
  PW_ACCUMULATE = 1u << 4,  /* add into the output (the reference's `!` forms of op_K/op_V_loc/op_V_Ps_nloc) */
  PW_MIXED = 1u << 5        /* FP32 FFT and FP32-MFMA projector products (mixed precision) */
};

/* ---- context ------------------------------------------------------------------- */

/* Create a context on HIP device `device`.  No reference counterpart (the reference
 * keeps this state inside Hamiltonian, Hamiltonian.jl:3-21). */
int pw_ctx_create(pw_ctx** ctx, int device);
int pw_ctx_destroy(pw_ctx* ctx);
const char* pw_last_error(pw_ctx* ctx);
/* Bands processed per fused FFT launch (intermediates live in HBM/L2 per batch). */
int pw_set_band_batch(pw_ctx* ctx, int nb);

/* Tuning knobs set by the autotuner (pwdft_hip/autotune.py); no reference counterpart. */
enum {
  PW_TUNE_GEMM_CN_SHAPE = 0,  /* A^H B products: -1 = chosen by N (default), 0 = 64x64 tile, 1 = 128x64,
                                 2 = 64x128, 3 = 128x128, 4 = 128x32, 5 = 64x32 */
  PW_TUNE_GEMM_NN_SHAPE = 1,  /* A B products: same codes */
  PW_TUNE_GEMM_CN_WG = 2,     /* A^H B split-K target workgroup count (0 = default 1024) */
  PW_TUNE_GEMM_NN_WG = 3,     /* A B split-K target workgroup count (0 = default 512) */
  PW_TUNE_BAND_BATCH = 4,     /* bands per fused FFT launch (same as pw_set_band_batch) */
  PW_TUNE_PLANE_COMPRESSED = 5, /* 1: always use the compressed z-plane kernels (they are used
                                   anyway when the padded plane does not fit in LDS); 2: always
                                   the global-plane pass (compressed planes in HBM, used anyway
                                   when even the compressed plane exceeds LDS) */
  PW_TUNE_OVERLAP = 6,          /* 1 (default): op_H runs its nonlocal products on a second
                                   stream beside the FFT passes; 0: one stream, in sequence */
  PW_TUNE_COL_CHUNK = 7,        /* z-columns per workgroup of the column FFT kernels (0: the
                                   default, as many as 48 KiB of LDS hold) */
  PW_TUNE_GEMM_SPLIT_BF16 = 8,  /* the mixed path's complex GEMMs on BF16 MFMA with each operand
                                   split in three BF16 parts (FP32-level accuracy): 1 always,
                                   0 never, -1 (default) for products of >= 2^33 complex
                                   multiply-adds */
  PW_TUNE_EIG_PATH = 9,         /* Hermitian eigensolves: 0 (default) the device tridiagonal
                                   reduction (one persistent launch, matrix in LDS) wherever
                                   the order fits; 1 rocSOLVER's ZHEEVD throughout */
  PW_TUNE_LOBPCG_LOCK = 10,     /* lock threshold of pw_diag_lobpcg_dev: 0 (default) the
                                   reference's 10 eps; k > 0: 10 * 10^-k (tests of the branch) */
  PW_TUNE_FUSED_BATCH = 11,     /* pw_op_H_batch_dev: 1 (default) one fused launch, a workgroup
                                   per (k-point, band), wherever the band's grid fits one
                                   workgroup's LDS (20^3 in FP64); on a cubic 20^3 grid its line
                                   transforms run in registers on a box kept at the active x
                                   only; 2: the fused launch with the whole box in LDS and
                                   wavefront-shared line transforms; 0: the three-stage column /
                                   plane / column launches and batched GEMMs */
  PW_TUNE_LOBPCG_GRAPH = 12,    /* pw_diag_lobpcg_batch_dev: 1 (default) replay each step from
                                   a captured hipGraph once it has run eagerly (orders 3 Nst <=
                                   64); 0: every step enqueued eagerly */
  PW_TUNE_FUSED_NT = 13,        /* threads per workgroup of the fused batched op_H: 256, 512 or
                                   1024; 0 (default): the workgroups one CU's LDS holds share
                                   1024 threads (whole-box kernel), 256 (register-line kernel,
                                   which takes 256 or 512) */
  PW_TUNE_EIG_SYNC = 14,        /* the device tridiagonal reduction's cross-workgroup exchanges:
                                   3 (default) by order: 2 up to n = 640, else 1;
                                   2 one exchange per column (the owner of the next
                                   column publishes it with its p segments and every
                                   workgroup forms the next reflector itself), device-coherent
                                   atomic accesses; 1 two exchanges per column (the owner
                                   publishes the next reflector), device-coherent atomic
                                   accesses, each producer waiting for its stores before the
                                   flag (n = 480: 10.8 ms, n = 960: 21.4 ms); 0 two exchanges
                                   ordered by agent-scope fences (12.2, 33.4 ms) */
  PW_TUNE_FFT_ORDER = 15,       /* radix plan of the FFT lines: 0 (default) largest radices
                                   first (72 = 9 x 8, compile-time specialised); 1 the same
                                   radices in reverse order; 2 radices 5, 4, 3, 2 only */
  PW_TUNE_BATCH_GEMM_SMALL = 16, /* pw_zgemm_batch_dev and the batched solvers: 1 (default) the
                                   small-product kernels (A B with k <= 64 and n <= 32: one
                                   thread per row; A^H B with m, n <= 32: one workgroup per
                                   product) where every product of the batch fits them; 0: the
                                   16 x 16-tile kernel for all */
  PW_TUNE_EIG_WG = 17,          /* the device tridiagonal reduction: at least this many
                                   workgroups share the matrix (default 64: n = 320 5.2 -> 4.2
                                   ms, n = 480 7.5 -> 6.7 ms for the whole eigensolve); 0: the
                                   fewest whose column shares fit in LDS, at least 8 */
  PW_TUNE_EIG_TRI = 18,         /* eigenpairs of the reduced tridiagonal matrix: 1 (default)
                                   only the wanted lowest ones on the device (multisection
                                   Sturm bisection, inverse iteration, orthogonalisation and
                                   Rayleigh-Ritz within runs of close eigenvalues) in the
                                   solvers and pw_heev_lowest_dev, rocSOLVER ZSTEDC for
                                   pw_heevd_dev; 2 the device path for pw_heevd_dev as well;
                                   0 ZSTEDC everywhere */
  PW_TUNE_CHOL_REC = 19         /* inv(cholesky(A).U) above order 64 in the solvers and in
                                   pw_chol_inv_batch_dev with one matrix: 1 (default) recursive
                                   2 x 2 blocks on the device (one-workgroup kernels for the
                                   diagonal blocks of order <= 64, complex GEMMs for the rest);
                                   0 rocSOLVER ZPOTRF + ZTRTRI */
};
int pw_tune_set(pw_ctx* ctx, int key, int value);

/* ---- set-up: the state op_H reads from Hamiltonian ------------------------------- */

/* FFT grid of PWGrid (PWGrid.jl:31-47: Ns, CellVolume). */
int pw_set_grid(pw_ctx* ctx, int n1, int n2, int n3, double cell_volume);

/* Wave-function sphere of k-point ik: pw.gvecw.idx_gw2r[ik] (PWGrid.jl:314) and the
 * G+k vectors (3 x ngw, column-major) of pw.gvec.G[:, idx_gw2g[ik]] .+ k (op_K.jl:69). */
int pw_set_kpoint(pw_ctx* ctx, int ik, int ngw, const int64_t* idx_gw2r, int index_base,
                  const double* gk);

/* Total local potential Ham.potentials.Total[:, ispin] (Potentials.jl:7-14), N values. */
int pw_set_vloc(pw_ctx* ctx, int ispin, const double* V);

/* Nonlocal GTH data: D (nbeta x nbeta, column-major, block diagonal, D[ib,jb] = h[l,i,j]
 * of PsPotNL prj2beta blocks, PsPotNL.jl:147-195) and betaNL[ik] (ngw x nbeta,
 * PsPotNL.jl:16-77).  pw_set_dij must precede pw_set_projectors. */
int pw_set_dij(pw_ctx* ctx, int nbeta, const double* D);
int pw_set_projectors(pw_ctx* ctx, int ik, int nbeta, const void* betaNL, int64_t ld);

/* ---- Hamiltonian application ----------------------------------------------------- */

/* Hpsi = (terms of) H psi for nb bands at (ik, ispin): op_H (op_H.jl:4-27) with the
 * term mask in flags; PW_ACCUMULATE adds into Hpsi as op_K!/op_V_loc!/op_V_Ps_nloc! do
 * (op_K.jl:44-75, op_V_loc.jl:33-92, op_V_Ps_nloc.jl:39-78).  Host pointers. */
int pw_op_H(pw_ctx* ctx, int ik, int ispin, int nb, const void* psi, int64_t ldpsi, void* Hpsi,
            int64_t ldh, unsigned flags);
/* Same on device pointers. */
int pw_op_H_dev(pw_ctx* ctx, int ik, int ispin, int nb, const void* psi, int64_t ldpsi, void* Hpsi,
                int64_t ldh, unsigned flags);
/* pw_op_H_dev with no device allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate (code under HIP stream capture, serving loops that forbid
 * hipMalloc).  Every workspace (the FFT work array, beta^H psi, the projector products'
 * split-K partial sums) lives in one fixed 64 MiB device array per GPU, held until the
 * process ends; the bands go through the FFT
 * passes in as many launches as that array needs.  Rejected with PW_ERR_ARG: a k-point
 * whose Nbeta x nb block of beta^H psi exceeds 8 MiB, or one of whose bands needs more
 * than 48 MiB of FFT work array.  Results equal pw_op_H_dev's to rounding (the split of
 * the projector products' K may differ).  Calls on one GPU share the array: calls that
 * overlap in time on one GPU (several contexts or threads) race. */
int pw_op_H_dev_noalloc(pw_ctx* ctx, int ik, int ispin, int nb, const void* psi, int64_t ldpsi, void* Hpsi,
                        int64_t ldh, unsigned flags);

/* op_H over several k-points in one launch per stage (the k loop of op_H(Ham, psiks) for a
 * BlochWavefunc, op_H.jl:29-50): k-point iks[i] applies to psi[i] (Ngw[iks[i]] x nb, leading
 * dimension ldpsi[i]) and writes Hpsi[i]; every k-point has the same band count nb.  Flags
 * as pw_op_H.  Device pointers.  With unchanged arguments, consecutive calls upload no
 * descriptors. */
int pw_op_H_batch_dev(pw_ctx* ctx, int nk, const int* iks, int ispin, int nb, void* const* psi,
                      const int64_t* ldpsi, void* const* Hpsi, const int64_t* ldh, unsigned flags);

/* Kprec for several k-points (Kprec.jl:10-28), device pointers. */
int pw_kprec_batch_dev(pw_ctx* ctx, int nk, const int* iks, int nb, void* const* in, const int64_t* ldin,
                       void* const* out, const int64_t* ldout);

/* Copy column blocks (rows[i] x cols[i]) from src[i] to dst[i] for a batch, one launch;
 * device pointers.  No reference counterpart (the eigensolver's block assignments). */
int pw_copy_cols_batch_dev(pw_ctx* ctx, int nbatch, const int* rows, const int* cols, void* const* src,
                           const int64_t* lds, void* const* dst, const int64_t* ldd);

/* Kinetic preconditioner Kprec (Kprec.jl:10-28): out = in / (1 + |G+k|^2), device pointers. */
int pw_kprec_dev(pw_ctx* ctx, int ik, int nb, const void* in, int64_t ldin, void* out, int64_t ldout);

/* betaNL_psi = (betaNL^H psi) as an nbeta x nb column-major complex matrix: the transpose
 * of calc_betaNL_psi (PsPotNL.jl:242-254), device pointers. */
int pw_betaNL_psi_dev(pw_ctx* ctx, int ik, int nb, const void* psi, int64_t ldpsi, void* out,
                      int64_t ldout, unsigned flags);

/* ---- density --------------------------------------------------------------------- */

/* calc_rhoe! (calc_rhoe.jl:9-80) split into: zero the device accumulator, add the bands
 * of one k-point with weights w[ist] = wk[ik] * Focc[ist, ik], and read rho (N doubles). */
int pw_rhoe_zero(pw_ctx* ctx);
int pw_rhoe_accum(pw_ctx* ctx, int ik, int nb, const void* psi, int64_t ldpsi, const double* w,
                  unsigned flags);
int pw_rhoe_accum_dev(pw_ctx* ctx, int ik, int nb, const void* psi, int64_t ldpsi, const double* w,
                      unsigned flags);
int pw_rhoe_get(pw_ctx* ctx, double* rho);
/* Copy of the accumulated density into a device array (N doubles). */
int pw_rhoe_get_dev(pw_ctx* ctx, double* rho);

/* ---- density side of the SCF iteration on the device ------------------------------ */

/* Exchange-correlation functionals of update! (Hamiltonian.jl:526-533). */
enum { PW_XC_VWN = 0, PW_XC_PBE = 1 };

/* Ham.potentials.Ps_loc (Hamiltonian.jl:119-141), N values, host pointer. */
int pw_set_vps_loc(pw_ctx* ctx, const double* Vps);

/* Symmetry operations of symmetrize_rhoe! (RhoeSymmetrizer.jl:111-236, SymmetryInfo.jl):
 * for operation s, M[9 s .. 9 s + 8] (row-major 3 x 3 integers) maps the crystal coordinates
 * m of a G-vector to M m (= R_s^{-T} m), t[3 s .. 3 s + 2] is its fractional translation;
 * gm[3 g .. 3 g + 2] are the integer crystal coordinates of rho-sphere entry g (ng = the
 * size given to pw_set_rho_sphere, which must come first).  The sphere must be closed
 * under every M (PW_ERR_ARG otherwise).  Host pointers. */
int pw_set_symmetry(pw_ctx* ctx, int nsym, const int* M, const double* t, int ng, const int* gm);

/* update!(Ham, Rhoe) for one spin (Hamiltonian.jl:510-545) on device arrays of N doubles:
 * V_H by Poisson_solve! (Poisson_solve.jl:22-39), V_xc by Slater + VWN or PBE, and
 * V_Total = V_Ps_loc + V_H + V_xc written to VH / Vxc / Vtot (each may be NULL); with
 * install != 0 V_Total also becomes the op_H potential of ispin (as pw_set_vloc).  E (host,
 * may be NULL) receives {E_Ps_loc, E_Hartree, E_xc} of this density, the local terms of
 * calc_energies (calc_energies.jl:245-333).  Synchronous when E is given. */
int pw_update_potential_dev(pw_ctx* ctx, int ispin, int xc, int install, const double* rho, double* VH,
                            double* Vxc, double* Vtot, double* E);

/* V_Total = V_Ps_loc + V_H + V_xc from device arrays (N doubles), installed as the op_H
 * potential of ispin, for drivers that mix the potentials themselves
 * (KS_solve_SCF_potmix.jl:236-266); Vtot (device, may be NULL) receives V_Total. */
int pw_set_vloc_parts_dev(pw_ctx* ctx, int ispin, const double* VH, const double* Vxc, double* Vtot);

/* Simple mixing of KS_solve_SCF! (KS_solve_SCF.jl:240-281): rho_out = beta rho_new +
 * (1 - beta) rho_in, then rho_out = max(rho_out, eps(Float64)); device arrays, rho_out may
 * alias either input. */
int pw_mix_rhoe_dev(pw_ctx* ctx, double beta, const double* rho_new, const double* rho_in, double* rho_out);

/* Vector primitives of the density mixers (mix_broyden.jl, mix_pulay.jl, mix_rpulay.jl,
 * mix_ppulay.jl, mix_anderson.jl) on device vectors of n doubles; the host keeps only the
 * small history matrices.  a, b, x are host arrays of device pointers.
 *   pw_vec_dots_dev:    out[q] = a[q] . b[q] for q < npair <= 64 (host out; synchronous,
 *                       summation order fixed, so repeated calls give the same bits);
 *   pw_vec_lincomb_dev: y = ay * y + sum_j coef[j] x[j], m <= 32 (y is not read when
 *                       ay == 0; y may be one of the x[j]). */
int pw_vec_dots_dev(pw_ctx* ctx, int64_t n, int npair, const double* const* a, const double* const* b, double* out);
int pw_vec_lincomb_dev(pw_ctx* ctx, int64_t n, int m, const double* coef, const double* const* x, double ay,
                       double* y);
/* A[:, j] *= s[j] (s: host, ncols values) for a column-major complex device block: the
 * occupation weights wk Focc of calc_grad (calc_grad.jl:1-57). */
int pw_scale_cols_dev(pw_ctx* ctx, int rows, int ncols, const double* s, void* A, int64_t ld);
/* y = max(y, lo) element-wise (the eps() clamp of KS_solve_SCF.jl:277-281 after mixing). */
int pw_vec_clamp_dev(pw_ctx* ctx, int64_t n, double lo, double* y);
/* mix_adaptive! (mix_adaptive.jl:34-54) on device vectors: mu (the density, updated), nu
 * (the new density), beta and f (per-point state, n doubles each). */
int pw_mix_adaptive_dev(pw_ctx* ctx, int64_t n, double beta0, double betamax, double* mu, const double* nu,
                        double* beta, double* f);

/* symmetrize_rhoe! in place on a device density: rho(G) on the rho sphere is replaced by
 * 1/Nsym sum_s rho(M_s m) exp(2 pi i t_s . M_s m); components outside the sphere are kept.
 * A no-op when pw_set_symmetry has not been called or gave one operation. */
int pw_symmetrize_rhoe_dev(pw_ctx* ctx, double* rho);

/* ---- per-geometry set-up on the device (relaxation / MD rebuilds) ------------------- */

/* betaNL[ik] of PsPotNL (PsPotNL.jl:16-77) built on the device for the current atoms
 * (pw_set_atoms) and k-point ik (pw_set_kpoint's G+k): column j is projector desc[4j..4j+3]
 * = (atom 0-based, l, m, iprj 1-based), beta = (-i)^l Ylm_real(l, m, g) eval_proj_G(l, iprj,
 * |g|) exp(-i g.tau) (Ylm_real.jl:5-107, PsPot_GTH.jl:338-391) with rc[4 s + l] the GTH rc of
 * species s.  Replaces pw_set_projectors (also derives beta D and the FP32 copies); nbeta
 * must equal pw_set_dij's. */
int pw_build_projectors_dev(pw_ctx* ctx, int ik, int nbeta, const int* desc, int nsp, const double* rc);
/* V_Ps_loc (Hamiltonian.jl:119-141) for the current atoms from the rho sphere and V_sp(|G|)
 * of pw_set_rho_sphere: V(r) = Re sum_G sum_a V_sp(a)(G) exp(-i G.tau_a) / Omega e^{iG.r};
 * becomes the V_Ps_loc of pw_update_potential_dev; copied to Vps (host, N) unless NULL. */
int pw_build_vps_loc_dev(pw_ctx* ctx, double* Vps);

/* ---- forces and energy gradients (the adjoint of the pseudopotential energies) ---- */

/* Atoms for the force kernels: positions tau (3 x natoms, bohr, column-major), 0-based
 * species per atom, and the atom of every projector (beta_atom[nbeta], 0-based; the
 * (atom, l, m, iprj) ordering of PsPotNL.jl:45-57).  Call again after atoms move
 * (together with pw_set_projectors, whose betaNL depends on tau). */
int pw_set_atoms(pw_ctx* ctx, int natoms, const double* tau, const int* species, int nbeta,
                 const int* beta_atom);

/* The rho-sphere G-vectors (pw.gvec: idx_g2r minus base, G as 3 x ng) and the local
 * pseudopotential form factors vg[ng x nsp] = V_sp(|G|) as eval_Vloc_G! returns them
 * (PsPot_GTH.jl:164-189, no 1/Omega), one column per species. */
int pw_set_rho_sphere(pw_ctx* ctx, int ng, const int64_t* idx, int base, const double* G, int nsp,
                      const double* vg);

/* E_Ps_loc = sum_G sum_sp S_sp(G) V_sp(G) conj(rho(G)) with rho(G) = FFT[rho] / N and
 * S_sp(G) = sum_{a in sp} exp(-i G.tau_a): the energy of calc_E_Ps_loc (calc_energies.jl)
 * written on the rho sphere, and its gradients at fixed rho
 *   dE_dvg[g + ng*sp] = d E_Ps_loc / d V_sp(G_g) = Re(S_sp(G) conj(rho(G))),
 *   dE_dtau[3a + x]   = d E_Ps_loc / d tau_a,x   = -F_Ps_loc (calc_forces_Ps_loc.jl:13-71).
 * rho: host array of N doubles (spin-summed), or NULL for the density accumulated on the
 * device by pw_rhoe_accum*.  Any output pointer may be NULL.  FP64. */
int pw_E_Ps_loc_grad(pw_ctx* ctx, const double* rho, double* E, double* dE_dvg, double* dE_dtau);

/* calc_forces_Ps_loc(Ham, Rhoe) (calc_forces_Ps_loc.jl:1-71): F (3 x natoms) overwritten. */
int pw_calc_forces_Ps_loc(pw_ctx* ctx, const double* rho, double* F);

/* The GTH nonlocal force of one k-point (calc_forces_Ps_nloc.jl:291-343 with calc_dbetaNL
 * :347-391 and calc_dbetaNL_psi :393-411), added into F (3 x natoms):
 *   F[:, a] += 2 sum_ist w[ist] sum_{i,j in a} h_ij Re(conj(<beta_i|psi>) <i(G+k) beta_j|psi>),
 * w[ist] = wk[ik] * Focc[ist, ik] (host array).  When E is not NULL the band energy
 * sum_ist w[ist] sum_ij h_ij conj(<beta_i|psi>) <beta_j|psi> (calc_E_Ps_nloc) is added to
 * *E.  Device pointers for psi; FP64. */
int pw_calc_forces_Ps_nloc_dev(pw_ctx* ctx, int ik, int nb, const void* psi, int64_t ldpsi,
                               const double* w, double* F, double* E);
/* Same with psi in host memory. */
int pw_calc_forces_Ps_nloc(pw_ctx* ctx, int ik, int nb, const void* psi, int64_t ldpsi,
                           const double* w, double* F, double* E);

/* The psi-dependent terms of calc_energies for one k-point with band weights
 * w[st] = wk[ik] * Focc[st, ik] (host, nb values): E[0] += sum_st w_st 1/2 sum_g |G+k|^2
 * |psi_g,st|^2 (calc_E_kin, calc_energies.jl:342-375) and E[1] += sum_st w_st sum_ij D_ij
 * Re(conj(B_i,st) B_j,st) with B = betaNL^H psi (calc_E_Ps_nloc for GTH,
 * calc_energies.jl:7-58).  psi on the device; synchronous. */
int pw_energies_psi_dev(pw_ctx* ctx, int ik, int nb, const void* psi, int64_t ldpsi, const double* w, double* E);

/* ---- dense products of diag_LOBPCG! / diag_davidson! ------------------------------ */

/* C = alpha * op(A) * B + beta * C with op(A) = A (transa 0, A m x k) or A^H (transa 1,
 * A k x m); complex, column-major, device pointers.  These are the X'HX, Q'HQ, Q'Q, Q*U,
 * HQ*U and W*U products of diag_LOBPCG.jl:85-203 and diag_davidson.jl:122-150.
 * flags & PW_MIXED computes in FP32 MFMA with FP32 accumulation. */
int pw_zgemm_dev(pw_ctx* ctx, int transa, int m, int n, int k, const double* alpha, const void* A,
                 int64_t lda, const void* B, int64_t ldb, const double* beta, void* C, int64_t ldc,
                 unsigned flags);

/* A batch of independent products C_i = alpha op(A_i) B_i + beta C_i with per-product sizes
 * (m_i, n_i, k_i) and leading dimensions, op as pw_zgemm_dev (transa shared); complex
 * column-major, device pointers, FP64.  The per-k Gram products and column updates of
 * diag_LOBPCG! run over all k-points at once through this (diag_LOBPCG.jl:85-203). */
int pw_zgemm_batch_dev(pw_ctx* ctx, int transa, int nbatch, const int* m, const int* n, const int* k,
                       const double* alpha, void* const* A, const int64_t* lda, void* const* B,
                       const int64_t* ldb, const double* beta, void* const* C, const int64_t* ldc);

/* ---- small dense FP64 problems of the eigensolvers, on the device ------------------- */

/* diag_LOBPCG!(Ham, psi) for k-point ik and spin ispin (diag_LOBPCG.jl:51-225, called per
 * k from diag_LOBPCG!(Ham, psiks), :1-28, which KS_solve_SCF! calls at KS_solve_SCF.jl:
 * 189-191): X (device, Ngw x nst, leading dimension ldx, orthonormal columns) is overwritten
 * with the Ritz vectors, evals (host, nst) with the Ritz values.  tol, niter_max and
 * nstates_conv (0 = nst) as the reference's tol, NiterMax and Nstates_conv: iteration stops
 * once nstates_conv Ritz values moved by less than tol.  flags: PW_MIXED for the mixed
 * precision products.  stats (host, 3 ints, may be NULL): iterations, op_H columns applied,
 * lock-branch iterations.  Requires 3 nst <= Ngw.  PW_ERR_NOT_POSDEF where the reference's
 * cholesky() throws PosDefException.  Synchronises once per iteration. */
int pw_diag_lobpcg_dev(pw_ctx* ctx, int ik, int ispin, int nst, void* X, int64_t ldx, double tol, int niter_max,
                       int nstates_conv, unsigned flags, double* evals, int* stats);

/* diag_davidson!(Ham, X) for k-point ik and spin ispin (diag_davidson.jl:50-191, called per k
 * from diag_davidson!(Ham, psiks), :12-38, which KS_solve_SCF! calls when the solver is
 * Davidson): X (device, Ngw x nst, leading dimension ldx, orthonormal columns) is
 * overwritten with the Ritz vectors of the last 2 nst search space, evals (host, nst) with
 * the Ritz values.  tol, niter_max, nstates_conv and flags as pw_diag_lobpcg_dev.  stats
 * (host, 3 ints, may be NULL): iterations, op_H columns applied, 0.  Requires 2 nst <= Ngw.
 * Synchronises once per iteration. */
int pw_diag_davidson_dev(pw_ctx* ctx, int ik, int ispin, int nst, void* X, int64_t ldx, double tol, int niter_max,
                         int nstates_conv, unsigned flags, double* evals, int* stats);

/* diag_LOBPCG!(Ham, psiks) over the nk k-points iks at once (diag_LOBPCG.jl:1-28 running
 * :51-225 for every k-point, KS_solve_SCF.jl:189-191): X[i] (device, Ngw[iks[i]] x nst,
 * leading dimension ldx[i], orthonormal columns) is overwritten with the Ritz vectors of
 * k-point iks[i], evals (host, nst x nk, column i for k-point iks[i]) with its Ritz values.
 * Every k-point iterates as diag_LOBPCG! would on its own -- the same |dlambda| test, the
 * same lock branch -- and leaves the batch once it has converged; each step of the k-points
 * still active is one batched launch per operation (op_H of all their bands in one launch),
 * and with 3 nst <= 64 (small dense problems on one workgroup each) the steps are captured
 * into hipGraphs once seen and replayed (PW_TUNE_LOBPCG_GRAPH).  Synchronises once per
 * iteration.  tol, niter_max, nstates_conv, flags as pw_diag_lobpcg_dev.  stats (host,
 * 3 x nk ints, may be NULL): per k-point iterations, op_H columns, lock-branch iterations.
 * Requires 3 nst <= Ngw of every k-point. */
int pw_diag_lobpcg_batch_dev(pw_ctx* ctx, int nk, const int* iks, int ispin, int nst, void* const* X,
                             const int64_t* ldx, double tol, int niter_max, int nstates_conv, unsigned flags,
                             double* evals, int* stats);

/* The host LAPACK calls of diag_LOBPCG! / diag_davidson! / ortho_sqrt (cholesky(Hermitian(C)),
 * inv(U), eigen(Hermitian(T), Hermitian(G)), eigen(Hermitian(S))) on device matrices
 * (diag_LOBPCG.jl:106-118, 176-203, 217-222).  Column-major complex, upper triangle read.
 *   pw_herm_fix_dev: mode 0 A <- Hermitian(A); 1 zero the strict lower triangle;
 *                    2 A <- (A + A^H) / 2.
 *   pw_potrf_dev:  A = U^H U, U into the upper triangle (PW_ERR_NOT_POSDEF on failure).
 *   pw_trtri_dev:  upper-triangular A <- inv(A).
 *   pw_heevd_dev:  eigenvectors into A (ascending), eigenvalues into w (host, n).
 *   pw_hegvd_dev:  T x = l G x; eigenvectors (G-orthonormal) into A, B overwritten. */
int pw_herm_fix_dev(pw_ctx* ctx, int n, void* A, int64_t lda, int mode);
int pw_potrf_dev(pw_ctx* ctx, int n, void* A, int64_t lda);
int pw_trtri_dev(pw_ctx* ctx, int n, void* A, int64_t lda);
int pw_heevd_dev(pw_ctx* ctx, int n, void* A, int64_t lda, double* w);
/* eigvals(Hermitian(A)) only (A is destroyed). */
int pw_heev_values_dev(pw_ctx* ctx, int n, void* A, int64_t lda, double* w);
/* The m lowest eigenpairs of Hermitian(A), as diag_LOBPCG! keeps the Nstates lowest Ritz
 * pairs of eigen(Hermitian(T)) (diag_LOBPCG.jl:176-185): eigenvectors into A's first m
 * columns (ascending), eigenvalues into w (host, m); the other columns of A are destroyed.
 * 1 <= m <= n. */
int pw_heev_lowest_dev(pw_ctx* ctx, int n, int m, void* A, int64_t lda, double* w);
int pw_hegvd_dev(pw_ctx* ctx, int n, void* A, int64_t lda, void* B, int64_t ldb, double* w);
/* eigen(Hermitian(A_i)) for nbatch matrices A_i = A + i*strideA (elements), each n x n with
 * leading dimension lda: the per-k-point Rayleigh-Ritz problems of diag_LOBPCG! run over
 * many k-points (diag_LOBPCG.jl:176-203, 217-222).  Eigenvectors into A_i (ascending),
 * eigenvalues into w (host, n x nbatch, column i for matrix i). */
int pw_heevd_batch_dev(pw_ctx* ctx, int n, int nbatch, void* A, int64_t lda, int64_t strideA, double* w);
/* pw_herm_fix_dev on each of nbatch matrices A_i = A + i*strideA (nbatch <= 65535). */
int pw_herm_fix_batch_dev(pw_ctx* ctx, int n, int nbatch, void* A, int64_t lda, int64_t strideA, int mode);
/* A_i <- inv(cholesky(Hermitian(A_i)).U) for nbatch matrices A_i = A + i*strideA, the strict
 * lower triangle zeroed: the C = U^H U factorisation and the inv(U) of every k-point's Gram
 * matrix in one call (diag_LOBPCG.jl:106-118, 193-203 over many k-points).
 * PW_ERR_NOT_POSDEF names the first matrix that is not positive definite. */
int pw_chol_inv_batch_dev(pw_ctx* ctx, int n, int nbatch, void* A, int64_t lda, int64_t strideA);
/* The same without waiting: the potrf and trtri info of matrix i go to d_info[i] and
 * d_info[nbatch + i] (device memory, 2 * nbatch ints; nonzero = not positive definite /
 * singular), for the caller to read when it next synchronises.  The batched LOBPCG
 * iteration reads them together with its eigenvalues, once per iteration. */
int pw_chol_inv_batch_async_dev(pw_ctx* ctx, int n, int nbatch, void* A, int64_t lda, int64_t strideA, int* d_info);
/* pw_heevd_batch_dev without waiting: eigenvalues into d_w (device, n x nbatch), the
 * convergence info of matrix i into d_info[i] (device; nonzero = no convergence). */
int pw_heevd_batch_async_dev(pw_ctx* ctx, int n, int nbatch, void* A, int64_t lda, int64_t strideA, double* d_w,
                             int* d_info);

/* ---- device memory (for device-resident psi across LOBPCG, calc_rhoe and energies) */

int pw_malloc(pw_ctx* ctx, size_t bytes, void** dptr);
int pw_free(pw_ctx* ctx, void* dptr);
int pw_memcpy_h2d(pw_ctx* ctx, void* dst, const void* src, size_t bytes);
int pw_memcpy_d2h(pw_ctx* ctx, void* dst, const void* src, size_t bytes);
int pw_memcpy_d2d(pw_ctx* ctx, void* dst, const void* src, size_t bytes);
int pw_synchronize(pw_ctx* ctx);

/* ---- graph capture ----------------------------------------------------------------- */

/* Record the device work of the _dev entry points called between pw_graph_begin and
 * pw_graph_end into a hipGraph (stream capture of the context stream) and replay it with
 * pw_graph_launch: one launch for a whole fixed sequence, e.g. the batched op_H of an
 * eigensolver step.  Run the sequence once before capturing it, so that every workspace and
 * descriptor array exists: during capture nothing may allocate, upload or synchronise, and
 * an entry point that would returns PW_ERR_STATE (end the capture with pw_graph_end and
 * free the result).  Replays read and write the same device addresses as the captured
 * calls.  No reference counterpart. */
int pw_graph_begin(pw_ctx* ctx);
int pw_graph_end(pw_ctx* ctx, int* graph_id);
int pw_graph_launch(pw_ctx* ctx, int graph_id);
int pw_graph_free(pw_ctx* ctx, int graph_id);

/* Strided 2-D copy of complex columns between device buffers: dst[:, j] = src[:, j]
 * for j < ncols, nrows complex each. */
int pw_copy_cols_dev(pw_ctx* ctx, int nrows, int ncols, const void* src, int64_t lds, void* dst,
                     int64_t ldd);

/* ---- timing ---------------------------------------------------------------------- */
/* Per-kernel timing slots (HIP events around each launch on the context stream, plus the
 * kernel's algorithmic HBM bytes per launch).  No reference counterpart. */
#define PW_PROF_SLOTS 8
enum {
  PW_PROF_COL_IFFT = 0,    /* sphere scatter + inverse z-FFT of the columns        */
  PW_PROF_PLANE_VLOC = 1,  /* y/x FFTs and the V(r) product on z-planes            */
  PW_PROF_COL_FFT = 2,     /* forward z-FFT + sphere gather + kinetic              */
  PW_PROF_GEMM_CN = 3,     /* A^H B products (betaNL^H psi, Gram/overlap)          */
  PW_PROF_GEMM_NN = 4,     /* A B products (beta (D B), Q U)                       */
  PW_PROF_PLANE_RHOE = 5,  /* y/x inverse FFTs and |psi|^2 accumulation            */
  PW_PROF_OTHER = 6
};
int pw_profile_enable(pw_ctx* ctx, int on);     /* on != 0 clears the counters */
int pw_profile_read(pw_ctx* ctx, int slot, double* total_ms, long* count, double* bytes);
/* Events on the context stream (the stream every kernel of the context runs on). */
int pw_event_record(pw_ctx* ctx, int slot);
int pw_event_elapsed_ms(pw_ctx* ctx, int slot0, int slot1, float* ms);

#ifdef __cplusplus
}
#endif

#endif /* PWDFT_HIP_H */
