// diag_LOBPCG! (diag_LOBPCG.jl:51-225) for one k-point and spin, entirely on the device
// behind the C ABI (pw_diag_lobpcg_dev), so that a Julia caller reaches it with one ccall.
//
// X, W, P and their H-images are column blocks of Q = [X W P] and HQ (Ngw x 3 Nst each) so
// that T = Q^H H Q and G = Q^H Q are single products per block column; only the upper block
// triangle of T and G is formed (the part Hermitian() reads).  The small dense problems the
// reference hands to LAPACK -- the Cholesky factor inverses of W^H W and P^H P, the
// generalised Rayleigh-Ritz eigenproblem T y = lambda G y of order 2 Nst or 3 Nst (reduced
// to standard form through G = U^H U), the eigenvalues of X^H H X -- stay in HBM.  The host
// gets the Nst lowest Ritz values once per iteration, for the reference's |d lambda|
// convergence test and lock decision, together with the status words of that iteration's
// factorisations: one synchronisation per iteration.
//
// The step order follows the reference (kept alike on purpose, so that iteration counts and
// converged results compare): R = HX - X (X^H H X), W = Kprec(R) (Kprec without 1/2),
// Cholesky re-orthonormalisation of W and of the new P, Q = [X W] in the first iteration and
// [X W P] after it, and the lock branch (diag_LOBPCG.jl:125-164) when some |d lambda| is at
// or below 10 eps, whose projection of the preconditioned W against Ylock (:130-133)
// rebinds a local in the reference and has no effect, so it is not performed.
//
// Precision: with PW_MIXED the Ngw-sized products other than the factorised Grams (W^H W,
// P^H P, G) run on FP32 MFMA with FP64 accumulation of the result; the Grams, the small
// dense problems and the eigenvalues stay FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

#include "pw_internal.h"

namespace pw {

namespace {

constexpr int kInfoSlots = 16;
const cplx<double> kOne = {1.0, 0.0}, kZero = {0.0, 0.0}, kMinusOne = {-1.0, 0.0};

__global__ void k_fill_i32(int* p, int n, int v) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

struct Lobpcg {
  pw_ctx* c;
  int ik, ispin, Nb, Nst;
  int prec;            // 1: mixed products
  unsigned hflags;     // op_H flags
  cplx<double> *Q, *HQ, *Tmp, *Sm, *Gm, *Cm, *S2, *S3;
  double *dw, *E;
  int* info;
  cplx<double>* L = nullptr;   // lock branch blocks (Nb x 5 Nst)
  long ncols = 0;
  int locks = 0;

  cplx<double>* q(int j) { return Q + (long)j * Nb; }
  cplx<double>* hq(int j) { return HQ + (long)j * Nb; }
  cplx<double>* tmp(int j) { return Tmp + (long)j * Nb; }

  int gemm(int ta, int m, int n, int k, const cplx<double>* A, long lda, const cplx<double>* B, long ldb,
           cplx<double>* C, long ldc, int p, cplx<double> alpha = kOne, cplx<double> beta = kZero) {
    return zgemm(c, ta, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, p);
  }
  int opH(const cplx<double>* src, cplx<double>* dst, int n) {
    ncols += n;
    return op_H_device(c, ik, ispin, n, src, Nb, dst, Nb, hflags);
  }
  int copy(int n, const cplx<double>* src, cplx<double>* dst) { return copy_cols(c, Nb, n, src, Nb, dst, Nb); }

  // U = chol(src^H src).U ; dst = src U^-1, hdst = hsrc U^-1 (S3 scratch, info slots s, s+1)
  int chol_inv_apply(const cplx<double>* src, const cplx<double>* hsrc, cplx<double>* dst, cplx<double>* hdst,
                     int n, int slot) {
    int rc = gemm(1, n, n, Nb, src, Nb, src, Nb, S3, n, 0);
    if (!rc) rc = chol_inv_enqueue(c, n, S3, n, info + slot);
    if (!rc) rc = gemm(0, Nb, n, n, src, Nb, S3, n, dst, Nb, prec);
    if (!rc) rc = gemm(0, Nb, n, n, hsrc, Nb, S3, n, hdst, Nb, prec);
    return rc;
  }

  // Generalised Hermitian eigenproblem T y = lambda G y of order n (T in Sm, G in Gm, both
  // leading dimension n, upper triangles read): the eigenvalues into dw, the nvec lowest
  // G-orthonormal eigenvectors into Cm (n x nvec, leading dimension n).  Info slots s..s+2.
  int rr_solve(int n, int nvec, int slot) {
    int rc = herm_fix_launch(c, n, 1, Sm, n, 0, 0);
    if (!rc) rc = chol_inv_enqueue(c, n, Gm, n, info + slot);             // Gm = U^-1
    if (!rc) rc = gemm(0, n, n, n, Sm, n, Gm, n, Cm, n, 0);                 // T U^-1
    if (!rc) rc = gemm(1, n, n, n, Gm, n, Cm, n, Sm, n, 0);                 // U^-H T U^-1
    if (!rc) rc = heev_enqueue(c, n, Sm, n, dw, nvec, E, info + slot + 2);
    if (!rc) rc = gemm(0, n, nvec, n, Gm, n, Sm, n, Cm, n, 0);              // U^-1 y
    return rc;
  }

  // Ngw-sized scratch of the lock branch, allocated on its first use.
  int lock_scratch() {
    if (L) return PW_OK;
    int rc = ensure(c, &c->d_lob2, &c->lob2_bytes, (size_t)Nb * 5 * Nst * sizeof(cplx<double>));
    if (rc) return rc;
    L = reinterpret_cast<cplx<double>*>(c->d_lob2);
    return PW_OK;
  }

  // Lock branch (diag_LOBPCG.jl:125-164).
  int lock_step(const std::vector<int>& ilock, const std::vector<int>& iact) {
    const int nl = (int)ilock.size(), na = (int)iact.size();
    int rc = lock_scratch();
    if (rc) return rc;
    ++locks;
    cplx<double>* Xl = L;
    cplx<double>* A = L + (long)nl * Nb;
    cplx<double>* HA = A + 2L * na * Nb;
    cplx<double>* T2 = HA + 2L * na * Nb;
    for (int j = 0; j < nl && !rc; ++j) rc = copy(1, q(ilock[j]), Xl + (long)j * Nb);
    for (int j = 0; j < na && !rc; ++j) rc = copy(1, q(iact[j]), A + (long)j * Nb);
    for (int j = 0; j < na && !rc; ++j) rc = copy(1, q(Nst + iact[j]), T2 + (long)j * Nb);
    if (!rc) rc = copy(na, T2, A + (long)na * Nb);
    // block -= Xlock (Xlock^H block)
    auto project = [&](cplx<double>* blk, int nc) {
      int r = gemm(1, nl, nc, Nb, Xl, Nb, blk, Nb, Sm, nl, prec);
      if (!r) r = gemm(0, Nb, nc, nl, Xl, Nb, Sm, nl, blk, Nb, prec, kMinusOne, kOne);
      return r;
    };
    if (!rc) rc = project(A, 2 * na);
    if (!rc) rc = opH(A, HA, 2 * na);
    if (!rc) rc = project(HA, 2 * na);
    cplx<double>* Aw = A + (long)na * Nb;
    cplx<double>* HAw = HA + (long)na * Nb;
    if (!rc) rc = chol_inv_apply(Aw, HAw, T2, Tmp, na, 9);
    if (!rc) rc = copy(na, T2, Aw);
    if (!rc) rc = copy(na, Tmp, HAw);
    const int n2 = 2 * na;
    if (!rc) rc = gemm(1, n2, n2, Nb, A, Nb, HA, Nb, Sm, n2, prec);
    if (!rc) rc = gemm(1, n2, n2, Nb, A, Nb, A, Nb, Gm, n2, 0);
    if (!rc) rc = rr_solve(n2, na, 11);
    // X = [Xlock, [Xact W] U]
    if (!rc) rc = copy(nl, Xl, q(0));
    if (!rc) rc = gemm(0, Nb, na, n2, A, Nb, Cm, n2, q(nl), Nb, prec);
    if (!rc) rc = opH(q(0), Tmp, Nst);
    if (!rc) rc = gemm(1, Nst, Nst, Nb, q(0), Nb, Tmp, Nb, S2, Nst, prec);
    if (!rc) rc = herm_fix_launch(c, Nst, 1, S2, Nst, 0, 0);
    if (!rc) rc = heev_enqueue(c, Nst, S2, Nst, dw, Nst, E, info + 14);
    if (!rc) rc = gemm(0, Nb, Nst, Nst, q(0), Nb, S2, Nst, Tmp, Nb, prec);
    if (!rc) rc = copy(Nst, Tmp, q(0));
    if (!rc) rc = opH(q(0), hq(0), Nst);
    return rc;
  }
};

int read_iteration(pw_ctx* c, const double* dw, int nst, const int* info, double* ev, int* hinfo) {
  hipError_t e = hipMemcpyAsync(ev, dw, nst * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(hinfo, info, kInfoSlots * sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_check(c, e, "LOBPCG iteration read-back");
  for (int s = 0; s < kInfoSlots; ++s) {
    if (hinfo[s] == 0) continue;
    const bool chol = s == 0 || s == 1 || s == 2 || s == 3 || s == 5 || s == 6 || s == 9 || s == 10 || s == 11 ||
                      s == 12;
    if (chol)
      return set_error(c, PW_ERR_NOT_POSDEF, "LOBPCG: Cholesky factorisation failed (status slot " + std::to_string(s) +
                                                 ", info " + std::to_string(hinfo[s]) + ")");
    return set_error(c, hinfo[s] == -1 ? PW_ERR_HIP : PW_ERR_ARG,
                     "LOBPCG: eigensolve failed (status slot " + std::to_string(s) + ", info " +
                         std::to_string(hinfo[s]) + ")");
  }
  return PW_OK;
}

}  // namespace

}  // namespace pw

using namespace pw;

extern "C" int pw_diag_lobpcg_dev(pw_ctx* c, int ik, int ispin, int nst, void* X, int64_t ldx, double tol,
                                  int niter_max, int nstates_conv, unsigned flags, double* evals, int* stats) {
  if (!c) return PW_ERR_ARG;
  hipSetDevice(c->device);
  if (c->N == 0) return set_error(c, PW_ERR_STATE, "pw_set_grid has not been called");
  if (ik < 0 || ik >= (int)c->kd.size() || c->kd[ik].ngw == 0)
    return set_error(c, PW_ERR_STATE, "k-point " + std::to_string(ik) + " is not set up");
  const int Nb = c->kd[ik].ngw;
  if (nst < 1 || !X || ldx < Nb || !evals || niter_max < 0 || nstates_conv < 0 || nstates_conv > nst || 3 * nst > Nb)
    return set_error(c, PW_ERR_ARG, "pw_diag_lobpcg_dev: bad arguments (need 1 <= Nst, 3 Nst <= Ngw)");
  if (c->capturing) return set_error(c, PW_ERR_STATE, "pw_diag_lobpcg_dev synchronises: not inside a graph capture");
  if (nstates_conv == 0) nstates_conv = nst;
  const int Nst = nst;
  const int n3 = 3 * Nst;
  // workspace: Q, HQ (Nb x 3 Nst), Tmp (Nb x 4 Nst), Sm, Gm, Cm ((3 Nst)^2), S2, S3 (Nst^2),
  // dw, E (3 Nst doubles), info words
  const size_t big = (size_t)Nb * 10 * Nst, small = 3 * (size_t)n3 * n3 + 2 * (size_t)Nst * Nst;
  const size_t need = (big + small) * sizeof(cplx<double>) + 2 * (size_t)n3 * sizeof(double) + kInfoSlots * sizeof(int);
  int rc = ensure(c, &c->d_lob, &c->lob_bytes, need);
  if (rc) return rc;
  Lobpcg L;
  L.c = c;
  L.ik = ik;
  L.ispin = ispin;
  L.Nb = Nb;
  L.Nst = Nst;
  L.prec = (flags & PW_MIXED) ? 1 : 0;
  L.hflags = PW_TERM_ALL | (flags & PW_MIXED);
  L.Q = reinterpret_cast<cplx<double>*>(c->d_lob);
  L.HQ = L.Q + (size_t)Nb * n3;
  L.Tmp = L.HQ + (size_t)Nb * n3;
  L.Sm = L.Tmp + (size_t)Nb * 4 * Nst;
  L.Gm = L.Sm + (size_t)n3 * n3;
  L.Cm = L.Gm + (size_t)n3 * n3;
  L.S2 = L.Cm + (size_t)n3 * n3;
  L.S3 = L.S2 + (size_t)Nst * Nst;
  L.dw = reinterpret_cast<double*>(L.S3 + (size_t)Nst * Nst);
  L.E = L.dw + n3;
  L.info = reinterpret_cast<int*>(L.E + n3);
  if (c->d_lob2) L.L = reinterpret_cast<cplx<double>*>(c->d_lob2);
  if (L.L && c->lob2_bytes < (size_t)Nb * 5 * Nst * sizeof(cplx<double>)) L.L = nullptr;
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, c->stream, L.info, kInfoSlots, 0);
  rc = hip_check(c, hipGetLastError(), "k_fill_i32");
  if (rc) return rc;

  const cplx<double>* Xin = reinterpret_cast<const cplx<double>*>(X);
  cplx<double>* x = L.q(0);
  cplx<double>* w = L.q(Nst);
  cplx<double>* p = L.q(2 * Nst);
  cplx<double>* hx = L.hq(0);
  cplx<double>* hw = L.hq(Nst);
  cplx<double>* hp = L.hq(2 * Nst);
  cplx<double>* t0 = L.tmp(0);
  cplx<double>* t1 = L.tmp(Nst);
  cplx<double>* t2 = L.tmp(2 * Nst);
  cplx<double>* t3 = L.tmp(3 * Nst);
  const int prec = L.prec;

  rc = copy_cols(c, Nb, Nst, Xin, ldx, x, Nb);
  if (!rc) rc = L.opH(x, hx, Nst);
  if (rc) return rc;
  const double SMALL = 10.0 * (c->lobpcg_lock_exp > 0 ? std::pow(10.0, -c->lobpcg_lock_exp)
                                                     : std::numeric_limits<double>::epsilon());
  std::vector<double> ev(Nst), ev_old(Nst, 0.0);
  std::vector<int> hinfo(kInfoSlots);
  bool have_rr = false;
  int it_done = 0;
  for (int it = 1; it <= niter_max; ++it) {
    it_done = it;
    rc = L.gemm(1, Nst, Nst, Nb, x, Nb, hx, Nb, L.S2, Nst, prec);  // X^H H X
    if (!rc) rc = herm_fix_launch(c, Nst, 1, L.S2, Nst, 0, 0);
    if (!rc && !have_rr) {
      // eigvals(Hermitian(X^H H X)); after a Rayleigh-Ritz step X holds its Ritz vectors and
      // the step's lowest Nst Ritz values (already in dw) are these eigenvalues to rounding
      rc = copy_cols(c, Nst, Nst, L.S2, Nst, L.S3, Nst);
      if (!rc) rc = heev_enqueue(c, Nst, L.S3, Nst, L.dw, 0, L.E, L.info + 7);
    }
    have_rr = false;
    if (!rc) rc = read_iteration(c, L.dw, Nst, L.info, ev.data(), hinfo.data());
    if (rc) return rc;
    // R = HX - X (X^H H X) into the W slot
    rc = L.copy(Nst, hx, w);
    if (!rc) rc = L.gemm(0, Nb, Nst, Nst, x, Nb, L.S2, Nst, w, Nb, prec, kMinusOne, kOne);
    if (rc) return rc;
    int nconv = 0;
    std::vector<int> ilock, iact;
    for (int i = 0; i < Nst; ++i) {
      const double d = std::fabs(ev[i] - ev_old[i]);
      if (d < tol) ++nconv;
      (d <= SMALL ? ilock : iact).push_back(i);
      ev_old[i] = ev[i];
    }
    if (nconv >= nstates_conv) break;
    rc = kprec_apply(c, ik, Nst, w, Nb, w, Nb);
    if (rc) return rc;
    if (!ilock.empty() && !iact.empty()) {
      rc = L.lock_step(ilock, iact);
      if (rc) return rc;
      continue;
    }
    rc = L.opH(w, hw, Nst);
    if (!rc) rc = L.chol_inv_apply(w, hw, t0, t1, Nst, 0);
    if (!rc) rc = L.copy(Nst, t0, w);
    if (!rc) rc = L.copy(Nst, t1, hw);
    if (rc) return rc;
    const int nq = it > 1 ? 3 * Nst : 2 * Nst;
    // T = Q^H H Q and G = Q^H Q, upper block triangle only; G in FP64 (it is factorised)
    for (int j = 0; j < nq / Nst && !rc; ++j) {
      const int m = (j + 1) * Nst;
      rc = L.gemm(1, m, Nst, Nb, L.Q, Nb, L.hq(j * Nst), Nb, L.Sm + (long)nq * j * Nst, nq, prec);
      if (!rc) rc = L.gemm(1, m, Nst, Nb, L.Q, Nb, L.q(j * Nst), Nb, L.Gm + (long)nq * j * Nst, nq, 0);
    }
    if (!rc) rc = L.rr_solve(nq, Nst, 2);
    if (rc) return rc;
    have_rr = true;
    // X = Q U, HX = HQ U; P = [W P] U[Nst:, :] (re-orthonormalised), HP likewise
    rc = L.gemm(0, Nb, Nst, nq, L.Q, Nb, L.Cm, nq, t0, Nb, prec);
    if (!rc) rc = L.gemm(0, Nb, Nst, nq, L.HQ, Nb, L.Cm, nq, t1, Nb, prec);
    if (rc) return rc;
    if (it > 1) {
      const cplx<double>* u23 = L.Cm + Nst;
      rc = L.gemm(0, Nb, Nst, 2 * Nst, w, Nb, u23, nq, t2, Nb, prec);
      if (!rc) rc = L.gemm(0, Nb, Nst, 2 * Nst, hw, Nb, u23, nq, t3, Nb, prec);
      if (!rc) rc = L.copy(Nst, t0, x);
      if (!rc) rc = L.copy(Nst, t1, hx);
      if (!rc) rc = L.chol_inv_apply(t2, t3, p, hp, Nst, 5);
    } else {
      rc = L.copy(Nst, w, p);
      if (!rc) rc = L.copy(Nst, hw, hp);
      if (!rc) rc = L.copy(Nst, t0, x);
      if (!rc) rc = L.copy(Nst, t1, hx);
    }
    if (rc) return rc;
  }
  // final Rayleigh-Ritz on X: S = (X^H H X + (X^H H X)^H) / 2, X <- X eigvecs(S)
  rc = L.gemm(1, Nst, Nst, Nb, x, Nb, hx, Nb, L.S2, Nst, prec);
  if (!rc) rc = herm_fix_launch(c, Nst, 1, L.S2, Nst, 0, 2);
  if (!rc) rc = heev_enqueue(c, Nst, L.S2, Nst, L.dw, Nst, L.E, L.info + 8);
  if (!rc) rc = L.gemm(0, Nb, Nst, Nst, x, Nb, L.S2, Nst, reinterpret_cast<cplx<double>*>(X), ldx, prec);
  if (!rc) rc = read_iteration(c, L.dw, Nst, L.info, evals, hinfo.data());
  if (rc) return rc;
  if (stats) {
    stats[0] = it_done;
    stats[1] = (int)std::min<long>(L.ncols, 2147483647L);
    stats[2] = L.locks;
  }
  return PW_OK;
}

// ------------------------------------------------------------------ diag_davidson!
//
// pw_diag_davidson_dev: diag_davidson! (diag_davidson.jl:50-191) for one k-point and spin on
// the device.  Q = [X R] and HQ = [HX HR] are the 2 Nst search space and its H-image; the
// reduced matrices Hred, Sred (2 Nst square) are assembled, factorised and solved in HBM
// (the same generalised solve as the Rayleigh-Ritz step of LOBPCG: G = U^H U, U^-H T U^-1),
// and the host reads the Nst lowest Ritz values once per iteration for the |d lambda| test.
// The step order is the reference's: R = X diag(evals) - HX, each column scaled by 1/|R_j|
// when |R_j| > eps, R = Kprec(R), HR = H R, Hred = [X^H HX or diag(evals), X^H HR; ., R^H HR],
// Sred = [I, X^H R; ., R^H R], both made Hermitian as 0.5 (A + A^H), eigen(Hred, Sred),
// X = Q y, HX = HQ y for the Nst lowest.

namespace pw {
namespace {

constexpr int DAV_NT = 256;

// dw[j] = Re S[j, j] (the initial Rayleigh quotients; S has leading dimension lds)
__global__ void k_diag_re(const cplx<double>* __restrict__ S, long lds, int n, double* __restrict__ dw) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) dw[j] = S[(long)j * lds + j].x;
}

// One workgroup per column j: R_j = ev_j X_j - HX_j, then R_j /= |R_j| when eps < |R_j|
// (diag_davidson.jl:82-89 and :102-112).
__global__ void __launch_bounds__(DAV_NT) k_dav_resid(const cplx<double>* __restrict__ X, const cplx<double>* __restrict__ HX,
                                                      const double* __restrict__ ev, cplx<double>* __restrict__ R,
                                                      long ld, int nb) {
  __shared__ double red[DAV_NT / 64];
  const int j = blockIdx.x;
  const double e = ev[j];
  const cplx<double>* x = X + (long)j * ld;
  const cplx<double>* h = HX + (long)j * ld;
  cplx<double>* r = R + (long)j * ld;
  double s = 0.0;
  for (int g = threadIdx.x; g < nb; g += DAV_NT) {
    const cplx<double> v = mk(e * x[g].x - h[g].x, e * x[g].y - h[g].y);
    r[g] = v;
    s += v.x * v.x + v.y * v.y;
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  double tot = 0.0;
  for (int q = 0; q < DAV_NT / 64; ++q) tot += red[q];
  const double res = sqrt(tot);
  if (!(2.220446049250313e-16 < res)) return;
  const double f = 1.0 / res;
  for (int g = threadIdx.x; g < nb; g += DAV_NT) r[g] = scal(f, r[g]);
}

// Hred and Sred (order 2 n, leading dimension 2 n) from the products already in their
// upper block column [., n:2n) (and, in the first iteration, Hred's leading n x n block):
// the leading blocks set (Hred: first ? X^H HX : diag(ev); Sred: I), the lower triangle
// mirrored from the upper one and every entry averaged as 0.5 (A + A^H).  One thread per
// pair i <= j writes both entries.
__global__ void k_dav_assemble(cplx<double>* __restrict__ H, cplx<double>* __restrict__ S, int n,
                               const double* __restrict__ ev, int first) {
  const int m = 2 * n;
  const long tot = (long)m * m;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < tot; t += (long)gridDim.x * blockDim.x) {
    const int i = (int)(t % m), j = (int)(t / m);
    if (i > j) continue;
    cplx<double> h, s;
    if (j < n) {  // leading block
      if (first) {
        const cplx<double> u = H[(long)j * m + i], l = H[(long)i * m + j];
        h = mk(0.5 * (u.x + l.x), 0.5 * (u.y - l.y));
      } else {
        h = mk(i == j ? ev[i] : 0.0, 0.0);
      }
      s = mk(i == j ? 1.0 : 0.0, 0.0);
    } else if (i < n) {  // upper-right block: X^H HR, X^H R (the mirror averages to itself)
      h = H[(long)j * m + i];
      s = S[(long)j * m + i];
    } else {  // trailing block: R^H HR, R^H R
      const cplx<double> hu = H[(long)j * m + i], hl = H[(long)i * m + j];
      const cplx<double> su = S[(long)j * m + i], sl = S[(long)i * m + j];
      h = mk(0.5 * (hu.x + hl.x), 0.5 * (hu.y - hl.y));
      s = mk(0.5 * (su.x + sl.x), 0.5 * (su.y - sl.y));
    }
    if (i == j) {
      h.y = 0.0;
      s.y = 0.0;
    }
    H[(long)j * m + i] = h;
    S[(long)j * m + i] = s;
    H[(long)i * m + j] = mk(h.x, -h.y);
    S[(long)i * m + j] = mk(s.x, -s.y);
  }
}

}  // namespace
}  // namespace pw

extern "C" int pw_diag_davidson_dev(pw_ctx* c, int ik, int ispin, int nst, void* X, int64_t ldx, double tol,
                                    int niter_max, int nstates_conv, unsigned flags, double* evals, int* stats) {
  if (!c) return PW_ERR_ARG;
  hipSetDevice(c->device);
  if (c->N == 0) return set_error(c, PW_ERR_STATE, "pw_set_grid has not been called");
  if (ik < 0 || ik >= (int)c->kd.size() || c->kd[ik].ngw == 0)
    return set_error(c, PW_ERR_STATE, "k-point " + std::to_string(ik) + " is not set up");
  const int Nb = c->kd[ik].ngw;
  if (nst < 1 || !X || ldx < Nb || !evals || niter_max < 0 || nstates_conv < 0 || nstates_conv > nst || 2 * nst > Nb)
    return set_error(c, PW_ERR_ARG, "pw_diag_davidson_dev: bad arguments (need 1 <= Nst, 2 Nst <= Ngw)");
  if (c->capturing) return set_error(c, PW_ERR_STATE, "pw_diag_davidson_dev synchronises: not inside a graph capture");
  if (nstates_conv == 0) nstates_conv = nst;
  const int Nst = nst, n2 = 2 * nst;
  // workspace: Q, HQ, Tmp (Nb x 2 Nst each), Hred, Sred, Cm (n2^2), S2 (Nst^2), dw, E (n2
  // doubles), status words
  const size_t big = (size_t)Nb * 6 * Nst, small = 3 * (size_t)n2 * n2 + (size_t)Nst * Nst;
  const size_t need = (big + small) * sizeof(cplx<double>) + 2 * (size_t)n2 * sizeof(double) + kInfoSlots * sizeof(int);
  int rc = ensure(c, &c->d_lob, &c->lob_bytes, need);
  if (rc) return rc;
  Lobpcg L;
  L.c = c;
  L.ik = ik;
  L.ispin = ispin;
  L.Nb = Nb;
  L.Nst = Nst;
  L.prec = (flags & PW_MIXED) ? 1 : 0;
  L.hflags = PW_TERM_ALL | (flags & PW_MIXED);
  L.Q = reinterpret_cast<cplx<double>*>(c->d_lob);
  L.HQ = L.Q + (size_t)Nb * n2;
  L.Tmp = L.HQ + (size_t)Nb * n2;
  L.Sm = L.Tmp + (size_t)Nb * n2;   // Hred
  L.Gm = L.Sm + (size_t)n2 * n2;    // Sred
  L.Cm = L.Gm + (size_t)n2 * n2;
  L.S2 = L.Cm + (size_t)n2 * n2;
  L.S3 = nullptr;
  L.dw = reinterpret_cast<double*>(L.S2 + (size_t)Nst * Nst);
  L.E = L.dw + n2;
  L.info = reinterpret_cast<int*>(L.E + n2);
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, c->stream, L.info, kInfoSlots, 0);
  rc = hip_check(c, hipGetLastError(), "k_fill_i32");
  if (rc) return rc;
  cplx<double>* x = L.q(0);
  cplx<double>* r = L.q(Nst);
  cplx<double>* hx = L.hq(0);
  cplx<double>* hr = L.hq(Nst);
  const int prec = L.prec;
  // HX, evals = Re diag(X^H HX), R = X diag(evals) - HX (scaled)
  rc = copy_cols(c, Nb, Nst, reinterpret_cast<const cplx<double>*>(X), ldx, x, Nb);
  if (!rc) rc = L.opH(x, hx, Nst);
  if (!rc) rc = L.gemm(1, Nst, Nst, Nb, x, Nb, hx, Nb, L.S2, Nst, prec);
  if (rc) return rc;
  hipLaunchKernelGGL(k_diag_re, dim3((Nst + 255) / 256), dim3(256), 0, c->stream, L.S2, (long)Nst, Nst, L.dw);
  hipLaunchKernelGGL(k_dav_resid, dim3(Nst), dim3(DAV_NT), 0, c->stream, x, hx, L.dw, r, (long)Nb, Nb);
  rc = hip_check(c, hipGetLastError(), "Davidson residual");
  std::vector<double> ev(Nst), ev_old(Nst);
  std::vector<int> hinfo(kInfoSlots);
  if (!rc) rc = read_iteration(c, L.dw, Nst, L.info, ev_old.data(), hinfo.data());
  if (rc) return rc;
  ev = ev_old;
  int it_done = 0;
  const int nblk = (int)std::min<long>(((long)n2 * n2 + 255) / 256, 1024);
  for (int it = 1; it <= niter_max; ++it) {
    it_done = it;
    rc = kprec_apply(c, ik, Nst, r, Nb, r, Nb);
    if (!rc) rc = L.opH(r, hr, Nst);
    // upper block column of Hred (X^H HR; R^H HR) and Sred (X^H R; R^H R); Sred in FP64
    if (!rc) rc = L.gemm(1, n2, Nst, Nb, L.Q, Nb, hr, Nb, L.Sm + (long)n2 * Nst, n2, prec);
    if (!rc) rc = L.gemm(1, n2, Nst, Nb, L.Q, Nb, r, Nb, L.Gm + (long)n2 * Nst, n2, 0);
    if (!rc && it == 1) rc = L.gemm(1, Nst, Nst, Nb, x, Nb, hx, Nb, L.Sm, n2, prec);
    if (rc) return rc;
    hipLaunchKernelGGL(k_dav_assemble, dim3(nblk), dim3(256), 0, c->stream, L.Sm, L.Gm, Nst, L.dw, it == 1 ? 1 : 0);
    rc = hip_check(c, hipGetLastError(), "k_dav_assemble");
    if (!rc) rc = L.rr_solve(n2, Nst, 2);
    // X = Q y, HX = HQ y
    if (!rc) rc = L.gemm(0, Nb, Nst, n2, L.Q, Nb, L.Cm, n2, L.tmp(0), Nb, prec);
    if (!rc) rc = L.gemm(0, Nb, Nst, n2, L.HQ, Nb, L.Cm, n2, L.tmp(Nst), Nb, prec);
    if (!rc) rc = L.copy(Nst, L.tmp(0), x);
    if (!rc) rc = L.copy(Nst, L.tmp(Nst), hx);
    if (rc) return rc;
    hipLaunchKernelGGL(k_dav_resid, dim3(Nst), dim3(DAV_NT), 0, c->stream, x, hx, L.dw, r, (long)Nb, Nb);
    rc = hip_check(c, hipGetLastError(), "Davidson residual");
    if (!rc) rc = read_iteration(c, L.dw, Nst, L.info, ev.data(), hinfo.data());
    if (rc) return rc;
    int nconv = 0;
    for (int i = 0; i < Nst; ++i)
      if (std::fabs(ev[i] - ev_old[i]) < tol) ++nconv;
    if (nconv >= nstates_conv) break;
    ev_old = ev;
  }
  rc = copy_cols(c, Nb, Nst, x, Nb, reinterpret_cast<cplx<double>*>(X), ldx);
  if (!rc) {
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) rc = hip_check(c, e, "Davidson final copy");
  }
  if (rc) return rc;
  std::copy(ev.begin(), ev.end(), evals);
  if (stats) {
    stats[0] = it_done;
    stats[1] = (int)std::min<long>(L.ncols, 2147483647L);
    stats[2] = 0;
  }
  return PW_OK;
}

// ------------------------------------------------------------------ all k-points at once
//
// pw_diag_lobpcg_batch_dev runs the iteration above for many k-points in lockstep.  Every
// operation of an iteration is one launch for all k-points still iterating: op_H of all
// their bands (pw_op_H_batch_dev; for grids whose box fits LDS one fused workgroup per
// (k-point, band)), the Gram products and column updates (zgemm_batch), copies and Kprec
// (elem_batch), and the small dense problems as strided stacks, one workgroup per matrix
// (pw_small.hip) or rocSOLVER's batched routines beyond order 64.  An iteration is two
// phases with one synchronisation between them:
//   check(S):  S_X = X^H H X (Hermitian), R = HX - X S_X into the W block, eigvals(S_X);
//              the host reads the eigenvalues and statuses and decides, per k-point, as
//              diag_LOBPCG.jl:88-97,125 does: converged (leaves the batch), lock branch
//              (:125-164, done per k-point with the single-k code above), or a plain step;
//   step(S):   :99-203 for the k-points S that take the plain step.
// A phase that has run once eagerly for the same k-points and shapes is captured into a
// hipGraph the next time and replayed from then on (orders 3 Nst <= 64, whose dense
// problems need no library workspace); the graphs persist in the context across calls (an
// SCF calls the solver once per iteration with the same k-points) and are discarded when
// any device buffer they address is reallocated.

namespace pw {

void free_lob_graphs(pw_ctx* c) {
  if (c->lob_graphs.empty()) return;
  hipStreamSynchronize(c->stream);
  for (auto& g : c->lob_graphs) {
    if (g.exec) {
      hipGraphExecDestroy(reinterpret_cast<hipGraphExec_t>(g.exec));
      c->graphs_live--;
    }
    if (g.graph) hipGraphDestroy(reinterpret_cast<hipGraph_t>(g.graph));
  }
  c->lob_graphs.clear();
  if (c->graphs_live == 0) {
    for (void* p : c->retired) hipFree(p);
    c->retired.clear();
  }
}

namespace {

constexpr int kMaxLobGraphs = 64;

GBatch gb(const cplx<double>* A, long lda, const cplx<double>* B, long ldb, cplx<double>* C, long ldc, int m, int n,
          int k) {
  GBatch g{};
  g.A = A;
  g.B = B;
  g.C = C;
  g.lda = lda;
  g.ldb = ldb;
  g.ldc = ldc;
  g.m = m;
  g.n = n;
  g.k = k;
  return g;
}

EBatch eb(const cplx<double>* in, long ldin, cplx<double>* out, long ldout, int m, int n, const double* gk2 = nullptr) {
  EBatch e{};
  e.in = in;
  e.out = out;
  e.gk2 = gk2;
  e.ldin = ldin;
  e.ldout = ldout;
  e.m = m;
  e.n = n;
  return e;
}

struct BatchLobpcg {
  pw_ctx* c;
  int nk, Nst, ispin;
  unsigned hflags;
  bool use_graph;
  std::vector<int> iks, rows;
  std::vector<cplx<double>*> blk;   // k-point i: [Q (3 Nst) | HQ (3 Nst) | Tmp (4 Nst)], rows[i] rows
  // stacked small matrices, by position in the set of the phase
  cplx<double> *sX, *sX2, *sC, *sT, *sG, *sM;
  double *dwx, *dw;   // eigvals of S_X (Nst per position), Rayleigh-Ritz values (3 Nst per position)
  int* info;          // 8 nk status words (layout below)
  // single-k workspace of the lock branch
  cplx<double> *lSm, *lGm, *lCm, *lS2, *lS3;
  double *ldw, *lE;
  int* linfo;
  std::vector<int> key0;

  cplx<double>* q(int i, int j) { return blk[i] + (long)j * rows[i]; }
  cplx<double>* hq(int i, int j) { return blk[i] + (long)(3 * Nst + j) * rows[i]; }
  cplx<double>* tmp(int i, int j) { return blk[i] + (long)(6 * Nst + j) * rows[i]; }
  cplx<double>* sx(int p) { return sX + (long)p * Nst * Nst; }
  cplx<double>* sc(int p) { return sC + (long)p * Nst * Nst; }
  int* inf_check() { return info; }          // nk: eigvals(S_X) and the final Rayleigh-Ritz
  int* inf_w() { return info + nk; }         // 2 nk: cholesky of W^H W
  int* inf_g() { return info + 3 * nk; }     // 2 nk: cholesky of G
  int* inf_rr() { return info + 5 * nk; }    // nk: Rayleigh-Ritz eigenproblem
  int* inf_p() { return info + 6 * nk; }     // 2 nk: cholesky of P^H P

  template <class F>
  int gemm(const std::vector<int>& S, int ta, cplx<double> alpha, cplx<double> beta, int per, F f) {
    std::vector<GBatch> g;
    g.reserve(S.size() * per);
    for (int r = 0; r < per; ++r)
      for (size_t p = 0; p < S.size(); ++p) g.push_back(f(r, (int)p, S[p]));
    return zgemm_batch(c, ta, g, alpha, beta);
  }
  template <class F>
  int elem(const std::vector<int>& S, int per, F f) {
    std::vector<EBatch> e;
    e.reserve(S.size() * per);
    for (int r = 0; r < per; ++r)
      for (size_t p = 0; p < S.size(); ++p) e.push_back(f(r, (int)p, S[p]));
    return elem_batch(c, e);
  }
  int chol(int n, int nb, cplx<double>* A, int* inf) {
    return pw_chol_inv_batch_async_dev(c, n, nb, A, n, (int64_t)n * n, inf);
  }
  int heev(int n, int nb, cplx<double>* A, double* w, int nvec, int* inf) {
    if (c->eig_path == 0 && n <= small_dense_max_n()) return heev_small_batch(c, n, nb, A, n, (long)n * n, w, nvec, inf);
    return pw_heevd_batch_async_dev(c, n, nb, A, n, (int64_t)n * n, w, inf);
  }

  // S_X = X^H H X, R = HX - X S_X (W block), eigvals(S_X) into dwx
  int check(const std::vector<int>& S) {
    const int ns = (int)S.size(), N = Nst;
    int rc = gemm(S, 1, kOne, kZero, 1, [&](int, int p, int i) {
      return gb(q(i, 0), rows[i], hq(i, 0), rows[i], sx(p), N, N, N, rows[i]);
    });
    if (!rc) rc = herm_fix_launch(c, N, ns, sX, N, (int64_t)N * N, 0);
    if (!rc) rc = elem(S, 1, [&](int, int, int i) { return eb(hq(i, 0), rows[i], q(i, N), rows[i], rows[i], N); });
    if (!rc) rc = gemm(S, 0, kMinusOne, kOne, 1, [&](int, int p, int i) {
      return gb(q(i, 0), rows[i], sx(p), N, q(i, N), rows[i], rows[i], N, N);
    });
    if (!rc) rc = copy_cols(c, N * N * ns, 1, sX, (long)N * N * ns, sX2, (long)N * N * ns);
    if (!rc) rc = heev(N, ns, sX2, dwx, 0, inf_check());
    return rc;
  }

  // diag_LOBPCG.jl:99-203 for the k-points S (W holds R); nq = 2 Nst in the first
  // iteration, 3 Nst after it
  int step(const std::vector<int>& S, int nq) {
    const int ns = (int)S.size(), N = Nst;
    std::vector<int> ik(ns);
    std::vector<void*> src(ns), dst(ns);
    std::vector<int64_t> ld(ns);
    for (int p = 0; p < ns; ++p) {
      const int i = S[p];
      ik[p] = iks[i];
      src[p] = q(i, N);
      dst[p] = hq(i, N);
      ld[p] = rows[i];
    }
    auto st = [&](int p) { return sT + (long)p * nq * nq; };
    auto sg = [&](int p) { return sG + (long)p * nq * nq; };
    auto sm = [&](int p) { return sM + (long)p * nq * nq; };
    // W = Kprec(R), HW
    int rc = elem(S, 1, [&](int, int, int i) {
      return eb(q(i, N), rows[i], q(i, N), rows[i], rows[i], N, c->kd[iks[i]].d_gk2);
    });
    if (!rc) rc = pw_op_H_batch_dev(c, ns, ik.data(), ispin, N, src.data(), ld.data(), dst.data(), ld.data(), hflags);
    // W, HW <- W U^-1, HW U^-1 with U = chol(W^H W)
    if (!rc) rc = gemm(S, 1, kOne, kZero, 1, [&](int, int p, int i) {
      return gb(q(i, N), rows[i], q(i, N), rows[i], sc(p), N, N, N, rows[i]);
    });
    if (!rc) rc = chol(N, ns, sC, inf_w());
    if (!rc) rc = gemm(S, 0, kOne, kZero, 2, [&](int r, int p, int i) {
      return gb(r ? hq(i, N) : q(i, N), rows[i], sc(p), N, tmp(i, r * N), rows[i], rows[i], N, N);
    });
    if (!rc) rc = elem(S, 2, [&](int r, int, int i) {
      return eb(tmp(i, r * N), rows[i], r ? hq(i, N) : q(i, N), rows[i], rows[i], N);
    });
    // T = Q^H H Q, G = Q^H Q
    if (!rc) rc = gemm(S, 1, kOne, kZero, 2, [&](int r, int p, int i) {
      return gb(q(i, 0), rows[i], r ? q(i, 0) : hq(i, 0), rows[i], r ? sg(p) : st(p), nq, nq, nq, rows[i]);
    });
    // eigen(Hermitian(T), Hermitian(G)): G = U^H U, U^-H T U^-1 y = l y, coefficients U^-1 y
    if (!rc) rc = herm_fix_launch(c, nq, ns, sT, nq, (int64_t)nq * nq, 0);
    if (!rc) rc = chol(nq, ns, sG, inf_g());
    if (!rc) rc = gemm(S, 0, kOne, kZero, 1, [&](int, int p, int) { return gb(st(p), nq, sg(p), nq, sm(p), nq, nq, nq, nq); });
    if (!rc) rc = gemm(S, 1, kOne, kZero, 1, [&](int, int p, int) { return gb(sg(p), nq, sm(p), nq, st(p), nq, nq, nq, nq); });
    if (!rc) rc = heev(nq, ns, sT, dw, N, inf_rr());
    if (!rc) rc = gemm(S, 0, kOne, kZero, 1, [&](int, int p, int) { return gb(sg(p), nq, st(p), nq, sm(p), nq, nq, N, nq); });
    // X U, HQ U into Tmp
    if (!rc) rc = gemm(S, 0, kOne, kZero, 2, [&](int r, int p, int i) {
      return gb(r ? hq(i, 0) : q(i, 0), rows[i], sm(p), nq, tmp(i, r * N), rows[i], rows[i], N, nq);
    });
    if (rc) return rc;
    if (nq == 3 * N) {
      // P = [W P] U[Nst:, :], HP likewise, then re-orthonormalised
      rc = gemm(S, 0, kOne, kZero, 2, [&](int r, int p, int i) {
        return gb(r ? hq(i, N) : q(i, N), rows[i], sm(p) + N, nq, tmp(i, (2 + r) * N), rows[i], rows[i], N, 2 * N);
      });
      if (!rc) rc = elem(S, 2, [&](int r, int, int i) {
        return eb(tmp(i, r * N), rows[i], r ? hq(i, 0) : q(i, 0), rows[i], rows[i], N);
      });
      if (!rc) rc = gemm(S, 1, kOne, kZero, 1, [&](int, int p, int i) {
        return gb(tmp(i, 2 * N), rows[i], tmp(i, 2 * N), rows[i], sc(p), N, N, N, rows[i]);
      });
      if (!rc) rc = chol(N, ns, sC, inf_p());
      if (!rc) rc = gemm(S, 0, kOne, kZero, 2, [&](int r, int p, int i) {
        return gb(tmp(i, (2 + r) * N), rows[i], sc(p), N, r ? hq(i, 2 * N) : q(i, 2 * N), rows[i], rows[i], N, N);
      });
    } else {
      // P = W, HP = HW; X, HX from Tmp
      rc = elem(S, 4, [&](int r, int, int i) {
        if (r == 0) return eb(q(i, N), rows[i], q(i, 2 * N), rows[i], rows[i], N);
        if (r == 1) return eb(hq(i, N), rows[i], hq(i, 2 * N), rows[i], rows[i], N);
        if (r == 2) return eb(tmp(i, 0), rows[i], q(i, 0), rows[i], rows[i], N);
        return eb(tmp(i, N), rows[i], hq(i, 0), rows[i], rows[i], N);
      });
    }
    return rc;
  }

  // Runs one phase: eagerly the first time its key is seen, captured into a graph the
  // second time, replayed from the graph after that.
  template <class F>
  int run(int kind, int nq, const std::vector<int>& S, F enqueue) {
    if (!use_graph) return enqueue();
    std::vector<int> key = key0;
    key.push_back(kind);
    key.push_back(nq);
    key.push_back((int)S.size());
    key.insert(key.end(), S.begin(), S.end());
    auto& G = c->lob_graphs;
    int idx = -1;
    for (int j = 0; j < (int)G.size(); ++j)
      if (G[j].gen == c->graph_gen && G[j].key == key) {
        idx = j;
        break;
      }
    if (idx < 0) {
      if ((int)G.size() >= kMaxLobGraphs) {
        int old = 0;
        for (int j = 1; j < (int)G.size(); ++j)
          if (G[j].stamp < G[old].stamp) old = j;
        hipStreamSynchronize(c->stream);
        if (G[old].exec) {
          hipGraphExecDestroy(reinterpret_cast<hipGraphExec_t>(G[old].exec));
          c->graphs_live--;
        }
        if (G[old].graph) hipGraphDestroy(reinterpret_cast<hipGraph_t>(G[old].graph));
        G.erase(G.begin() + old);
      }
      LobGraph g;
      g.key = std::move(key);
      g.gen = c->graph_gen;
      g.stamp = ++c->lob_stamp;
      G.push_back(std::move(g));
      return enqueue();
    }
    G[idx].stamp = ++c->lob_stamp;
    if (G[idx].failed) return enqueue();
    if (!G[idx].exec) {
      hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        G[idx].failed = true;
        return enqueue();
      }
      c->capturing = true;
      const int rc = enqueue();
      hipGraph_t gr = nullptr;
      e = hipStreamEndCapture(c->stream, &gr);
      c->capturing = false;
      hipGraphExec_t x = nullptr;
      if (rc == PW_OK && e == hipSuccess && gr) e = hipGraphInstantiate(&x, gr, nullptr, nullptr, 0);
      if (rc != PW_OK || e != hipSuccess || !x) {
        if (gr) hipGraphDestroy(gr);
        (void)hipGetLastError();
        c->err.clear();
        G[idx].failed = true;
        return enqueue();
      }
      G[idx].graph = gr;
      G[idx].exec = x;
      c->graphs_live++;
    }
    return hip_check(c, hipGraphLaunch(reinterpret_cast<hipGraphExec_t>(G[idx].exec), c->stream), "LOBPCG step graph");
  }

  // Statuses of the phases since the last read: 0 everywhere, or the first failure.
  int check_info(const int* h) {
    for (int s = 0; s < 8 * nk; ++s) {
      if (h[s] == 0) continue;
      const int region = s < nk ? 0 : s < 3 * nk ? 1 : s < 5 * nk ? 2 : s < 6 * nk ? 3 : 4;
      if (region == 0 || region == 3)
        return set_error(c, PW_ERR_ARG, "LOBPCG batch: eigensolve failed (" + std::to_string(h[s]) + ")");
      return set_error(c, PW_ERR_NOT_POSDEF, "LOBPCG batch: Cholesky factorisation failed (" +
                                                 std::string(region == 1 ? "W^H W" : region == 2 ? "G" : "P^H P") +
                                                 ", info " + std::to_string(h[s]) + ")");
    }
    return PW_OK;
  }
};

}  // namespace

}  // namespace pw

extern "C" int pw_diag_lobpcg_batch_dev(pw_ctx* c, int nk, const int* iks, int ispin, int nst, void* const* X,
                                        const int64_t* ldx, double tol, int niter_max, int nstates_conv,
                                        unsigned flags, double* evals, int* stats) {
  if (!c) return PW_ERR_ARG;
  hipSetDevice(c->device);
  if (c->N == 0) return set_error(c, PW_ERR_STATE, "pw_set_grid has not been called");
  if (nk < 1 || !iks || !X || !ldx || !evals || nst < 1 || niter_max < 0 || nstates_conv < 0 || nstates_conv > nst)
    return set_error(c, PW_ERR_ARG, "pw_diag_lobpcg_batch_dev: bad arguments");
  if (c->capturing) return set_error(c, PW_ERR_STATE, "pw_diag_lobpcg_batch_dev synchronises: not inside a capture");
  const int N = nst, N3 = 3 * nst;
  BatchLobpcg B;
  B.c = c;
  B.nk = nk;
  B.Nst = N;
  B.ispin = ispin;
  B.hflags = PW_TERM_ALL | (flags & PW_MIXED);
  B.iks.assign(iks, iks + nk);
  B.rows.resize(nk);
  size_t big = 0;
  for (int i = 0; i < nk; ++i) {
    if (iks[i] < 0 || iks[i] >= (int)c->kd.size() || c->kd[iks[i]].ngw == 0)
      return set_error(c, PW_ERR_STATE, "k-point " + std::to_string(iks[i]) + " is not set up");
    B.rows[i] = c->kd[iks[i]].ngw;
    if (3 * N > B.rows[i] || ldx[i] < B.rows[i] || !X[i])
      return set_error(c, PW_ERR_ARG, "pw_diag_lobpcg_batch_dev: need 3 Nst <= Ngw and ldx >= Ngw for every k-point");
    big += (size_t)B.rows[i] * 10 * N;
  }
  if (nstates_conv == 0) nstates_conv = N;
  if (nk > 65535) return set_error(c, PW_ERR_ARG, "pw_diag_lobpcg_batch_dev: more than 65535 k-points");
  // workspace: the k-points' blocks; the small stacks, the lock branch's single-k matrices,
  // eigenvalues and status words
  int rc = ensure(c, &c->d_lobb, &c->lobb_bytes, big * sizeof(cplx<double>));
  if (rc) return rc;
  const size_t s1 = (size_t)N * N, s9 = (size_t)N3 * N3;
  const size_t ncplx = (size_t)nk * (3 * s1 + 3 * s9) + 3 * s9 + 2 * s1;
  const size_t ndbl = (size_t)nk * (N + N3) + 2 * (size_t)N3;
  const size_t need = ncplx * sizeof(cplx<double>) + ndbl * sizeof(double) + ((size_t)8 * nk + kInfoSlots) * sizeof(int);
  rc = ensure(c, &c->d_lobs, &c->lobs_bytes, need);
  if (rc) return rc;
  {
    cplx<double>* p = reinterpret_cast<cplx<double>*>(c->d_lobb);
    B.blk.resize(nk);
    for (int i = 0; i < nk; ++i) {
      B.blk[i] = p;
      p += (size_t)B.rows[i] * 10 * N;
    }
    cplx<double>* s = reinterpret_cast<cplx<double>*>(c->d_lobs);
    B.sX = s;
    B.sX2 = B.sX + nk * s1;
    B.sC = B.sX2 + nk * s1;
    B.sT = B.sC + nk * s1;
    B.sG = B.sT + nk * s9;
    B.sM = B.sG + nk * s9;
    B.lSm = B.sM + nk * s9;
    B.lGm = B.lSm + s9;
    B.lCm = B.lGm + s9;
    B.lS2 = B.lCm + s9;
    B.lS3 = B.lS2 + s1;
    double* d = reinterpret_cast<double*>(B.lS3 + s1);
    B.dwx = d;
    B.dw = B.dwx + (size_t)nk * N;
    B.ldw = B.dw + (size_t)nk * N3;
    B.lE = B.ldw + N3;
    B.info = reinterpret_cast<int*>(B.lE + N3);
    B.linfo = B.info + 8 * nk;
  }
  B.use_graph = c->lob_graph && c->eig_path == 0 && N3 <= small_dense_max_n();
  B.key0 = {N, ispin, (int)B.hflags, c->fused_batch, nk};
  B.key0.insert(B.key0.end(), iks, iks + nk);
  // stale captured steps (a buffer they address was reallocated) are dropped here
  {
    auto& G = c->lob_graphs;
    bool stale = false;
    for (auto& g : G) stale = stale || g.gen != c->graph_gen;
    if (stale) {
      hipStreamSynchronize(c->stream);
      for (auto& g : G)
        if (g.gen != c->graph_gen) {
          if (g.exec) {
            hipGraphExecDestroy(reinterpret_cast<hipGraphExec_t>(g.exec));
            c->graphs_live--;
          }
          if (g.graph) hipGraphDestroy(reinterpret_cast<hipGraph_t>(g.graph));
          g.exec = g.graph = nullptr;
        }
      G.erase(std::remove_if(G.begin(), G.end(), [&](const LobGraph& g) { return g.gen != c->graph_gen; }), G.end());
      if (c->graphs_live == 0) {
        for (void* p : c->retired) hipFree(p);
        c->retired.clear();
      }
    }
  }
  hipLaunchKernelGGL(k_fill_i32, dim3((8 * nk + kInfoSlots + 255) / 256), dim3(256), 0, c->stream, B.info,
                     8 * nk + kInfoSlots, 0);
  rc = hip_check(c, hipGetLastError(), "k_fill_i32");
  if (rc) return rc;

  std::vector<int> all(nk);
  for (int i = 0; i < nk; ++i) all[i] = i;
  std::vector<long> ncols(nk, N);
  std::vector<int> iters(nk, 0), nlocks(nk, 0);
  std::vector<double> ev_old((size_t)nk * N, 0.0), ev((size_t)nk * N);
  std::vector<int> hinfo(8 * nk + kInfoSlots);
  const double SMALL = 10.0 * (c->lobpcg_lock_exp > 0 ? std::pow(10.0, -c->lobpcg_lock_exp)
                                                     : std::numeric_limits<double>::epsilon());
  // X into the blocks, HX
  rc = B.elem(all, 1, [&](int, int, int i) {
    return eb(reinterpret_cast<const cplx<double>*>(X[i]), ldx[i], B.q(i, 0), B.rows[i], B.rows[i], N);
  });
  if (!rc) {
    std::vector<void*> src(nk), dst(nk);
    std::vector<int64_t> ld(nk);
    for (int i = 0; i < nk; ++i) {
      src[i] = B.q(i, 0);
      dst[i] = B.hq(i, 0);
      ld[i] = B.rows[i];
    }
    rc = pw_op_H_batch_dev(c, nk, iks, ispin, N, src.data(), ld.data(), dst.data(), ld.data(), B.hflags);
  }
  if (rc) return rc;
  auto readback = [&](int ns) -> int {
    hipError_t e = hipMemcpyAsync(ev.data(), B.dwx, (size_t)ns * N * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(hinfo.data(), B.info, hinfo.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_check(c, e, "LOBPCG batch read-back");
    int r = B.check_info(hinfo.data());
    if (r) return r;
    for (int s = 0; s < kInfoSlots; ++s)
      if (hinfo[8 * nk + s] != 0)
        return set_error(c, hinfo[8 * nk + s] == -1 ? PW_ERR_HIP : PW_ERR_NOT_POSDEF,
                         "LOBPCG batch: lock branch factorisation or eigensolve failed (slot " + std::to_string(s) +
                             ")");
    return PW_OK;
  };
  std::vector<int> active = all;
  for (int it = 1; it <= niter_max && !active.empty(); ++it) {
    rc = B.run(0, 0, active, [&]() { return B.check(active); });
    if (!rc) rc = readback((int)active.size());
    if (rc) return rc;
    std::vector<int> keep, locked;
    std::vector<std::vector<int>> lk_lock, lk_act;
    for (size_t p = 0; p < active.size(); ++p) {
      const int i = active[p];
      iters[i] = it;
      int nconv = 0;
      std::vector<int> il, ia;
      for (int b = 0; b < N; ++b) {
        const double v = ev[p * N + b];
        const double dl = std::fabs(v - ev_old[(size_t)i * N + b]);
        if (dl < tol) ++nconv;
        (dl <= SMALL ? il : ia).push_back(b);
        ev_old[(size_t)i * N + b] = v;
      }
      if (nconv >= nstates_conv) continue;
      if (!il.empty() && !ia.empty()) {
        locked.push_back(i);
        lk_lock.push_back(il);
        lk_act.push_back(ia);
      } else {
        keep.push_back(i);
      }
    }
    // the lock branch, k-point by k-point (diag_LOBPCG.jl:125-164)
    for (size_t j = 0; j < locked.size(); ++j) {
      const int i = locked[j];
      rc = kprec_apply(c, iks[i], N, B.q(i, N), B.rows[i], B.q(i, N), B.rows[i]);
      if (rc) return rc;
      Lobpcg L;
      L.c = c;
      L.ik = iks[i];
      L.ispin = ispin;
      L.Nb = B.rows[i];
      L.Nst = N;
      L.prec = (flags & PW_MIXED) ? 1 : 0;
      L.hflags = B.hflags;
      L.Q = B.q(i, 0);
      L.HQ = B.hq(i, 0);
      L.Tmp = B.tmp(i, 0);
      L.Sm = B.lSm;
      L.Gm = B.lGm;
      L.Cm = B.lCm;
      L.S2 = B.lS2;
      L.S3 = B.lS3;
      L.dw = B.ldw;
      L.E = B.lE;
      L.info = B.linfo;
      if (c->d_lob2 && c->lob2_bytes >= (size_t)L.Nb * 5 * N * sizeof(cplx<double>))
        L.L = reinterpret_cast<cplx<double>*>(c->d_lob2);
      rc = L.lock_step(lk_lock[j], lk_act[j]);
      if (rc) return rc;
      ncols[i] += L.ncols;
      nlocks[i]++;
    }
    if (!keep.empty()) {
      const int nq = it > 1 ? 3 * N : 2 * N;
      rc = B.run(1, nq, keep, [&]() { return B.step(keep, nq); });
      if (rc) return rc;
      for (int i : keep) ncols[i] += N;
    }
    active = keep;
    active.insert(active.end(), locked.begin(), locked.end());
    std::sort(active.begin(), active.end());
  }
  // final Rayleigh-Ritz of every k-point: S = (X^H H X + (X^H H X)^H) / 2, X <- X eigvecs(S)
  rc = B.gemm(all, 1, kOne, kZero, 1, [&](int, int p, int i) {
    return gb(B.q(i, 0), B.rows[i], B.hq(i, 0), B.rows[i], B.sx(p), N, N, N, B.rows[i]);
  });
  if (!rc) rc = herm_fix_launch(c, N, nk, B.sX, N, (int64_t)N * N, 2);
  if (!rc) rc = B.heev(N, nk, B.sX, B.dwx, N, B.inf_check());
  if (!rc) rc = B.gemm(all, 0, kOne, kZero, 1, [&](int, int p, int i) {
    return gb(B.q(i, 0), B.rows[i], B.sx(p), N, reinterpret_cast<cplx<double>*>(X[i]), ldx[i], B.rows[i], N, N);
  });
  if (!rc) rc = readback(nk);
  if (rc) return rc;
  std::copy(ev.begin(), ev.begin() + (size_t)nk * N, evals);
  if (stats)
    for (int i = 0; i < nk; ++i) {
      stats[3 * i] = iters[i];
      stats[3 * i + 1] = (int)std::min<long>(ncols[i], 2147483647L);
      stats[3 * i + 2] = nlocks[i];
    }
  return PW_OK;
}
