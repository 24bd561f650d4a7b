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
