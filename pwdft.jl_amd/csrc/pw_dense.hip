// Small dense FP64 linear algebra of the eigensolvers on the device: the Cholesky
// factorisation and triangular inverse of the Gram matrices and the (generalised) Hermitian
// eigenproblems of the Rayleigh-Ritz step (diag_LOBPCG.jl:106-118, 176-203;
// ortho_sqrt.jl:6-9).  These are the LAPACK calls the reference makes on the host
// (ZPOTRF, ZTRTRI via inv(U), ZHEGV, ZHEEV); here they run through rocSOLVER on the context
// stream so that the subspace matrices never leave HBM.  Only the upper triangle of the
// input is read, as Julia's Hermitian(A) does.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <string>
#include <vector>

#include "pw_internal.h"

namespace pw {

static int ensure_solver(pw_ctx* c) {
  if (c->rb) return PW_OK;
  rocblas_handle h = nullptr;
  if (rocblas_create_handle(&h) != rocblas_status_success)
    return set_error(c, PW_ERR_HIP, "rocblas_create_handle failed");
  if (rocblas_set_stream(h, c->stream) != rocblas_status_success) {
    rocblas_destroy_handle(h);
    return set_error(c, PW_ERR_HIP, "rocblas_set_stream failed");
  }
  c->rb = h;
  int rc = ensure(c, &c->d_dinfo, &c->dinfo_bytes, 64 * sizeof(double) + 16 * sizeof(int));
  return rc;
}

void free_dense(pw_ctx* c) {
  if (c->rb) rocblas_destroy_handle(reinterpret_cast<rocblas_handle>(c->rb));
  c->rb = nullptr;
  hipFree(c->d_dinfo);
  c->d_dinfo = nullptr;
  c->dinfo_bytes = 0;
  hipFree(c->d_dwork);
  c->d_dwork = nullptr;
  c->dwork_bytes = 0;
  hipFree(c->d_eig);
  c->d_eig = nullptr;
  c->eig_bytes = 0;
  hipFree(c->d_chol);
  c->d_chol = nullptr;
  c->chol_bytes = 0;
  hipFree(c->d_lob);
  c->d_lob = nullptr;
  c->lob_bytes = 0;
  hipFree(c->d_lob2);
  c->d_lob2 = nullptr;
  c->lob2_bytes = 0;
  hipFree(c->d_lobb);
  c->d_lobb = nullptr;
  c->lobb_bytes = 0;
  hipFree(c->d_lobs);
  c->d_lobs = nullptr;
  c->lobs_bytes = 0;
}

static int st_check(pw_ctx* c, rocblas_status s, const char* what) {
  if (s == rocblas_status_success) return PW_OK;
  return set_error(c, PW_ERR_HIP, std::string(what) + ": rocSOLVER status " + std::to_string((int)s));
}

static int read_info(pw_ctx* c, const int* d_info, int* info) {
  hipError_t e = hipMemcpyAsync(info, d_info, sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  return hip_check(c, e, "solver info");
}

// mode 0: A <- Hermitian(A) (upper triangle mirrored, real diagonal); 1: strict lower
// triangle zeroed (an upper-triangular factor); 2: A <- (A + A^H) / 2.
// Matrix blockIdx.y of a strided stack (stride 0 for a single matrix).
__global__ void k_herm_fix(cplx<double>* __restrict__ A, int n, long lda, long stride, int mode) {
  A += (long)blockIdx.y * stride;
  const long tot = (long)n * n;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < tot; t += (long)gridDim.x * blockDim.x) {
    const int i = (int)(t % n), j = (int)(t / n);
    if (i > j) {
      if (mode == 1) {
        A[(long)j * lda + i] = mk(0.0, 0.0);
      } else if (mode == 0) {
        const cplx<double> u = A[(long)i * lda + j];
        A[(long)j * lda + i] = mk(u.x, -u.y);
      }
    } else if (i == j && mode != 1) {
      A[(long)j * lda + i].y = 0.0;
    }
  }
}

__global__ void k_herm_avg(cplx<double>* __restrict__ A, int n, long lda, long stride) {
  A += (long)blockIdx.y * stride;
  const long tot = (long)n * n;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < tot; t += (long)gridDim.x * blockDim.x) {
    const int i = (int)(t % n), j = (int)(t / n);
    if (i < j) {  // one thread per pair writes both entries
      const cplx<double> u = A[(long)j * lda + i], l = A[(long)i * lda + j];
      const cplx<double> m = mk(0.5 * (u.x + l.x), 0.5 * (u.y - l.y));
      A[(long)j * lda + i] = m;
      A[(long)i * lda + j] = mk(m.x, -m.y);
    } else if (i == j) {
      A[(long)j * lda + i].y = 0.0;
    }
  }
}

}  // namespace pw

using namespace pw;

#define CTX_CHECK(c) \
  if (!(c)) return PW_ERR_ARG

namespace pw {

int herm_fix_launch(pw_ctx* c, int n, int nbatch, void* A, int64_t lda, int64_t stride, int mode) {
  const long tot = (long)n * n;
  const int blocks = (int)std::min<long>((tot + 255) / 256, nbatch > 1 ? 64 : 2048);
  cplx<double>* a = reinterpret_cast<cplx<double>*>(A);
  if (mode == 2)
    hipLaunchKernelGGL(k_herm_avg, dim3(blocks, nbatch), dim3(256), 0, c->stream, a, n, (long)lda, (long)stride);
  else
    hipLaunchKernelGGL(k_herm_fix, dim3(blocks, nbatch), dim3(256), 0, c->stream, a, n, (long)lda, (long)stride,
                       mode);
  return hip_check(c, hipGetLastError(), "k_herm_fix");
}

}  // namespace pw

namespace pw {

int solver_handle(pw_ctx* c, void** rb) {
  int rc = ensure_solver(c);
  if (rc) return rc;
  *rb = c->rb;
  return PW_OK;
}

// ---- inv(cholesky(Hermitian(A)).U) above the one-workgroup order, by 2 x 2 blocks
// (A = U^H U, U = [U11 U12; 0 U22]): V11 = inv(U11) by recursion on A11; T = V11^H A12 (= U12);
// A22 -= T^H T; V22 = inv(U22) by recursion on A22; V12 = -V11 (T V22).  The recursion ends in
// blocks of order <= small_dense_max_n() (k_chol_inv_small, one workgroup each); the
// off-diagonal blocks are complex GEMMs.  Each leaf writes its potrf / trtri statuses to its
// own slot, and k_chol_rec_info folds them into the caller's two info words (the failing
// leading minor counted from the whole matrix).  Stable as blocked right-looking Cholesky;
// the inverse's off-diagonal blocks come from the same products as a blocked triangular
// inverse.
constexpr int CR_MAXLEAF = 128;
struct CholLeaves {
  int off[CR_MAXLEAF];
  int n;
};

__global__ void k_chol_rec_info(const int* __restrict__ linfo, CholLeaves lv, int* __restrict__ info) {
  int p = 0, t = 0;
  for (int l = 0; l < lv.n; ++l) {
    if (p == 0 && linfo[2 * l] != 0) p = lv.off[l] + linfo[2 * l];
    if (t == 0 && linfo[2 * l + 1] != 0) t = lv.off[l] + linfo[2 * l + 1];
  }
  info[0] = p;
  info[1] = t;
}

struct CholRec {
  cplx<double>* ws;
  int* linfo;
  CholLeaves lv;
};

static int chol_rec(pw_ctx* c, int n, cplx<double>* A, long lda, int off, CholRec& r) {
  if (n <= small_dense_max_n()) {
    if (r.lv.n >= CR_MAXLEAF) return set_error(c, PW_ERR_ARG, "cholesky inverse: too many blocks");
    const int slot = r.lv.n++;
    r.lv.off[slot] = off;
    return chol_inv_small_batch(c, n, 1, A, lda, 0, r.linfo + 2 * slot);
  }
  const cplx<double> one = mk(1.0, 0.0), mone = mk(-1.0, 0.0), zero = mk(0.0, 0.0);
  const int n1 = n / 2, n2 = n - n1;
  cplx<double>* A11 = A;
  cplx<double>* A12 = A + (long)n1 * lda;
  cplx<double>* A22 = A12 + n1;
  cplx<double>* T = r.ws;
  cplx<double>* W = T + (size_t)n1 * n2;
  r.ws = W + (size_t)n1 * n2;
  int rc = chol_rec(c, n1, A11, lda, off, r);
  if (!rc) rc = zgemm(c, 1, n1, n2, n1, one, A11, lda, A12, lda, zero, T, n1, 0);
  if (!rc) rc = zgemm(c, 1, n2, n2, n1, mone, T, n1, T, n1, one, A22, lda, 0);
  if (!rc) rc = chol_rec(c, n2, A22, lda, off + n1, r);
  if (!rc) rc = zgemm(c, 0, n1, n2, n2, one, T, n1, A22, lda, zero, W, n1, 0);
  if (!rc) rc = zgemm(c, 0, n1, n2, n1, mone, A11, lda, W, n1, zero, A12, lda, 0);
  // the block's strict lower triangle (the caller's unread entries below A11's and A22's
  // diagonals and all of A21) zeroed, so the level above multiplies a triangular factor
  if (!rc) rc = herm_fix_launch(c, n, 1, A, lda, 0, 1);
  r.ws = T;
  return rc;
}

static int chol_inv_rec(pw_ctx* c, int n, cplx<double>* A, long lda, int* d_info) {
  // blocks T and W of every level on the current path: sum of 2 n1 n2 < n^2 complex
  const size_t ws = (size_t)n * n * sizeof(cplx<double>);
  int rc = ensure(c, &c->d_chol, &c->chol_bytes, ws + 2 * CR_MAXLEAF * sizeof(int));
  if (rc) return rc;
  CholRec r;
  r.ws = reinterpret_cast<cplx<double>*>(c->d_chol);
  r.linfo = reinterpret_cast<int*>(reinterpret_cast<char*>(c->d_chol) + ws);
  r.lv.n = 0;
  rc = chol_rec(c, n, A, lda, 0, r);
  if (rc) return rc;
  hipLaunchKernelGGL(k_chol_rec_info, dim3(1), dim3(1), 0, c->stream, r.linfo, r.lv, d_info);
  return hip_check(c, hipGetLastError(), "k_chol_rec_info");
}

// U = cholesky(Hermitian(A)).U and its inverse in place (strict lower triangle zeroed);
// potrf's info into d_info[0], trtri's into d_info[1].  No synchronisation.
int chol_inv_enqueue(pw_ctx* c, int n, cplx<double>* A, long lda, int* d_info) {
  if (c->eig_path == 0 && n <= small_dense_max_n()) return chol_inv_small_batch(c, n, 1, A, lda, 0, d_info);
  if (c->eig_path == 0 && c->chol_rec && n <= small_dense_max_n() * CR_MAXLEAF)
    return chol_inv_rec(c, n, A, lda, d_info);
  int rc = ensure_solver(c);
  if (rc) return rc;
  rocblas_handle h = reinterpret_cast<rocblas_handle>(c->rb);
  rocblas_double_complex* a = reinterpret_cast<rocblas_double_complex*>(A);
  rc = st_check(c, rocsolver_zpotrf(h, rocblas_fill_upper, n, a, (int)lda, d_info), "zpotrf");
  if (rc) return rc;
  rc = st_check(c, rocsolver_ztrtri(h, rocblas_fill_upper, rocblas_diagonal_non_unit, n, a, (int)lda, d_info + 1),
                "ztrtri");
  if (rc) return rc;
  return herm_fix_launch(c, n, 1, A, lda, 0, 1);
}

__global__ void k_heev1(cplx<double>* A, double* w, int vec, int* info) {
  w[0] = A[0].x;
  if (vec) A[0] = mk(1.0, 0.0);
  *info = 0;
}

// Eigenvalues of Hermitian(A) (upper triangle) into d_w, eigenvectors of the nvec lowest
// into A's first nvec columns (nvec = 0: values only); E holds n doubles of workspace and
// d_info one status word.  No synchronisation.
int heev_enqueue(pw_ctx* c, int n, cplx<double>* A, long lda, double* d_w, int nvec, double* E, int* d_info) {
  if (n == 1) {
    hipLaunchKernelGGL(k_heev1, dim3(1), dim3(1), 0, c->stream, A, d_w, nvec > 0 ? 1 : 0, d_info);
    return hip_check(c, hipGetLastError(), "k_heev1");
  }
  if (c->eig_path == 0 && n <= small_dense_max_n()) return heev_small_batch(c, n, 1, A, lda, 0, d_w, nvec, d_info);
  int rc = ensure_solver(c);
  if (rc) return rc;
  if (c->eig_path == 0 && n <= heevd_dist_max_n())
    return heevd_dist(c, c->rb, n, A, lda, d_w, nvec, d_info, c->eig_tri >= 1);
  rocblas_handle h = reinterpret_cast<rocblas_handle>(c->rb);
  return st_check(c, rocsolver_zheevd(h, nvec > 0 ? rocblas_evect_original : rocblas_evect_none, rocblas_fill_upper,
                                      n, reinterpret_cast<rocblas_double_complex*>(A), (int)lda, d_w, E, d_info),
                  "zheevd");
}

}  // namespace pw

extern "C" {

int pw_herm_fix_dev(pw_ctx* c, int n, void* A, int64_t lda, int mode) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (n < 0 || !A || lda < (n > 0 ? n : 1) || mode < 0 || mode > 2)
    return set_error(c, PW_ERR_ARG, "pw_herm_fix_dev: bad arguments");
  if (n == 0) return PW_OK;
  return herm_fix_launch(c, n, 1, A, lda, 0, mode);
}

int pw_herm_fix_batch_dev(pw_ctx* c, int n, int nbatch, void* A, int64_t lda, int64_t strideA, int mode) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (n < 0 || nbatch < 0 || nbatch > 65535 || (n > 0 && nbatch > 0 && !A) || lda < (n > 0 ? n : 1) ||
      strideA < lda * (int64_t)n || mode < 0 || mode > 2)
    return set_error(c, PW_ERR_ARG, "pw_herm_fix_batch_dev: bad arguments");
  if (n == 0 || nbatch == 0) return PW_OK;
  return herm_fix_launch(c, n, nbatch, A, lda, strideA, mode);
}

// potrf + trtri + lower-triangle clear of a strided stack; one potrf info and one trtri
// info per matrix into dinfo[0..nbatch) and dinfo[nbatch..2 nbatch).  No synchronisation.
static int chol_inv_batch_enqueue(pw_ctx* c, int n, int nbatch, void* A, int64_t lda, int64_t strideA, int* dinfo) {
  if (c->eig_path == 0 && n <= small_dense_max_n())
    return chol_inv_small_batch(c, n, nbatch, reinterpret_cast<cplx<double>*>(A), (long)lda, (long)strideA, dinfo);
  if (nbatch == 1 && c->eig_path == 0 && c->chol_rec && n <= small_dense_max_n() * CR_MAXLEAF)
    return chol_inv_rec(c, n, reinterpret_cast<cplx<double>*>(A), (long)lda, dinfo);
  rocblas_handle h = reinterpret_cast<rocblas_handle>(c->rb);
  rocblas_double_complex* a = reinterpret_cast<rocblas_double_complex*>(A);
  int rc = st_check(c, rocsolver_zpotrf_strided_batched(h, rocblas_fill_upper, n, a, (int)lda, (rocblas_stride)strideA,
                                                    dinfo, nbatch),
                "zpotrf_strided_batched");
  if (rc) return rc;
  rc = st_check(c, rocsolver_ztrtri_strided_batched(h, rocblas_fill_upper, rocblas_diagonal_non_unit, n, a, (int)lda,
                                                    (rocblas_stride)strideA, dinfo + nbatch, nbatch),
                "ztrtri_strided_batched");
  if (rc) return rc;
  return herm_fix_launch(c, n, nbatch, A, lda, strideA, 1);
}

static bool chol_inv_args_ok(int n, int nbatch, const void* A, int64_t lda, int64_t strideA) {
  return !(n < 0 || nbatch < 0 || nbatch > 65535 || (n > 0 && nbatch > 0 && !A) || lda < (n > 0 ? n : 1) ||
           strideA < lda * (int64_t)n);
}

int pw_chol_inv_batch_dev(pw_ctx* c, int n, int nbatch, void* A, int64_t lda, int64_t strideA) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (!chol_inv_args_ok(n, nbatch, A, lda, strideA))
    return set_error(c, PW_ERR_ARG, "pw_chol_inv_batch_dev: bad arguments");
  if (n == 0 || nbatch == 0) return PW_OK;
  int rc = ensure_solver(c);
  if (rc) return rc;
  rc = ensure(c, &c->d_dwork, &c->dwork_bytes, 2 * (size_t)nbatch * sizeof(int));
  if (rc) return rc;
  int* dinfo = reinterpret_cast<int*>(c->d_dwork);
  rc = chol_inv_batch_enqueue(c, n, nbatch, A, lda, strideA, dinfo);
  if (rc) return rc;
  std::vector<int> info(2 * (size_t)nbatch);
  hipError_t e = hipMemcpyAsync(info.data(), dinfo, info.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_check(c, e, "batched cholesky info");
  for (int i = 0; i < nbatch; ++i) {
    if (info[i] != 0)
      return set_error(c, PW_ERR_NOT_POSDEF, "cholesky of matrix " + std::to_string(i) + ": leading minor " +
                                                 std::to_string(info[i]) + " not positive definite");
    if (info[nbatch + i] != 0)
      return set_error(c, PW_ERR_NOT_POSDEF, "triangular inverse of matrix " + std::to_string(i) +
                                                 ": singular diagonal " + std::to_string(info[nbatch + i]));
  }
  return PW_OK;
}

int pw_potrf_dev(pw_ctx* c, int n, void* A, int64_t lda) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (n < 0 || !A || lda < (n > 0 ? n : 1)) return set_error(c, PW_ERR_ARG, "pw_potrf_dev: bad arguments");
  if (n == 0) return PW_OK;
  int rc = ensure_solver(c);
  if (rc) return rc;
  rocblas_handle h = reinterpret_cast<rocblas_handle>(c->rb);
  int* dinfo = reinterpret_cast<int*>(c->d_dinfo);
  rc = st_check(c, rocsolver_zpotrf(h, rocblas_fill_upper, n, reinterpret_cast<rocblas_double_complex*>(A), (int)lda,
                                    dinfo),
                "zpotrf");
  if (rc) return rc;
  int info = 0;
  rc = read_info(c, dinfo, &info);
  if (rc) return rc;
  if (info != 0) return set_error(c, PW_ERR_NOT_POSDEF, "cholesky: leading minor " + std::to_string(info) + " not positive definite");
  return PW_OK;
}

int pw_trtri_dev(pw_ctx* c, int n, void* A, int64_t lda) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (n < 0 || !A || lda < (n > 0 ? n : 1)) return set_error(c, PW_ERR_ARG, "pw_trtri_dev: bad arguments");
  if (n == 0) return PW_OK;
  int rc = ensure_solver(c);
  if (rc) return rc;
  rocblas_handle h = reinterpret_cast<rocblas_handle>(c->rb);
  int* dinfo = reinterpret_cast<int*>(c->d_dinfo);
  rc = st_check(c, rocsolver_ztrtri(h, rocblas_fill_upper, rocblas_diagonal_non_unit, n,
                                    reinterpret_cast<rocblas_double_complex*>(A), (int)lda, dinfo),
                "ztrtri");
  if (rc) return rc;
  int info = 0;
  rc = read_info(c, dinfo, &info);
  if (rc) return rc;
  if (info != 0) return set_error(c, PW_ERR_NOT_POSDEF, "triangular inverse: singular diagonal " + std::to_string(info));
  return PW_OK;
}

static int heevd_impl(pw_ctx* c, int n, int m, void* A, int64_t lda, double* w, bool vectors, bool lowest_api);

int pw_heevd_dev(pw_ctx* c, int n, void* A, int64_t lda, double* w) {
  return heevd_impl(c, n, n, A, lda, w, true, false);
}

int pw_heev_values_dev(pw_ctx* c, int n, void* A, int64_t lda, double* w) {
  return heevd_impl(c, n, n, A, lda, w, false, false);
}

int pw_heev_lowest_dev(pw_ctx* c, int n, int m, void* A, int64_t lda, double* w) {
  CTX_CHECK(c);
  if (n < 1 || m < 1 || m > n) return set_error(c, PW_ERR_ARG, "pw_heev_lowest_dev: need 1 <= m <= n");
  return heevd_impl(c, n, m, A, lda, w, true, true);
}

// Eigenvalues of the m lowest eigenpairs into w (host) and, when vectors, their eigenvectors
// into A's first m columns.  lowest_api: the caller keeps only those m (the device
// tridiagonal eigensolver may then compute just them).
static int heevd_impl(pw_ctx* c, int n, int m, void* A, int64_t lda, double* w, bool vectors, bool lowest_api) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (n < 0 || !A || !w || lda < (n > 0 ? n : 1)) return set_error(c, PW_ERR_ARG, "pw_heevd_dev: bad arguments");
  if (n == 0) return PW_OK;
  int rc = ensure_solver(c);
  if (rc) return rc;
  rc = ensure(c, &c->d_dwork, &c->dwork_bytes, 2 * (size_t)n * sizeof(double));
  if (rc) return rc;
  double* D = reinterpret_cast<double*>(c->d_dwork);
  rocblas_handle h = reinterpret_cast<rocblas_handle>(c->rb);
  int* dinfo = reinterpret_cast<int*>(c->d_dinfo);
  if (c->eig_path == 0 && n <= heevd_dist_max_n()) {
    const bool tri_dev = c->eig_tri == 2 || (lowest_api && c->eig_tri == 1);
    rc = n <= small_dense_max_n()
             ? heev_small_batch(c, n, 1, reinterpret_cast<cplx<double>*>(A), (long)lda, 0, D, vectors ? n : 0, dinfo)
             : heevd_dist(c, c->rb, n, reinterpret_cast<cplx<double>*>(A), (long)lda, D, vectors ? m : 0, dinfo,
                          tri_dev);
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(w, D, m * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e != hipSuccess) return hip_check(c, e, "eigenvalues download");
    int info = 0;
    rc = read_info(c, dinfo, &info);
    if (rc) return rc;
    if (info == -1) return set_error(c, PW_ERR_HIP, "eigen: the device tridiagonal reduction timed out");
    if (info != 0) return set_error(c, PW_ERR_ARG, "eigen: no convergence (" + std::to_string(info) + ")");
    return PW_OK;
  }
  rc = st_check(c, rocsolver_zheevd(h, vectors ? rocblas_evect_original : rocblas_evect_none, rocblas_fill_upper,
                                    n, reinterpret_cast<rocblas_double_complex*>(A), (int)lda, D, D + n, dinfo),
                "zheevd");
  if (rc) return rc;
  hipError_t e = hipMemcpyAsync(w, D, m * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e != hipSuccess) return hip_check(c, e, "eigenvalues download");
  int info = 0;
  rc = read_info(c, dinfo, &info);
  if (rc) return rc;
  if (info != 0) return set_error(c, PW_ERR_ARG, "eigen: no convergence (" + std::to_string(info) + ")");
  return PW_OK;
}

int pw_heevd_batch_dev(pw_ctx* c, int n, int nbatch, void* A, int64_t lda, int64_t strideA, double* w) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (n < 0 || nbatch < 0 || (n > 0 && nbatch > 0 && (!A || !w)) || lda < (n > 0 ? n : 1) ||
      strideA < lda * (int64_t)n)
    return set_error(c, PW_ERR_ARG, "pw_heevd_batch_dev: bad arguments");
  if (n == 0 || nbatch == 0) return PW_OK;
  int rc = ensure_solver(c);
  if (rc) return rc;
  // eigenvalues, off-diagonals and one info per matrix
  const size_t nd = 2 * (size_t)n * nbatch;
  rc = ensure(c, &c->d_dwork, &c->dwork_bytes, nd * sizeof(double) + (size_t)nbatch * sizeof(int));
  if (rc) return rc;
  double* D = reinterpret_cast<double*>(c->d_dwork);
  double* E = D + (size_t)n * nbatch;
  int* dinfo = reinterpret_cast<int*>(D + nd);
  rocblas_handle h = reinterpret_cast<rocblas_handle>(c->rb);
  if (c->eig_path == 0 && n <= small_dense_max_n())
    rc = heev_small_batch(c, n, nbatch, reinterpret_cast<cplx<double>*>(A), (long)lda, (long)strideA, D, n, dinfo);
  else
    rc = st_check(c, rocsolver_zheevd_strided_batched(h, rocblas_evect_original, rocblas_fill_upper, n,
                                                      reinterpret_cast<rocblas_double_complex*>(A), (int)lda,
                                                      (rocblas_stride)strideA, D, n, E, n, dinfo, nbatch),
                  "zheevd_strided_batched");
  if (rc) return rc;
  std::vector<int> info(nbatch);
  hipError_t e = hipMemcpyAsync(w, D, (size_t)n * nbatch * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(info.data(), dinfo, (size_t)nbatch * sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_check(c, e, "batched eigenvalues download");
  for (int i = 0; i < nbatch; ++i)
    if (info[i] != 0)
      return set_error(c, PW_ERR_ARG, "eigen: no convergence for matrix " + std::to_string(i) + " (" +
                                          std::to_string(info[i]) + ")");
  return PW_OK;
}

int pw_chol_inv_batch_async_dev(pw_ctx* c, int n, int nbatch, void* A, int64_t lda, int64_t strideA, int* d_info) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (!chol_inv_args_ok(n, nbatch, A, lda, strideA) || (n > 0 && nbatch > 0 && !d_info))
    return set_error(c, PW_ERR_ARG, "pw_chol_inv_batch_async_dev: bad arguments");
  if (n == 0 || nbatch == 0) return PW_OK;
  int rc = ensure_solver(c);
  if (rc) return rc;
  return chol_inv_batch_enqueue(c, n, nbatch, A, lda, strideA, d_info);
}

int pw_heevd_batch_async_dev(pw_ctx* c, int n, int nbatch, void* A, int64_t lda, int64_t strideA, double* d_w,
                             int* d_info) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (n < 0 || nbatch < 0 || (n > 0 && nbatch > 0 && (!A || !d_w || !d_info)) || lda < (n > 0 ? n : 1) ||
      strideA < lda * (int64_t)n)
    return set_error(c, PW_ERR_ARG, "pw_heevd_batch_async_dev: bad arguments");
  if (n == 0 || nbatch == 0) return PW_OK;
  int rc = ensure_solver(c);
  if (rc) return rc;
  // off-diagonals of the tridiagonal forms
  rc = ensure(c, &c->d_dwork, &c->dwork_bytes, (size_t)n * nbatch * sizeof(double));
  if (rc) return rc;
  double* E = reinterpret_cast<double*>(c->d_dwork);
  rocblas_handle h = reinterpret_cast<rocblas_handle>(c->rb);
  if (c->eig_path == 0 && n <= small_dense_max_n())
    return heev_small_batch(c, n, nbatch, reinterpret_cast<cplx<double>*>(A), (long)lda, (long)strideA, d_w, n,
                            d_info);
  return st_check(c, rocsolver_zheevd_strided_batched(h, rocblas_evect_original, rocblas_fill_upper, n,
                                                       reinterpret_cast<rocblas_double_complex*>(A), (int)lda,
                                                       (rocblas_stride)strideA, d_w, n, E, n, d_info, nbatch),
                  "zheevd_strided_batched");
}

int pw_hegvd_dev(pw_ctx* c, int n, void* A, int64_t lda, void* B, int64_t ldb, double* w) {
  CTX_CHECK(c);
  hipSetDevice(c->device);
  if (n < 0 || !A || !B || !w || lda < (n > 0 ? n : 1) || ldb < (n > 0 ? n : 1))
    return set_error(c, PW_ERR_ARG, "pw_hegvd_dev: bad arguments");
  if (n == 0) return PW_OK;
  int rc = ensure_solver(c);
  if (rc) return rc;
  rc = ensure(c, &c->d_dwork, &c->dwork_bytes, 2 * (size_t)n * sizeof(double));
  if (rc) return rc;
  double* D = reinterpret_cast<double*>(c->d_dwork);
  rocblas_handle h = reinterpret_cast<rocblas_handle>(c->rb);
  int* dinfo = reinterpret_cast<int*>(c->d_dinfo);
  rc = st_check(c, rocsolver_zhegvd(h, rocblas_eform_ax, rocblas_evect_original, rocblas_fill_upper, n,
                                    reinterpret_cast<rocblas_double_complex*>(A), (int)lda,
                                    reinterpret_cast<rocblas_double_complex*>(B), (int)ldb, D, D + n, dinfo),
                "zhegvd");
  if (rc) return rc;
  hipError_t e = hipMemcpyAsync(w, D, n * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e != hipSuccess) return hip_check(c, e, "eigenvalues download");
  int info = 0;
  rc = read_info(c, dinfo, &info);
  if (rc) return rc;
  if (info > n) return set_error(c, PW_ERR_NOT_POSDEF, "eigen(T, G): G not positive definite");
  if (info != 0) return set_error(c, PW_ERR_ARG, "eigen(T, G): no convergence (" + std::to_string(info) + ")");
  return PW_OK;
}

}  // extern "C"
