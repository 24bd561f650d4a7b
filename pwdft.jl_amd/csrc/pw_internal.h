// Internal definitions shared by the HIP translation units of libpwdft_hip.
// Target: gfx950 (MI355X) only.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pwdft_hip.h"

namespace pw {

template <typename T>
struct alignas(2 * sizeof(T)) cplx {
  T x, y;
};

template <typename T>
__host__ __device__ inline cplx<T> mk(T a, T b) { cplx<T> r; r.x = a; r.y = b; return r; }
template <typename T>
__host__ __device__ inline cplx<T> operator+(cplx<T> a, cplx<T> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename T>
__host__ __device__ inline cplx<T> operator-(cplx<T> a, cplx<T> b) { return mk(a.x - b.x, a.y - b.y); }
template <typename T>
__host__ __device__ inline cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
template <typename T>
__host__ __device__ inline cplx<T> cmulc(cplx<T> a, cplx<T> b) {
  return mk(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <typename T>
__host__ __device__ inline cplx<T> scal(T s, cplx<T> a) { return mk(s * a.x, s * a.y); }
// multiply by +i (s > 0) or -i (s < 0)
template <int S, typename T>
__host__ __device__ inline cplx<T> mul_i(cplx<T> a) {
  return S > 0 ? mk(-a.y, a.x) : mk(a.y, -a.x);
}
template <typename TO, typename TI>
__host__ __device__ inline cplx<TO> cvt(cplx<TI> a) { return mk((TO)a.x, (TO)a.y); }

// One FFT dimension: length n = prod(radix[0..npass)).
struct FftDim {
  int n;
  int npass;
  int radix[8];
  int L;     // lanes of one wavefront that own one line
  int gpw;   // lines per wavefront (64 / L)
  int plan;  // n when a compile-time specialised line FFT exists for this radix plan, else 0
};

// Lanes per line so that no lane holds more than `me` elements across one radix pass
// (shared by the host plan and the compile-time specialised line FFTs).
__host__ __device__ constexpr int lanes_for_pass(int n, int R, int me) {
  return (n / R + ((me / R) > 0 ? (me / R) : 1) - 1) / ((me / R) > 0 ? (me / R) : 1);
}
__host__ __device__ constexpr int cmax(int a, int b) { return a > b ? a : b; }
__host__ __device__ constexpr int lanes_for(int n, int r0, int r1, int r2, int me) {
  return cmax(cmax(lanes_for_pass(n, r0, me), lanes_for_pass(n, r1, me)), r2 > 1 ? lanes_for_pass(n, r2, me) : 1);
}

inline void plan_lanes(FftDim& d, int me) {
  int L = 1;
  for (int p = 0; p < d.npass; ++p) L = cmax(L, lanes_for_pass(d.n, d.radix[p], me));
  d.L = L;
  d.gpw = L <= 64 ? 64 / L : 0;
  // compile-time specialisations (pw_fft.hip: line_static) for these radix plans
  const int r0 = d.npass > 0 ? d.radix[0] : 0, r1 = d.npass > 1 ? d.radix[1] : 0;
  const int r2 = d.npass > 2 ? d.radix[2] : 1;
  d.plan = 0;
  if (d.npass == 2 && ((d.n == 72 && r0 == 9 && r1 == 8) || (d.n == 27 && r0 == 9 && r1 == 3) ||
                       (d.n == 45 && r0 == 9 && r1 == 5) || (d.n == 20 && r0 == 5 && r1 == 4)))
    d.plan = d.n;
  if (d.npass == 3 && d.n == 108 && r0 == 9 && r1 == 4 && r2 == 3) d.plan = 108;
}

// LDS strides: odd line lengths keep strided butterfly reads off a single bank group.
__host__ __device__ inline int odd_stride(int n) { return (n % 2 == 0) ? n + 1 : n; }

// Sphere bookkeeping of one k-point.  Sphere entries are sorted by (column, z);
// a column is one (x, y) pair that holds at least one entry.
struct KData {
  int ngw = 0;
  int ncol = 0;
  int nxact = 0;
  int nbeta = 0;
  int* d_ent_igw = nullptr;   // [ngw]   row of psi for each entry
  int* d_ent_pos = nullptr;   // [ngw]   column * cstride + z (LDS slot in the column tile)
  int* d_col_ptr = nullptr;   // [ncol+1]
  int* d_col_ij = nullptr;    // [ncol]  i + pstride * j (LDS slot inside a padded z-plane)
  int* d_col_cy = nullptr;    // [ncol]  a * ystride + j, x = xact[a] (slot in a compressed plane)
  int* d_xact = nullptr;      // [nxact] distinct x of the columns
  double* d_gk2 = nullptr;    // [ngw]   |G+k|^2
  double* d_gk = nullptr;     // [3*ngw] G+k (x,y,z per entry, column-major 3 x ngw)
  int* d_gpos = nullptr;      // [ngw]   psi row -> x + ps (y + n2 z), ps = odd_stride(n1): slot in a whole padded grid
  void* d_beta64 = nullptr;   // [ngw * nbeta] complex double, column-major
  void* d_betaD64 = nullptr;  // beta * D
  void* d_beta32 = nullptr;   // complex float copies for the mixed path
  void* d_betaD32 = nullptr;
  std::vector<int> h_xact;
};

// One k-point of a batched op_H launch (pw_op_H_batch_dev): its sphere bookkeeping, its
// psi / Hpsi columns and its block of the column work array.
struct KBatch {
  const int* ent_igw;
  const int* ent_pos;
  const int* col_ptr;
  const int* col_ij;
  const int* xact;
  const double* gk2;
  const cplx<double>* psi;
  cplx<double>* hpsi;
  long ldpsi, ldh;
  long woff;   // first element of this k-point's block of W (nb x n3 x ncol)
  int ncol, nxact;
  int blk0;    // first column-kernel workgroup of this k-point (a multiple of 8)
  int pad;
};

// One product of a batched complex GEMM (pw_zgemm_batch_dev).
struct GBatch {
  const cplx<double>* A;
  const cplx<double>* B;
  cplx<double>* C;
  long lda, ldb, ldc;
  int m, n, k;
  int tile0;   // first workgroup of this product
};

// One column block of a batched element-wise launch (copy, Kprec): out = in, or
// out = in / (1 + |G+k|^2) when gk2 is set.
struct EBatch {
  const cplx<double>* in;
  cplx<double>* out;
  const double* gk2;
  long ldin, ldout;
  int m, n;
  int blk0;
  int pad;
};

// A captured step of the batched LOBPCG driver (pw_lobpcg.hip), keyed by the k-points and
// shapes it was captured for; usable while the context's graph_gen is unchanged.
struct LobGraph {
  std::vector<int> key;
  void* exec = nullptr;    // hipGraphExec_t
  void* graph = nullptr;   // hipGraph_t
  unsigned gen = 0;
  bool failed = false;     // capture was refused: this step runs eagerly
  long stamp = 0;          // last use (the least recently used entry is evicted)
};

}  // namespace pw

struct pw_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // second stream for the nonlocal products of op_H, which run beside the FFT passes;
  // forked from and joined back into `stream` with the two events (default priority)
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  bool overlap = true;
  int n1 = 0, n2 = 0, n3 = 0;
  long N = 0;
  double omega = 0.0;
  pw::FftDim dim[3];
  // twiddle tables exp(+2 pi i t / n) per dimension, double and float
  void* d_tw64[3] = {nullptr, nullptr, nullptr};
  void* d_tw32[3] = {nullptr, nullptr, nullptr};
  std::vector<pw::KData> kd;
  int nspin = 0;
  std::vector<double*> d_v64;  // V_loc / N per spin
  std::vector<float*> d_v32;
  int nbeta = 0;
  std::vector<double> h_D;     // nbeta x nbeta column-major
  // workspaces (grow on demand)
  void* d_work = nullptr;  size_t work_bytes = 0;     // FFT column/plane intermediates
  void* d_yglob = nullptr; size_t yglob_bytes = 0;    // compressed planes beyond LDS (global-plane pass)
  void* d_gpart = nullptr; size_t gpart_bytes = 0;    // GEMM split-K partials
  void* d_bnl = nullptr;   size_t bnl_bytes = 0;      // beta^H psi
  void* d_rpart = nullptr; size_t rpart_bytes = 0;    // rho partial planes
  double* d_rho = nullptr;                            // accumulated density [N]
  void* d_tmp = nullptr;   size_t tmp_bytes = 0;      // staging for host-pointer entry points
  void* d_wts = nullptr;   size_t wts_bytes = 0;      // band weights of calc_rhoe
  int band_batch = 64;
  // 1: force the compressed plane kernels, 2: the global-plane pass (tuning / tests)
  int plane_compressed = 0;
  int tune_col_chunk = 0;
  // FP32-path complex GEMMs on BF16 MFMA with 3-part operands: 1 always, 0 never, -1 by size
  int gemm_split_bf16 = -1;
  // tuning (pw_tune_set): [0] = A^H B products, [1] = A B products
  // -1: by N (pw_gemm.hip: zgemm_t); 0: 2x2 wavefronts (64x64 tile), 1: 4x2, 2: 2x4, 3: 4x4,
  // 4: 4x1 (128 x 32), 5: 2x1 (64 x 32)
  int tune_gemm_shape[2] = {-1, -1};
  int tune_gemm_wg[2] = {0, 0};     // split-K target workgroups (0 = 1024)
  std::string err;
  // per-kernel timing (pw_profile_*): event pairs recorded around launches on `stream`
  bool profile = false;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_used;  // (slot, (start, stop))
  double prof_ms[PW_PROF_SLOTS] = {0};
  long prof_count[PW_PROF_SLOTS] = {0};
  double prof_bytes[PW_PROF_SLOTS] = {0};
  // forces and energy gradients (pw_forces.hip)
  pw::KData krho;                      // rho-sphere G-vectors (pw_set_rho_sphere)
  int krho_g0 = -1;                    // sphere entry of G = 0
  int nsp_loc = 0;
  double* d_vg = nullptr;              // [ng * nsp] V_sp(|G|), column-major
  int natoms = 0;
  int species_max = -1;                // largest species index of pw_set_atoms
  int nbeta_atoms = 0;                 // length of beta_atom of pw_set_atoms
  double* d_tau = nullptr;             // [3 * natoms]
  int* d_species = nullptr;            // [natoms]
  int* d_beta_atom = nullptr;          // [nbeta]
  bool dcsr_valid = false;             // D columns on the device match h_D
  int* d_dcol_ptr = nullptr;           // [nbeta + 1]  non-zeros of column j of D
  int* d_drow = nullptr;               // row index of each non-zero
  double* d_dval = nullptr;            // value of each non-zero
  void* d_fbeta = nullptr; size_t fbeta_bytes = 0;  // (G+k)_alpha * beta, alpha = x, y, z
  void* d_fmat = nullptr;  size_t fmat_bytes = 0;   // beta^H psi and ((G+k)_alpha beta)^H psi
  void* d_fgrid = nullptr; size_t fgrid_bytes = 0;  // rho / N on the grid, delta_G0, rho(G)
  void* d_fpart = nullptr; size_t fpart_bytes = 0;  // per-block partial sums
  // batched launches: device copies of descriptor arrays, keyed by content (pw::desc_get)
  std::vector<std::vector<char>> desc_host;
  std::vector<void*> desc_dev;
  std::vector<size_t> desc_cap;
  int desc_next = 0;
  void* d_bnlb = nullptr;  size_t bnlb_bytes = 0;   // beta^H psi of every k-point of a batch
  // stream capture into hipGraphs (pw_graph_*): while capturing, nothing may allocate,
  // upload or synchronise
  bool capturing = false;
  // inside pw_op_H_dev_noalloc: workspaces point into the fixed array and must not grow
  bool no_alloc = false;
  std::vector<void*> graphs;       // hipGraphExec_t per id (nullptr once freed)
  std::vector<void*> graph_defs;   // hipGraph_t per id
  int graphs_live = 0;
  std::vector<void*> retired;      // workspaces replaced while a graph may still use them
  // dense solvers (pw_dense.hip): rocBLAS handle bound to `stream`, info words, eigenvalues
  void* rb = nullptr;
  void* d_dinfo = nullptr; size_t dinfo_bytes = 0;
  void* d_dwork = nullptr; size_t dwork_bytes = 0;
  void* d_eig = nullptr;   size_t eig_bytes = 0;     // reflectors and exchange buffers of heevd_dist
  void* d_chol = nullptr;  size_t chol_bytes = 0;    // blocks and leaf statuses of the recursive Cholesky inverse
  int chol_rec = 1;                                   // Cholesky inverse above order 64: 1 recursive on the device, 0 rocSOLVER
  int eig_path = 0;                                   // 0: device reduction where it fits, 1: rocSOLVER only
  int eig_tri = 1;                                    // tridiagonal eigenpairs: 0 ZSTEDC, 1 device (solver calls), 2 also pw_heevd_dev
  int eig_min_wg = 64;                                // device reduction: minimum workgroups (0: fewest that fit LDS)
  int eig_light = 3;                                  // device reduction: exchanges as device-coherent atomics (no cache flushes)
  void* d_lob = nullptr;   size_t lob_bytes = 0;     // LOBPCG blocks and small matrices (pw_lobpcg.hip)
  void* d_lob2 = nullptr;  size_t lob2_bytes = 0;    // LOBPCG lock-branch blocks
  int lobpcg_lock_exp = 0;                            // 0: lock at 10 eps; k: at 10 * 10^-k
  void* d_lobb = nullptr;  size_t lobb_bytes = 0;    // batched LOBPCG: Q, HQ, Tmp of every k-point
  void* d_lobs = nullptr;  size_t lobs_bytes = 0;    // batched LOBPCG: stacked small matrices, statuses
  bool lob_graph = true;                              // batched LOBPCG steps replayed from hipGraphs
  std::vector<pw::LobGraph> lob_graphs;               // captured steps (pw_lobpcg.hip)
  long lob_stamp = 0;
  // bumped whenever device memory that a captured step may address is allocated or freed;
  // a captured step whose generation differs is discarded, never replayed
  unsigned graph_gen = 0;
  int fused_batch = 1;                                // batched op_H: 1 fused launch (register-line kernel where it
                                                      // applies), 2 fused launch with the whole box in LDS, 0 three stages
  int fft_order = 0;                                  // radix mode of the FFT plans (pw_tune_set PW_TUNE_FFT_ORDER)
  int batch_gemm_small = 1;                           // batched GEMM: small-product kernels where they apply
  int fused_nt = 0;                                   // threads per fused workgroup (0: by LDS, 256/512/1024)
  // density side of an SCF iteration (pw_scf.hip)
  double* d_vps = nullptr;                            // V_Ps_loc(r) [N]
  void* d_sgrid = nullptr; size_t sgrid_bytes = 0;    // grid fields of the potential update
  void* d_ssph = nullptr;  size_t ssph_bytes = 0;     // rho-sphere fields
  void* d_spart = nullptr; size_t spart_bytes = 0;    // reduction partials
  int nsym = 0;                                       // symmetry operations (0: none set)
  int sym_ng = 0;                                     // rho-sphere size the tables were built for
  int* d_symM = nullptr;                              // [9 nsym] m' = M m, row-major 3 x 3
  double* d_symt = nullptr;                           // [3 nsym] fractional translations
  int* d_gm = nullptr;                                // [3 ng] crystal coordinates of the rho-sphere G
  int* d_gtab = nullptr;                              // box of m -> sphere entry (-1: not in the sphere)
  int gtab_lo[3] = {0, 0, 0};
  int gtab_span[3] = {0, 0, 0};
};

namespace pw {

// Brackets one kernel launch with events when profiling is on; `bytes` is the kernel's
// algorithmic HBM traffic for that launch.
struct ProfScope {
  pw_ctx* c;
  int slot;
  hipEvent_t e1 = nullptr;
  ProfScope(pw_ctx* ctx, int s, double bytes);
  ~ProfScope();
};

int set_error(pw_ctx* c, int code, const std::string& msg);
// Destroys the batched LOBPCG driver's captured steps (pw_lobpcg.hip).
void free_lob_graphs(pw_ctx* c);
int hip_check(pw_ctx* c, hipError_t e, const char* what);
int ensure(pw_ctx* c, void** buf, size_t* have, size_t need);
// Looks up the device address of the no-allocation workspace (pw_noalloc.hip) for the
// context's device, once per device; called by pw_ctx_create with that device current.
int noalloc_init(pw_ctx* c);

// Hpsi of the bands psi at k-point ik, spin ispin, into hpsi (terms and precision by flags)
// on the context stream (pw_capi.hip).
int op_H_device(pw_ctx* c, int ik, int ispin, int nb, const cplx<double>* psi, long ldpsi, cplx<double>* hpsi,
                long ldh, unsigned flags);

// FFT drivers (pw_fft.hip)
// join: when set, the stream waits for this event before the first gather pass (the event
// marks work on another stream that writes hpsi first).
int vloc_apply(pw_ctx* c, int ik, int ispin, int nb, const cplx<double>* psi, long ldpsi,
               cplx<double>* hpsi, long ldh, unsigned flags, hipEvent_t join = nullptr);
int kinetic_apply(pw_ctx* c, int ik, int nb, const cplx<double>* psi, long ldpsi, cplx<double>* hpsi,
                  long ldh, unsigned flags);
int rhoe_accum(pw_ctx* c, int ik, int nb, const cplx<double>* psi, long ldpsi, const double* d_w,
               unsigned flags);
int kprec_apply(pw_ctx* c, int ik, int nb, const cplx<double>* in, long ldin, cplx<double>* out,
                long ldout);
// out = FFT[ V(r) * IFFT[psi] ] restricted to the sphere k (FP64, no kinetic term, overwrite);
// V64 holds V(r) / N, the convention of d_v64.
int vloc_apply_k(pw_ctx* c, KData& k, const double* V64, int nb, const cplx<double>* psi, long ldpsi,
                 cplx<double>* out, long ldout);
// Sphere bookkeeping for the index list idx (grid offsets, minus base) with vectors gk (3 x ngw).
int build_kdata(pw_ctx* c, KData& k, int ngw, const int64_t* idx, int base, const double* gk, const char* who);
void free_k(KData& k);
template <typename T>
int upload(pw_ctx* c, T** dst, const T* src, size_t n);
void free_forces(pw_ctx* c);
// Real fields on the sphere k back to the grid (FP64):
// out[f N + r] = scale Re sum_G fG[f ldf + g] e^{+i G.r} for f < nf (added to out when acc).
int g2r_apply(pw_ctx* c, KData& k, int nf, const cplx<double>* fG, long ldf, double* out, double scale, int acc);
void free_scf(pw_ctx* c);
// Non-zeros of each column of D (pw_set_dij) on the device: d_dcol_ptr / d_drow / d_dval.
int ensure_dcsr(pw_ctx* c);
// out[0..n) = 0 with a kernel (safe inside stream capture, unlike a memset node).
int zero_fill(pw_ctx* c, double* out, long n);

// op_H (kinetic + local + nonlocal terms selected by flags) of every band of the k-points
// iks in one launch, one workgroup per (k-point, band) holding the band's whole grid in
// LDS (pw_fft.hip).  Returns PW_ERR_ARG without launching when the grid does not fit.
int op_H_fused_batch(pw_ctx* c, int nk, const int* iks, int ispin, int nb, const cplx<double>* const* psi,
                     const int64_t* ldpsi, cplx<double>* const* hpsi, const int64_t* ldh, unsigned flags);
bool op_H_fused_fits(const pw_ctx* c, bool mixed);
// Batched op_V_loc (+ kinetic) over the k-points iks (pw_fft.hip); psi / hpsi per k-point.
int vloc_apply_batch(pw_ctx* c, int nk, const int* iks, int ispin, int nb, const cplx<double>* const* psi,
                     const int64_t* ldpsi, cplx<double>* const* hpsi, const int64_t* ldh, unsigned flags);
// Batched complex GEMM C_i = alpha op(A_i) B_i + beta C_i (pw_batch.hip).
int zgemm_batch(pw_ctx* c, int transa, const std::vector<GBatch>& g, cplx<double> alpha, cplx<double> beta);
// Device copy of a descriptor array (bytes at `data`).  Arrays are cached by content in up
// to 256 slots, so a call sequence that repeats (the steps of an eigensolver iteration)
// uploads nothing after its first pass.  A miss fills a new slot with an upload ordered on
// the stream (no wait); only once all slots are taken does a miss synchronise the stream
// before it overwrites one, since kernels still queued may read it.
int desc_get(pw_ctx* c, const void* data, size_t bytes, const void** out);
void free_desc(pw_ctx* c);
void free_dense(pw_ctx* c);
// Hermitian eigensolver with the device tridiagonal reduction (pw_eig.hip): eigenvalues into
// d_w (device), eigenvectors of the nvec lowest into A's first nvec columns; enqueued on the
// context stream, status (0, rocSOLVER's code, or -1 for a timed-out reduction) into d_info.
int heevd_dist(pw_ctx* c, void* rb, int n, cplx<double>* A, long lda, double* d_w, int nvec, int* d_info,
               bool lowest = false);
int heevd_dist_max_n();
// dst[:, 0:ncols] = src[:, 0:ncols] (nrows rows) on the context stream.
int copy_cols(pw_ctx* c, int nrows, int ncols, const cplx<double>* src, long lds, cplx<double>* dst, long ldd);
// One launch over the column blocks e (copies, or Kprec where gk2 is set) (pw_batch.hip).
int elem_batch(pw_ctx* c, std::vector<EBatch>& e);
// Hermitian fix-ups of a strided stack (pw_herm_fix_dev modes 0, 1, 2).
int herm_fix_launch(pw_ctx* c, int n, int nbatch, void* A, int64_t lda, int64_t stride, int mode);
// Cholesky factor and its inverse in place (infos into d_info[0..1]), no synchronisation.
int chol_inv_enqueue(pw_ctx* c, int n, cplx<double>* A, long lda, int* d_info);
// Eigenvalues into d_w and the nvec lowest eigenvectors into A (E: n doubles), no
// synchronisation; the device reduction where the order fits, rocSOLVER's ZHEEVD beyond.
int heev_enqueue(pw_ctx* c, int n, cplx<double>* A, long lda, double* d_w, int nvec, double* E, int* d_info);
int solver_handle(pw_ctx* c, void** rb);
// Batched small dense problems (order <= small_dense_max_n(), one workgroup per matrix,
// pw_small.hip): Hermitian eigenpairs (eigenvalues into d_w + b n, the nvec lowest vectors
// into matrix b, status into d_info[b]) and Cholesky factor inverses (statuses into
// d_info[b] and d_info[nbatch + b]).  No synchronisation.
int heev_small_batch(pw_ctx* c, int n, int nbatch, cplx<double>* A, long lda, long stride, double* d_w, int nvec,
                     int* d_info);
int chol_inv_small_batch(pw_ctx* c, int n, int nbatch, cplx<double>* A, long lda, long stride, int* d_info);
int small_dense_max_n();

// GEMM drivers (pw_gemm.hip)
// C (M x N) = alpha * op(A) * B + beta * C ; op(A) = A (transa = 0, A is M x K) or
// A^H (transa = 1, A is K x M).  All pointers are device pointers.
int zgemm(pw_ctx* c, int transa, int M, int N, int K, cplx<double> alpha, const cplx<double>* A,
          long lda, const cplx<double>* B, long ldb, cplx<double> beta, cplx<double>* C, long ldc,
          int prec);
int zgemm_f32in(pw_ctx* c, int transa, int M, int N, int K, cplx<double> alpha, const cplx<float>* A,
                long lda, const cplx<double>* B, long ldb, cplx<double> beta, cplx<double>* C,
                long ldc, int prec);

}  // namespace pw
