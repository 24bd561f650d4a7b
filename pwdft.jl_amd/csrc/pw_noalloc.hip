// op_H with no device allocation: pw_op_H_dev_noalloc runs the same kernels as pw_op_H_dev,
// with every workspace they use -- the FFT work array W, B = beta^H psi and the split-K
// partial sums of the projector products -- carved from one fixed device array, `scratch`,
// instead of the context's growable buffers.  For callers that may not call hipMalloc
// (code under stream capture, tight serving loops).
//
// The array is a file-scope __device__ variable: every device the library runs on holds
// its own copy, at its own address, until the process ends.  pw_ctx_create looks that
// address up once per device (noalloc_init) and host code uses only the looked-up pointer,
// never the variable's name (which is a host-side stand-in, not a device address).
//
// Layout of the array: [split-K partials | B | W].  The band loop of the FFT passes takes
// as many bands per launch as W's share holds, and the projector products split K only as
// far as the partials' share holds, so any number of bands runs; a k-point whose B or a
// single band of W does not fit is rejected with PW_ERR_ARG.
//
// Every call on one device shares that device's array and leaves its contents behind:
// nothing is assumed to start at zero (W and B are fully written before they are read,
// the partials are written by the GEMM before its reduction reads them).  Calls that
// overlap on one device -- two contexts' streams, or two host threads -- race on it.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "pw_internal.h"

namespace pw {

constexpr size_t kScratchBytes = size_t(64) << 20;     // whole array
constexpr size_t kScratchPart = size_t(8) << 20;       // split-K partial sums
constexpr size_t kScratchBnl = size_t(8) << 20;        // B = beta^H psi
constexpr int kMaxDevices = 64;

__device__ __align__(256) char scratch[kScratchBytes];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices];

// With c->device current: the device address of `scratch` on that device, looked up once.
int noalloc_init(pw_ctx* c) {
  if (c->device < 0 || c->device >= kMaxDevices)
    return set_error(c, PW_ERR_ARG, "no-allocation workspace: device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[c->device]) return PW_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return hip_check(c, e, "hipGetSymbolAddress(scratch)");
  if (!p) return set_error(c, PW_ERR_HIP, "hipGetSymbolAddress(scratch) returned a null address");
  g_scratch_ptr[c->device] = p;
  return PW_OK;
}

static void* scratch_ptr(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_ptr[device];
}

}  // namespace pw

using namespace pw;

extern "C" int pw_op_H_dev_noalloc(pw_ctx* c, int ik, int ispin, int nb, const void* psi, int64_t ldpsi,
                                   void* hpsi, int64_t ldh, unsigned flags) {
  if (!c) return PW_ERR_ARG;
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return hip_check(c, e, "hipSetDevice");
  char* base = reinterpret_cast<char*>(scratch_ptr(c->device));
  if (!base) return set_error(c, PW_ERR_STATE, "no-allocation workspace not set up for this device");
  if (c->N == 0) return set_error(c, PW_ERR_STATE, "pw_set_grid has not been called");
  if (ik < 0 || ik >= (int)c->kd.size() || c->kd[ik].ngw == 0)
    return set_error(c, PW_ERR_STATE, "k-point " + std::to_string(ik) + " is not set up");
  const KData& k = c->kd[ik];
  if (nb < 0 || !psi || !hpsi || ldpsi < k.ngw || ldh < k.ngw)
    return set_error(c, PW_ERR_ARG, "pw_op_H_dev_noalloc: bad nb, pointer or leading dimension");
  if (nb == 0) return PW_OK;
  const unsigned terms = flags & PW_TERM_ALL;
  const size_t esz = (flags & PW_MIXED) ? sizeof(cplx<float>) : sizeof(cplx<double>);
  const size_t wband = (size_t)c->n3 * k.ncol * esz;
  const size_t wcap = kScratchBytes - kScratchPart - kScratchBnl;
  if ((terms & PW_TERM_VLOC) && wband > wcap)
    return set_error(c, PW_ERR_ARG, "pw_op_H_dev_noalloc: one band's FFT work array (" + std::to_string(wband) +
                                        " bytes) exceeds the fixed workspace (" + std::to_string(wcap) + " bytes)");
  const size_t bnl = (terms & PW_TERM_NLOC) ? (size_t)k.nbeta * nb * sizeof(cplx<double>) : 0;
  if (bnl > kScratchBnl)
    return set_error(c, PW_ERR_ARG, "pw_op_H_dev_noalloc: beta^H psi (" + std::to_string(k.nbeta) + " x " +
                                        std::to_string(nb) + ") exceeds the fixed workspace (" +
                                        std::to_string(kScratchBnl) + " bytes); call with fewer bands");
  // the context's own buffers are set aside, not freed: queued work may still use them
  void* save_work = c->d_work;
  size_t save_work_b = c->work_bytes;
  void* save_bnl = c->d_bnl;
  size_t save_bnl_b = c->bnl_bytes;
  void* save_part = c->d_gpart;
  size_t save_part_b = c->gpart_bytes;
  void* save_yg = c->d_yglob;
  size_t save_yg_b = c->yglob_bytes;
  c->d_gpart = base;
  c->gpart_bytes = kScratchPart;
  c->d_bnl = base + kScratchPart;
  c->bnl_bytes = kScratchBnl;
  c->d_work = base + kScratchPart + kScratchBnl;
  c->work_bytes = wcap;
  c->d_yglob = nullptr;   // the global-plane pass (planes beyond LDS) has no share: rejected
  c->yglob_bytes = 0;
  c->no_alloc = true;
  const int rc = op_H_device(c, ik, ispin, nb, reinterpret_cast<const cplx<double>*>(psi), ldpsi,
                             reinterpret_cast<cplx<double>*>(hpsi), ldh, flags);
  c->no_alloc = false;
  c->d_work = save_work;
  c->work_bytes = save_work_b;
  c->d_bnl = save_bnl;
  c->bnl_bytes = save_bnl_b;
  c->d_gpart = save_part;
  c->gpart_bytes = save_part_b;
  c->d_yglob = save_yg;
  c->yglob_bytes = save_yg_b;
  return rc;
}
