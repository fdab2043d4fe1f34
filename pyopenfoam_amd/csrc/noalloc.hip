// pfm_pcg_solve_noalloc: pfm_pcg_solve for callers that cannot allocate
// device memory (code under HIP graph capture, serving loops, harnesses that
// forbid hipMalloc).  The solver workspace -- state, partial sums, residual
// history and the four Krylov vectors -- lives in one file-scope device array
// of kScratchBytes per GPU, held until the process ends, instead of a buffer
// the caller passes in.  The PCG kernels are those of pfm_pcg_solve; they take
// the array's device address as their workspace pointer.

#include <mutex>

#include "pfm_internal.h"

namespace {

constexpr size_t kScratchBytes = size_t(68) << 20;

}  // namespace

// In host code `scratch` names a host-side stand-in, not device memory: its
// device address is taken once per device (scratch_for) and only that address
// is passed to kernels.
static __device__ __align__(256) char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on device `dev`, which must be current: each
// GPU holds its own copy of the array at its own address.
int scratch_for(int dev, void **out) {
  PFM_REQUIRE(dev >= 0 && dev < kMaxDevices, "device ordinal %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    PFM_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    PFM_REQUIRE(p != nullptr, "no device address for the scratch array");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return PFM_OK;
}

}  // namespace

extern "C" {

int64_t pfm_pcg_noalloc_max_bytes(void) { return (int64_t)kScratchBytes; }

int pfm_pcg_solve_noalloc(pfm_ldu h, const pfm_solve_params *prm, const double *diag,
                          const double *packed, const double *rD, const double *b, double *x,
                          double *history_h, pfm_solve_result *res, void *stream) {
  PFM_REQUIRE(h && prm && res, "null handle, params or result");
  PFM_REQUIRE(prm->max_iter >= 0, "max_iter must be >= 0");
  const int64_t need = pfm_pcg_workspace_bytes(h, prm->max_iter);
  PFM_REQUIRE(need >= 0 && (size_t)need <= kScratchBytes,
              "PCG on %d cells with max_iter %d needs %lld workspace bytes; "
              "pfm_pcg_solve_noalloc holds at most %zu",
              h->n, prm->max_iter, (long long)need, kScratchBytes);
  int dev = -1;
  PFM_CHECK_HIP(hipGetDevice(&dev));
  PFM_REQUIRE(dev == h->device, "current device %d is not the handle's device %d", dev,
              h->device);
  void *work = nullptr;
  const int rc = scratch_for(dev, &work);
  if (rc != PFM_OK) return rc;
  return pfm::pcg_solve(h, prm, diag, packed, rD, b, x, work, history_h, res, stream, false);
}

}  // extern "C"
