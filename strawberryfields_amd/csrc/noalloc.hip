// Allocation-free variant of the fidelity loss (no counterpart in the
// reference): for callers that cannot allocate device memory -- code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
//
// sfamd_fidelity_loss needs a caller-provided partial-sum buffer of
// sfamd_overlap_partials(n) doubles.  This variant takes it from a file-scope
// device array instead.  Each GPU holds its own copy of the array until the
// process ends; its address is looked up once per device (with that device
// current) on the first call there, and only that address is handed to the
// kernels.  Every call on one GPU shares the array: the partial-sum kernel
// writes every slot the final reduction reads, so nothing needs clearing
// between calls, but calls that overlap on one GPU (different streams) race.
#include <mutex>

#include "common.h"

// the overlap reduction uses at most 2 * 1024 partial doubles (1024
// workgroups, real and imaginary parts), whatever the state size
static __device__ __align__(256) double scratch[SFAMD_NOALLOC_SCRATCH_DOUBLES];

namespace {

constexpr int64_t kScratchDoubles = SFAMD_NOALLOC_SCRATCH_DOUBLES;
constexpr int kMaxDevices = 64;

double* g_scratch[kMaxDevices];
std::mutex g_scratch_mutex;

// device address of `scratch` on the current device (looked up once)
int scratch_ptr(double** out) {
  int dev = 0;
  HIP_RET(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return SFAMD_E_ARG;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    HIP_RET(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return SFAMD_E_WORKSPACE;
    g_scratch[dev] = (double*)p;
  }
  *out = g_scratch[dev];
  return 0;
}

}  // namespace

extern "C" int sfamd_fidelity_loss_noalloc(const void* target, const void* psi, int64_t n_elems, double* out_dev,
                                           void* g_psi, int dtype, void* stream) {
  CHECK_ARG(target && psi && out_dev && n_elems >= 0);
  if (sfamd_overlap_partials(n_elems) > kScratchDoubles) return SFAMD_E_WORKSPACE;
  double* partial = nullptr;
  RET_IF(scratch_ptr(&partial));
  return sfamd_fidelity_loss(target, psi, n_elems, partial, out_dev, g_psi, dtype, stream);
}
