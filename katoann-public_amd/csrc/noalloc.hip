// The allocation-free search entry point (no reference counterpart): for
// callers that cannot allocate device memory, such as code under HIP graph
// capture, tight serving loops and harnesses that forbid hipMalloc. It runs
// katoann_search_ex's pipeline (device pointers) in a static device array of
// KATOANN_NOALLOC_SCRATCH_BYTES per GPU instead of buffers of its own, and
// calls no hipMalloc, hipFree or other allocating routine.
//
// Each GPU the library runs on holds its own copy of the array until the
// process ends. Host code never passes the array's name to a kernel (on the
// host it is a stand-in, not a device address): on a device's first call the
// array's device address is looked up with hipGetSymbolAddress, with that
// device current, and kept per device. Calls on one GPU share its array and
// would race, so they are serialised here; the array keeps the last call's
// contents, and every buffer the pipeline needs cleared is cleared by its own
// kernels.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/katoann_c.h"

static __device__ __align__(256) unsigned char scratch[KATOANN_NOALLOC_SCRATCH_BYTES];

extern std::string &katoann_error_slot();
extern "C" int katoann_search_in_scratch(const katoann_index *h, const float *queries, uint32_t nq,
                                         const katoann_search_params *p, const katoann_search_extras *x,
                                         uint32_t *ids, float *dists, void *scratch, uint64_t bytes);

namespace {

constexpr int kMaxDevices = 64;

struct DeviceScratch {
  std::once_flag once;
  hipError_t status = hipSuccess;
  void *ptr = nullptr;  // device address of this GPU's copy of `scratch`
  std::mutex mu;        // one call at a time on this GPU's copy
};

DeviceScratch g_dev[kMaxDevices];

}  // namespace

extern "C" int katoann_search_noalloc(const katoann_index *h, const float *queries, uint32_t nq,
                                      const katoann_search_params *p, const katoann_search_extras *x,
                                      uint32_t *ids, float *dists) {
  int dev = -1;
  const hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    katoann_error_slot() = std::string("allocation-free search: no current device: ") + hipGetErrorString(e);
    return KATOANN_RUNTIME_ERROR;
  }
  DeviceScratch &d = g_dev[dev];
  std::call_once(d.once, [&d] { d.status = hipGetSymbolAddress(&d.ptr, HIP_SYMBOL(scratch)); });
  if (d.status != hipSuccess || !d.ptr) {
    katoann_error_slot() =
        std::string("allocation-free search: hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(d.status);
    return KATOANN_RUNTIME_ERROR;
  }
  std::lock_guard<std::mutex> g(d.mu);
  return katoann_search_in_scratch(h, queries, nq, p, x, ids, dists, d.ptr, KATOANN_NOALLOC_SCRATCH_BYTES);
}
