// Fixed device workspace of lbswim_system_step_noalloc: a file-scope device
// array of NOALLOC_SCRATCH_BYTES, so that entry point never allocates.  Each
// GPU the library runs on holds its own copy of the array at its own
// address, found once per device with hipGetSymbolAddress; host code uses
// only that address (the array's name on the host is not a device address).
#include <hip/hip_runtime.h>

#include <mutex>

#include "internal.hpp"

namespace lbs {
namespace {

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
void* g_scratch[MAX_DEVICES] = {};
std::mutex g_scratch_mu;

}  // namespace

hipError_t noalloc_scratch(void** out) {
  *out = nullptr;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return hipSuccess;
}

}  // namespace lbs
