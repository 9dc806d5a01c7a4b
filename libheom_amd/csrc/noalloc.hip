// The device scratch of heom_solve_dev_noalloc: a file-scope __device__ array
// (each GPU holds its own copy until the process ends), addressed through
// hipGetSymbolAddress on each device.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/libheom_amd.h"
#include "common.h"
#include "noalloc.h"

namespace hamd {

static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
void* g_scratch[kMaxDevices] = {};
std::mutex g_mu;

}  // namespace

void* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices)
    throw HeomError{HEOM_ERR_ARG, "noalloc scratch: device ordinal out of range"};
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_scratch[device] == nullptr) {
    int current = 0;
    HAMD_HIP(hipGetDevice(&current));
    HAMD_HIP(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    (void)hipSetDevice(current);
    HAMD_HIP(e);
    if (p == nullptr) throw HeomError{HEOM_ERR_STATE, "noalloc scratch: no device address"};
    g_scratch[device] = p;
  }
  return g_scratch[device];
}

}  // namespace hamd
