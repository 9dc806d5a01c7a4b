// Device scratch of the entry point that allocates nothing
// (heom_solve_dev_noalloc): one fixed-size array per GPU.
#pragma once

#include <cstddef>

#include "kernels.h"

namespace hamd {

// Layout of the scratch: the Dormand-Prince control block, the error-norm
// partials (one double per workgroup of the widest stage pass), then rho_S of
// every output time.
constexpr size_t kNoallocScratchBytes = size_t(8) << 20;
constexpr size_t kNoallocCtrlOff = 0;
constexpr size_t kNoallocPartialsOff = 256;
constexpr size_t kNoallocRhoSOff = kNoallocPartialsOff + sizeof(double) * kMaxGrid;
constexpr size_t kNoallocRhoSBytes = kNoallocScratchBytes - kNoallocRhoSOff;

// Device address of the scratch on `device`, looked up once per device (with
// that device current) and kept for the life of the process.
void* noalloc_scratch(int device);

}  // namespace hamd
