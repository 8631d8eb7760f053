// Allocation-free pinv (noalloc.hip): a static device scratch array per GPU.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include "common.hpp"

namespace farcur {

// Largest core the allocation-free pinv takes (its scratch is sized for it).
constexpr int kNoAllocMaxR = 256;
// Bytes of the static scratch array (held on every GPU the library runs on).
size_t noalloc_scratch_bytes();
// Takes the scratch array's device address on `device` (which must be the
// current device); once per device, at context creation. Returns 0 or the
// HIP error code.
int noalloc_init(int device);
// Device address of the scratch array on `device` (NULL before noalloc_init).
unsigned char* noalloc_scratch(int device);
// Enqueues the pinv of the r x r matrix W (device) on s: U (device, r x r),
// sigma (device, r, descending; may be NULL), *kept (device; may be NULL).
// base = noalloc_scratch(device). No allocation.
void launch_pinv_noalloc(unsigned char* base, const c128* W, int r, double atol, double rtol, int rmax, int threads,
                         c128* U, double* sigma, int* kept, hipStream_t s);

}  // namespace farcur
