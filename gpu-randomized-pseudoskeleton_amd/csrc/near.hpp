// Near-field dense blocks and the compressed operator's near part (near.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"

namespace farcur {

// One dense near-field block Z (m x n, ld m, complex128) of RWG ids rows x cols.
struct NearJob {
    const int32_t* rows;
    const int32_t* cols;
    c128* Z;
    int32_t m, n;
};

// A block and the orientation applied: 0: y[rows] += Z x[cols], 1: y[cols] += Z^T x[rows].
struct NearMv {
    NearJob blk;
    int32_t o, pad;
};

// Row tiles of the fill: 64 rows of job tile_job[t] from tile_row0[t]. rec:
// FP64 RWG records. Triangle pairs whose centroids are closer than
// near_factor x the longer longest edge use singularity extraction.
void launch_near_fill(const NearJob* jobs, int ntiles, const int32_t* tile_job, const int32_t* tile_row0,
                      const double* rec, double k, double eta, double near_factor, hipStream_t s);
// y (FP64 complex, accumulated with atomics) += the listed block products.
void launch_near_matvec(const NearMv* jobs, int njobs, const c128* x, c128* y, hipStream_t s);
void launch_to_c64(const c128* x, c64* y, int n, hipStream_t s);
void launch_add_c64(c128* y, const c64* a, int n, hipStream_t s);

}  // namespace farcur
