// Accuracy of compressed blocks against their FP64 dense blocks, on the
// device (verify.hip): the relative Frobenius error of every block, its
// matvec error for one vector and the FP64 dense product with a few vectors.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"
#include "kernels.hpp"

namespace farcur {

// Most right-hand sides of the dense reference product per call.
constexpr int kVerifyMaxRhs = 16;

// One block (or, with ref_only, one orientation of a block used only for the
// dense product). Device pointers.
struct VerifyJob {
    const int32_t* rows;  // m RWG ids (the test functions)
    const int32_t* cols;  // n RWG ids
    int32_t m, n, r, ref_only;
    const void* R;        // r x n (ld r), panel precision
    const c128* P;        // P = C U, m x r row-major (P[i * r + k]), FP64
    double* sums;         // [0] sum |Z - CUR|^2, [1] sum |Z|^2 (atomic adds)
    c128* yerr;           // m: (Z - C U R) v, v = V[cols, 0] (NULL: not wanted)
    c128* yref;           // m: Z v
    c128* ys;             // m x s (ld m): Z V[cols, :] (NULL: not wanted)
};

// One block's FP64 interpolation matrix P = C U (m x r row-major).
struct VerifyInterp {
    const void* C;  // m x r (ld m), panel precision
    const c128* U;  // r x r (ld r)
    c128* P;
    int32_t m, r;
};

void launch_verify_interp(const VerifyInterp* jobs, int njobs, int f32, hipStream_t s);
// Row tiles (job tile_job[t], first row tile_row0[t], 64 rows each). V: N x s
// column-major complex128 (s <= kVerifyMaxRhs, s >= 1).
void launch_verify(const VerifyJob* jobs, int ntiles, const int32_t* tile_job, const int32_t* tile_row0, int f32,
                   const c128* V, int N, int s, const FillParams& fp, hipStream_t st);

}  // namespace farcur
