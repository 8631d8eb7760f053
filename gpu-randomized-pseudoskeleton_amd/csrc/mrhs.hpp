// Multi-RHS product of the compressed far field (descriptors and launchers).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"

namespace farcur {

enum : int { kMrhsValu = 0, kMrhsMfmaF32 = 1, kMrhsMfmaBf16x3 = 2 };

// Interpolation matrix of one block: P = C U (m x r, column-major), formed in
// FP64 and stored in the panel precision.
struct MrhsInterp {
    const void* c;     // C, m x r column-major (panel precision)
    const c128* u;     // U, r x r column-major
    void* p;           // P, m x r column-major (panel precision)
    int32_t m, r;
};

// One K segment of a grouped GEMM problem: A (M x K) times B (K x s).
//   A[i][k] = a_trans ? a[i * lda + k] : a[k * lda + i]   (panel precision)
//   B[k][j] = b_ids ? B[b_ids[k] * s + j] : B[b_off + k * s + j]
struct MrhsSeg {
    const void* a;
    int32_t lda, a_trans;
    int32_t k, pad;
    const int32_t* b_ids;
    size_t b_off;
};

// One problem: rows [0, m) of the output, the sum over segments
// [seg_begin, seg_end). Row i goes to Out[out_ids[i] * s + j] when out_ids is
// set, else to Out[out_off + i * s + j]. Every output row belongs to exactly one
// problem, so the kernel stores without atomics.
struct MrhsProb {
    const int32_t* out_ids;
    size_t out_off;
    int32_t m, seg_begin, seg_end, pad;
};

int mrhs_chunk();
void launch_mrhs_interp(const MrhsInterp* jobs, int nb, int f32, hipStream_t st);
// ntiles 64-row tiles (problem tile_prob[t], first row tile_row0[t]) times
// ceil(s / 64) right-hand-side chunks.
void launch_mrhs_gemm(const MrhsProb* probs, int ntiles, const int32_t* tile_prob, const int32_t* tile_row0,
                      const MrhsSeg* segs, int f32, int mode, const float2* B, int s, float2* Out, hipStream_t st);

}  // namespace farcur
