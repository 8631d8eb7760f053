// Multi-RHS product of the compressed far field (descriptors and launchers).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"

namespace farcur {

enum : int { kMrhsValu = 0, kMrhsMfmaF32 = 1, kMrhsMfmaBf16x3 = 2 };

// Interpolation matrix of one block, P = C U, formed in FP64 and stored
// transposed (P^T: r x m, column-major, the layout of R) in the panel precision.
struct MrhsInterp {
    const void* c;     // C, m x r column-major (panel precision)
    const c128* u;     // U, r x r column-major
    void* p;           // P^T, r x m column-major (panel precision)
    int32_t m, r;
};

// One K segment of a grouped GEMM problem: A (M x K) times B (K x s).
//   A[i][k] = a_trans ? a[i * lda + k] : a[k * lda + i]   (panel precision)
//   B[k][j] = b_ids ? B[b_ids[k] * s + j] : B[b_off + k * s + j]
struct MrhsSeg {
    const void* a;
    int32_t lda, a_trans;
    int32_t k, kbase;  // kbase: first packed K index of the segment in its problem
    const int32_t* b_ids;
    size_t b_off;
};

// Rows [row0, row0 + rows) of a row-segmented problem: A[i][k] =
// a[k * lda + (i - row0)] (column-major, panel precision).
struct MrhsRowSeg {
    const void* a;
    int32_t lda, rows, row0, pad;
};

// One problem: rows [0, m) of the output. Row i goes to Out[out_ids[i] * s + j]
// when out_ids is set, else to Out[out_off + i * s + j]. Every output row
// belongs to exactly one problem, so the kernel stores without atomics.
// K-segmented problems (launched without row segments): the sum over the K
// segments [seg_begin, seg_end) of MrhsSeg, packed along K (k = total K,
// segments sorted by kbase, A read transposed, B dense). Row-segmented problems: rows come
// from the MrhsRowSeg range [seg_begin, seg_end) (sorted by row0), times one
// gathered B, B[k][j] = B[b_ids[k] * s + j] for k below the problem's k.
struct MrhsProb {
    const int32_t* out_ids;
    const int32_t* b_ids;
    size_t out_off;
    int32_t m, k, seg_begin, seg_end;
};

// Single right-hand side (k_matvec1): y[dst] += A2^T (A1 x[src]) with A1 r x K
// and A2 r x M column-major in the panel precision (o = 0: A1 = R_b, A2 = P_b^T,
// src = cols_b, dst = rows_b; o = 1: A1 = P_b^T, A2 = R_b, src = rows_b,
// dst = cols_b). y is accumulated with FP32 atomics (zero it first).
struct Mv1Job {
    const void* a1;
    const void* a2;
    const int32_t* src;
    const int32_t* dst;
    int32_t r, K, M, pad;
};
size_t matvec1_lds_bytes(int rmax);
void launch_matvec1(const Mv1Job* jobs, int njobs, int rmax, int f32, const float2* x, float2* y, hipStream_t st);

int mrhs_chunk();
void launch_mrhs_interp(const MrhsInterp* jobs, int nb, int f32, hipStream_t st);
// ntiles 64-row tiles (problem tile_prob[t], first row tile_row0[t]) times
// ceil(s / 64) right-hand-side chunks.
// rsegs non-NULL: row-segmented problems (segs unused); else K-segmented.
void launch_mrhs_gemm(const MrhsProb* probs, int ntiles, const int32_t* tile_prob, const int32_t* tile_row0,
                      const MrhsSeg* segs, const MrhsRowSeg* rsegs, int f32, int mode, const float2* B, int s,
                      float2* Out, hipStream_t st);

}  // namespace farcur
