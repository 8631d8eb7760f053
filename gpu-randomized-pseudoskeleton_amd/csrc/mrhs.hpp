// Multi-RHS product of the compressed far field (descriptors and launchers).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"

namespace farcur {

enum : int { kMrhsValu = 0, kMrhsMfmaF32 = 1, kMrhsMfmaBf16x3 = 2 };

// One contribution: a block in one orientation. Pass A forms
// T2 = U' (A1 X) with X = V[x_ids]; pass B adds L T2 to the target group.
struct MrhsContrib {
    const void* a1;        // R (r x n, ld r) or C (m x r, ld m) read transposed
    int32_t a1_ld, a1_trans;
    int32_t k1;            // columns of A1 (n, or m when transposed)
    int32_t r;
    const int32_t* x_ids;  // k1 RWG ids selecting rows of V
    const c128* u;         // r x r
    int32_t u_trans, pad0;
    const void* l;         // C (m x r, ld m) or R (r x n) read transposed
    int32_t l_ld, l_trans;
    size_t t2_off;         // offset (complex elements) of T2 (r x s, row-major)
};

// A target group: rows ids[0..m) of Y, contributions [c_begin, c_end) of the
// contribution-of-group list.
struct MrhsGroup {
    const int32_t* ids;
    int32_t m;
    int32_t c_begin, c_end, pad;
};

int mrhs_max_r();
int mrhs_chunk();
// rmax: largest r of the listed contributions (<= mrhs_max_r()).
void launch_mrhs_pass_a(const MrhsContrib* cs, int nc, int rmax, int f32, const float2* V, int s, float2* T2,
                        hipStream_t st);
void launch_mrhs_pass_b(const MrhsGroup* groups, int ntiles, const int32_t* tile_group, const int32_t* tile_row0,
                        const MrhsContrib* cs, const int32_t* contrib_of_group, int f32, int mode, const float2* T2,
                        int s, float2* Y, hipStream_t st);

}  // namespace farcur
