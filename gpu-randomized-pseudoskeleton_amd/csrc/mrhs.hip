// Multi-RHS product of the compressed far field, Y = sum_b P_b C_b U_b R_b V.
//
// Pass A, per contribution (block b, orientation o) and 64 right-hand sides:
//   T1 = A1 X  (A1 = R_b, X = V[cols_b]; or, transposed, A1 = C_b^T, X = V[rows_b])
//   T2 = U' T1 (U' = U_b or U_b^T), stored r x s in HBM.
// Pass B, per target group (a set of RWG rows), 64-row tile and 64 RHS:
//   Y[group] = sum over the group's contributions of L T2  (L = C_b or R_b^T),
// i.e. one GEMM whose K dimension concatenates every contribution's r.
// Every Y row belongs to exactly one group, so pass B writes without atomics.
// Pass B runs on VALU (FP32), FP32 MFMA (v_mfma_f32_16x16x4f32) or split BF16x3
// MFMA (v_mfma_f32_16x16x32_bf16: x = hi + lo, xy ~ hi hi + hi lo + lo hi) with
// FP32 accumulation.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "mrhs.hpp"

namespace farcur {

namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

template <typename PC>
__device__ __forceinline__ float2 ld_c(const PC* p, size_t i) {
    const c128 v = to_c128(p[i]);
    return make_float2((float)v.re, (float)v.im);
}
__device__ __forceinline__ float2 ld_c(const c64* p, size_t i) { return make_float2(p[i].re, p[i].im); }
__device__ __forceinline__ float2 ld_u(const c128* p, size_t i) { return make_float2((float)p[i].re, (float)p[i].im); }

constexpr int kNC = 64;   // right-hand sides per work item
constexpr int kKC = 32;   // K chunk of pass B
constexpr int kMaxRA = 256;

// ---------------------------------------------------------------- pass A
// T1 (r x 64 complex) lives in dynamic LDS sized for the launch's largest r.
template <typename PC>
__global__ __launch_bounds__(256) void k_mrhs_pass_a(const MrhsContrib* __restrict__ cs, int nc, int nchunks,
                                                     const float2* __restrict__ V, int s,
                                                     float2* __restrict__ T2) {
    extern __shared__ __align__(16) float2 t1_flat[];
    float2(*t1)[kNC] = reinterpret_cast<float2(*)[kNC]>(t1_flat);
    const int tid = threadIdx.x, j = tid % kNC, q = tid / kNC;  // 4 row phases
    for (int w = blockIdx.x; w < nc * nchunks; w += gridDim.x) {
        const MrhsContrib& c = cs[w / nchunks];
        const int j0 = (w % nchunks) * kNC;
        const int r = c.r, K1 = c.k1;
        const PC* A1 = (const PC*)c.a1;
        const bool jv = j0 + j < s;
        // T1[i][j] = sum_k A1[i][k] V[x_ids[k]][j0 + j]
        for (int i = q; i < r; i += 4) {
            float ar = 0, ai = 0;
            for (int k = 0; k < K1; ++k) {
                const float2 a = c.a1_trans ? ld_c(A1, (size_t)i * c.a1_ld + k) : ld_c(A1, (size_t)k * c.a1_ld + i);
                const float2 x = jv ? V[(size_t)c.x_ids[k] * s + j0 + j] : make_float2(0.f, 0.f);
                ar += a.x * x.x - a.y * x.y;
                ai += a.x * x.y + a.y * x.x;
            }
            t1[i][j] = make_float2(ar, ai);
        }
        __syncthreads();
        // T2[i][j] = sum_l U'[i][l] T1[l][j]
        for (int i = q; i < r; i += 4) {
            float ar = 0, ai = 0;
            for (int l = 0; l < r; ++l) {
                const float2 u = c.u_trans ? ld_u(c.u, (size_t)i * r + l) : ld_u(c.u, (size_t)l * r + i);
                const float2 x = t1[l][j];
                ar += u.x * x.x - u.y * x.y;
                ai += u.x * x.y + u.y * x.x;
            }
            if (jv) T2[c.t2_off + (size_t)i * s + j0 + j] = make_float2(ar, ai);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- pass B
struct PassBLds {
    float lr[64][kKC + 1], li[64][kKC + 1];  // L chunk, re / im planes
    float tr[kKC][kNC + 1], ti[kKC][kNC + 1];
};

// Loads the K chunk [k0, k0 + 32) of contribution c into LDS (zero padded).
template <typename PC>
__device__ __forceinline__ void load_chunk(const MrhsContrib& c, int row0, int m, int j0, int s, int k0,
                                           const float2* __restrict__ T2, PassBLds& L) {
    const PC* Lp = (const PC*)c.l;
    for (int idx = threadIdx.x; idx < 64 * kKC; idx += blockDim.x) {
        const int i = idx % 64, k = idx / 64;  // rows fastest: coalesced for C (col-major)
        float2 v = make_float2(0.f, 0.f);
        if (row0 + i < m && k0 + k < c.r)
            v = c.l_trans ? ld_c(Lp, (size_t)(row0 + i) * c.l_ld + k0 + k) : ld_c(Lp, (size_t)(k0 + k) * c.l_ld + row0 + i);
        L.lr[i][k] = v.x;
        L.li[i][k] = v.y;
    }
    for (int idx = threadIdx.x; idx < kKC * kNC; idx += blockDim.x) {
        const int j = idx % kNC, k = idx / kNC;
        float2 v = make_float2(0.f, 0.f);
        if (k0 + k < c.r && j0 + j < s) v = T2[c.t2_off + (size_t)(k0 + k) * s + j0 + j];
        L.tr[k][j] = v.x;
        L.ti[k][j] = v.y;
    }
}

__device__ __forceinline__ void split_bf16(float x, __bf16& hi, __bf16& lo) {
    hi = (__bf16)x;
    lo = (__bf16)(x - (float)hi);
}

template <typename PC, int kMode>
__global__ __launch_bounds__(256) void k_mrhs_pass_b(const MrhsGroup* __restrict__ groups, int ntiles_total,
                                                     const int32_t* __restrict__ tile_group,
                                                     const int32_t* __restrict__ tile_row0,
                                                     const MrhsContrib* __restrict__ cs,
                                                     const int32_t* __restrict__ contrib_of_group, int nchunks,
                                                     const float2* __restrict__ T2, int s, float2* __restrict__ Y) {
    __shared__ PassBLds L;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int w = blockIdx.x; w < ntiles_total * nchunks; w += gridDim.x) {
        const int tile = w / nchunks, j0 = (w % nchunks) * kNC;
        const MrhsGroup& G = groups[tile_group[tile]];
        const int row0 = tile_row0[tile], m = G.m;
        // accumulators
        float vr[4][4], vi[4][4];    // VALU: rows (tid/16)*4 + a, cols (tid%16)*4 + b
        floatx4 accr[2][2], acci[2][2];  // MFMA: wave sub-tile rows 32*(wv/2) + 16*rt, cols 32*(wv%2) + 16*ct
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) vr[a][b] = vi[a][b] = 0.f;
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) accr[a][b] = acci[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
        for (int ci = G.c_begin; ci < G.c_end; ++ci) {
            const MrhsContrib& c = cs[contrib_of_group[ci]];
            for (int k0 = 0; k0 < c.r; k0 += kKC) {
                __syncthreads();
                load_chunk<PC>(c, row0, m, j0, s, k0, T2, L);
                __syncthreads();
                if constexpr (kMode == kMrhsValu) {
                    const int ib = (tid / 16) * 4, jb = (tid % 16) * 4;
                    for (int k = 0; k < kKC; ++k) {
                        float ar[4], ai[4], br[4], bi[4];
#pragma unroll
                        for (int a = 0; a < 4; ++a) {
                            ar[a] = L.lr[ib + a][k];
                            ai[a] = L.li[ib + a][k];
                            br[a] = L.tr[k][jb + a];
                            bi[a] = L.ti[k][jb + a];
                        }
#pragma unroll
                        for (int a = 0; a < 4; ++a)
#pragma unroll
                            for (int b = 0; b < 4; ++b) {
                                vr[a][b] += ar[a] * br[b] - ai[a] * bi[b];
                                vi[a][b] += ar[a] * bi[b] + ai[a] * br[b];
                            }
                    }
                } else if constexpr (kMode == kMrhsMfmaF32) {
                    // v_mfma_f32_16x16x4f32: lane supplies A[l%16][l/16], B[l/16][l%16]
                    const int li = lane & 15, lk = lane >> 4;
#pragma unroll
                    for (int k4 = 0; k4 < kKC; k4 += 4) {
                        float a_r[2], a_i[2], b_r[2], b_i[2];
#pragma unroll
                        for (int t = 0; t < 2; ++t) {
                            const int row = 32 * (wv >> 1) + 16 * t + li;
                            a_r[t] = L.lr[row][k4 + lk];
                            a_i[t] = L.li[row][k4 + lk];
                            const int col = 32 * (wv & 1) + 16 * t + li;
                            b_r[t] = L.tr[k4 + lk][col];
                            b_i[t] = L.ti[k4 + lk][col];
                        }
#pragma unroll
                        for (int rt = 0; rt < 2; ++rt)
#pragma unroll
                            for (int ct = 0; ct < 2; ++ct) {
                                accr[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a_r[rt], b_r[ct], accr[rt][ct], 0, 0, 0);
                                accr[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(-a_i[rt], b_i[ct], accr[rt][ct], 0, 0, 0);
                                acci[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a_r[rt], b_i[ct], acci[rt][ct], 0, 0, 0);
                                acci[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a_i[rt], b_r[ct], acci[rt][ct], 0, 0, 0);
                            }
                    }
                } else {
                    // v_mfma_f32_16x16x32_bf16: lane supplies A[l%16][8*(l/16) + 0..7], B[8*(l/16) + 0..7][l%16]
                    const int li = lane & 15, lk = (lane >> 4) * 8;
                    bf16x8 arh[2], arl[2], aih[2], ail[2], naih[2], nail[2], brh[2], brl[2], bih[2], bil[2];
#pragma unroll
                    for (int t = 0; t < 2; ++t) {
                        const int row = 32 * (wv >> 1) + 16 * t + li;
                        const int col = 32 * (wv & 1) + 16 * t + li;
#pragma unroll
                        for (int e = 0; e < 8; ++e) {
                            __bf16 h, l;
                            split_bf16(L.lr[row][lk + e], h, l);
                            arh[t][e] = h;
                            arl[t][e] = l;
                            split_bf16(L.li[row][lk + e], h, l);
                            aih[t][e] = h;
                            ail[t][e] = l;
                            split_bf16(-L.li[row][lk + e], h, l);
                            naih[t][e] = h;
                            nail[t][e] = l;
                            split_bf16(L.tr[lk + e][col], h, l);
                            brh[t][e] = h;
                            brl[t][e] = l;
                            split_bf16(L.ti[lk + e][col], h, l);
                            bih[t][e] = h;
                            bil[t][e] = l;
                        }
                    }
#pragma unroll
                    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
                        for (int ct = 0; ct < 2; ++ct) {
#define FARCUR_BF3(acc, AH, AL, BH, BL)                                                     \
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(AH[rt], BH[ct], acc, 0, 0, 0);            \
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(AH[rt], BL[ct], acc, 0, 0, 0);            \
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(AL[rt], BH[ct], acc, 0, 0, 0);
                            FARCUR_BF3(accr[rt][ct], arh, arl, brh, brl)
                            FARCUR_BF3(accr[rt][ct], naih, nail, bih, bil)
                            FARCUR_BF3(acci[rt][ct], arh, arl, bih, bil)
                            FARCUR_BF3(acci[rt][ct], aih, ail, brh, brl)
#undef FARCUR_BF3
                        }
                }
            }
        }
        // store
        if constexpr (kMode == kMrhsValu) {
            const int ib = (tid / 16) * 4, jb = (tid % 16) * 4;
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const int i = row0 + ib + a;
                if (i >= m) continue;
                const int gid = G.ids[i];
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const int j = j0 + jb + b;
                    if (j < s) Y[(size_t)gid * s + j] = make_float2(vr[a][b], vi[a][b]);
                }
            }
        } else {
            // D[4*(l/16) + v][l%16] of each 16x16 tile
#pragma unroll
            for (int rt = 0; rt < 2; ++rt)
#pragma unroll
                for (int ct = 0; ct < 2; ++ct)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int i = row0 + 32 * (wv >> 1) + 16 * rt + 4 * (lane >> 4) + v;
                        const int j = j0 + 32 * (wv & 1) + 16 * ct + (lane & 15);
                        if (i < m && j < s) Y[(size_t)G.ids[i] * s + j] = make_float2(accr[rt][ct][v], acci[rt][ct][v]);
                    }
        }
    }
}

}  // namespace

int mrhs_max_r() { return kMaxRA; }
int mrhs_chunk() { return kNC; }

void launch_mrhs_pass_a(const MrhsContrib* cs, int nc, int rmax, int f32, const float2* V, int s, float2* T2,
                        hipStream_t st) {
    if (nc <= 0) return;
    const int nchunks = (s + kNC - 1) / kNC;
    const int work = nc * nchunks;
    const int grid = work < 16384 ? work : 16384;
    const size_t lds = (size_t)rmax * kNC * sizeof(float2);
    static bool attr = false;
    if (!attr) {
        const int maxb = kMaxRA * kNC * (int)sizeof(float2);
        (void)hipFuncSetAttribute((const void*)k_mrhs_pass_a<c64>, hipFuncAttributeMaxDynamicSharedMemorySize, maxb);
        (void)hipFuncSetAttribute((const void*)k_mrhs_pass_a<c128>, hipFuncAttributeMaxDynamicSharedMemorySize, maxb);
        attr = true;
    }
    if (f32)
        hipLaunchKernelGGL(k_mrhs_pass_a<c64>, dim3(grid), dim3(256), lds, st, cs, nc, nchunks, V, s, T2);
    else
        hipLaunchKernelGGL(k_mrhs_pass_a<c128>, dim3(grid), dim3(256), lds, st, cs, nc, nchunks, V, s, T2);
}

void launch_mrhs_pass_b(const MrhsGroup* groups, int ntiles, const int32_t* tile_group, const int32_t* tile_row0,
                        const MrhsContrib* cs, const int32_t* contrib_of_group, int f32, int mode, const float2* T2,
                        int s, float2* Y, hipStream_t st) {
    if (ntiles <= 0) return;
    const int nchunks = (s + kNC - 1) / kNC;
    const int work = ntiles * nchunks;
    const int grid = work < 16384 ? work : 16384;
#define FARCUR_LAUNCH_B(PC, MODE)                                                                                   \
    hipLaunchKernelGGL((k_mrhs_pass_b<PC, MODE>), dim3(grid), dim3(256), 0, st, groups, ntiles, tile_group,        \
                       tile_row0, cs, contrib_of_group, nchunks, T2, s, Y)
    if (f32) {
        if (mode == kMrhsValu) FARCUR_LAUNCH_B(c64, kMrhsValu);
        else if (mode == kMrhsMfmaF32) FARCUR_LAUNCH_B(c64, kMrhsMfmaF32);
        else FARCUR_LAUNCH_B(c64, kMrhsMfmaBf16x3);
    } else {
        if (mode == kMrhsValu) FARCUR_LAUNCH_B(c128, kMrhsValu);
        else if (mode == kMrhsMfmaF32) FARCUR_LAUNCH_B(c128, kMrhsMfmaF32);
        else FARCUR_LAUNCH_B(c128, kMrhsMfmaBf16x3);
    }
#undef FARCUR_LAUNCH_B
}

}  // namespace farcur
