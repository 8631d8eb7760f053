// Multi-RHS product of the compressed far field, Y = sum_b P_b A_b P_b' V, with
// every block in interpolative form A_b = P_b R_b, P_b = C_b U_b.
//
// The product is evaluated as P (R X) and, for the transposed orientation of a
// symmetric operator, R^T (P^T X). Both intermediates stay at the size of the
// result: the core U = W^+ is ill-conditioned (its norm is 1 / sigma_min of the
// kept W), and the grouping C (U (R X)) would round an intermediate of norm
// |U| |R X| to FP32 and cancel it afterwards, losing log10(cond W) digits.
// P is formed once per block in FP64 (k_mrhs_interp).
//
// Pass A, per contribution (block b, orientation o):
//   T = A1 X  with A1 = R_b, X = V[cols_b]  (o = 0)  or A1 = P_b^T, X = V[rows_b]  (o = 1),
// stored r x s in HBM.
// Pass B, per target group (a set of RWG rows):
//   Y[group] = sum over the group's contributions of L T  (L = P_b, or R_b^T),
// one GEMM whose K dimension concatenates every contribution's r.
// Both passes run the same grouped gather-GEMM kernel on 64 x 64 output tiles
// with K chunks of 32: on VALU (FP32), FP32 MFMA (v_mfma_f32_16x16x4f32) or
// split BF16x3 MFMA (v_mfma_f32_16x16x32_bf16: x = hi + lo,
// xy ~ hi hi + hi lo + lo hi) with FP32 accumulation.
#include <hip/hip_runtime.h>

#include <type_traits>

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

constexpr int kNC = 64;   // right-hand sides per tile
constexpr int kKC = 32;   // K chunk

// ---------------------------------------------------------------- P = C U
// 64 rows of P at a time: computed with lanes over rows (C read coalesced, U
// broadcast), staged in LDS, written as P^T with lanes over the r columns.
constexpr int kInterpMaxR = 64;
template <typename PC>
__global__ __launch_bounds__(256) void k_mrhs_interp(const MrhsInterp* __restrict__ jobs, int nb) {
    __shared__ PC tile[64][kInterpMaxR + 1];
    for (int b = blockIdx.x; b < nb; b += gridDim.x) {
        const MrhsInterp J = jobs[b];
        const PC* C = (const PC*)J.c;
        PC* Pt = (PC*)J.p;
        const int m = J.m, r = J.r;
        if (r > kInterpMaxR) {  // P^T entry P[i][k] at i * r + k, straight from global
            for (int idx = threadIdx.x; idx < m * r; idx += blockDim.x) {
                const int i = idx / r, k = idx - i * r;
                double pr = 0, pi = 0;
                for (int l = 0; l < r; ++l) {
                    const c128 c = to_c128(C[(size_t)l * m + i]);
                    const c128 u = J.u[(size_t)k * r + l];
                    pr += c.re * u.re - c.im * u.im;
                    pi += c.re * u.im + c.im * u.re;
                }
                store_c(Pt + idx, pr, pi);
            }
            continue;
        }
        for (int i0 = 0; i0 < m; i0 += 64) {
            const int rows = min(64, m - i0);
            for (int idx = threadIdx.x; idx < 64 * r; idx += blockDim.x) {
                const int i = idx % 64, k = idx / 64;
                if (i >= rows) continue;
                double pr = 0, pi = 0;
                for (int l = 0; l < r; ++l) {
                    const c128 c = to_c128(C[(size_t)l * m + i0 + i]);
                    const c128 u = J.u[(size_t)k * r + l];
                    pr += c.re * u.re - c.im * u.im;
                    pi += c.re * u.im + c.im * u.re;
                }
                store_c(&tile[i][k], pr, pi);
            }
            __syncthreads();
            for (int idx = threadIdx.x; idx < rows * r; idx += blockDim.x) {
                const int i = idx / r, k = idx - i * r;
                Pt[(size_t)(i0 + i) * r + k] = tile[i][k];
            }
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------- grouped GEMM
struct GemmLds {
    float ar[64][kKC + 1], ai[64][kKC + 1];  // A chunk, re / im planes
    float br[kKC][kNC + 1], bi[kKC][kNC + 1];
};

__device__ __forceinline__ void split_bf16(float x, __bf16& hi, __bf16& lo) {
    hi = (__bf16)x;
    lo = (__bf16)(x - (float)hi);
}

// BF16x3 staging: each element split once into hi / lo as it is staged, real
// and imaginary planes, A as [row][k] and B as [col][k], so that an MFMA
// operand (8 consecutive k) is one 16-byte LDS read. Rows padded by 8 (16 B).
constexpr int kKP = kKC + 8;
struct GemmLdsB {
    __bf16 ah[2][64][kKP], al[2][64][kKP];
    __bf16 bh[2][kNC][kKP], bl[2][kNC][kKP];
};
__device__ __forceinline__ void stage_a(GemmLdsB& L, int i, int k, float2 v) {
    __bf16 h, l;
    split_bf16(v.x, h, l);
    L.ah[0][i][k] = h;
    L.al[0][i][k] = l;
    split_bf16(v.y, h, l);
    L.ah[1][i][k] = h;
    L.al[1][i][k] = l;
}
__device__ __forceinline__ void stage_b(GemmLdsB& L, int k, int j, float2 v) {
    __bf16 h, l;
    split_bf16(v.x, h, l);
    L.bh[0][j][k] = h;
    L.bl[0][j][k] = l;
    split_bf16(v.y, h, l);
    L.bh[1][j][k] = h;
    L.bl[1][j][k] = l;
}
// Eight consecutive k of one A row or B column at once: four 16-byte writes
// (hi / lo of the real and imaginary planes) instead of 32 two-byte ones.
__device__ __forceinline__ void stage8(__bf16 (*h)[64][kKP], __bf16 (*l)[64][kKP], int i, int k0, const float2* v) {
    bf16x8 hr, lr, hi, li;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        __bf16 a, b;
        split_bf16(v[e].x, a, b);
        hr[e] = a;
        lr[e] = b;
        split_bf16(v[e].y, a, b);
        hi[e] = a;
        li[e] = b;
    }
    *(bf16x8*)&h[0][i][k0] = hr;
    *(bf16x8*)&l[0][i][k0] = lr;
    *(bf16x8*)&h[1][i][k0] = hi;
    *(bf16x8*)&l[1][i][k0] = li;
}
typedef short short8 __attribute__((ext_vector_type(8)));
__device__ __forceinline__ bf16x8 neg_bf16x8(bf16x8 x) {
    short8 b = __builtin_bit_cast(short8, x);
    b ^= (short)0x8000;
    return __builtin_bit_cast(bf16x8, b);
}

// FP32 staging (VALU and FP32 MFMA): real and imaginary planes.
__device__ __forceinline__ void stage_a(GemmLds& L, int i, int k, float2 v) {
    L.ar[i][k] = v.x;
    L.ai[i][k] = v.y;
}
__device__ __forceinline__ void stage_b(GemmLds& L, int k, int j, float2 v) {
    L.br[k][j] = v.x;
    L.bi[k][j] = v.y;
}

// Software-pipelined chunks: a thread fetches its 8 A and 8 B entries of the
// next chunk into registers before the current chunk's products, and stores
// them to LDS afterwards (the launch has 256 threads: 64 x 32 / 256 = 8 each).
// A thread's 8 entries of a B column (and, in pass A, of an A row) are 8
// consecutive k, lanes running over the rows / columns (coalesced loads).
constexpr int kGemmThreads = 256;
constexpr int kPerThread = 64 * kKC / kGemmThreads;
static_assert(kKC * kNC == 64 * kKC, "A and B chunks have the same size");
static_assert(kGemmThreads / 64 * kPerThread == kKC && kNC == 64, "four groups of eight k per row or column");
struct ChunkRegs {
    float2 a[kPerThread], b[kPerThread];
};
// pass A layout: A rows from the row table (i = idx % 64), B rows gathered by b_ids
template <typename PC>
__device__ __forceinline__ void fetch_rows(ChunkRegs& cr, const void* const* rt_a, const int* rt_lda,
                                           const int* rt_lrow, const int32_t* __restrict__ b_ids, int K, int j0,
                                           int s, int k0, const float2* __restrict__ B) {
    const int i = threadIdx.x % 64, kg = (threadIdx.x / 64) * kPerThread;
    const PC* A = (const PC*)rt_a[i];
#pragma unroll
    for (int e = 0; e < kPerThread; ++e) {
        const int k = kg + e;
        float2 v = make_float2(0.f, 0.f);
        if (A && k0 + k < K) v = ld_c(A, (size_t)(k0 + k) * rt_lda[i] + rt_lrow[i]);
        cr.a[e] = v;
        float2 w = make_float2(0.f, 0.f);
        if (k0 + k < K && j0 + i < s) w = B[(size_t)b_ids[k0 + k] * s + j0 + i];
        cr.b[e] = w;
    }
}
template <typename Lds>
__device__ __forceinline__ void store_rows(Lds& L, const ChunkRegs& cr) {
    const int i = threadIdx.x % 64, kg = (threadIdx.x / 64) * kPerThread;
    if constexpr (std::is_same<Lds, GemmLdsB>::value) {
        stage8(L.ah, L.al, i, kg, cr.a);
        stage8(L.bh, L.bl, i, kg, cr.b);
    } else {
#pragma unroll
        for (int e = 0; e < kPerThread; ++e) {
            stage_a(L, i, kg + e, cr.a[e]);
            stage_b(L, kg + e, i, cr.b[e]);
        }
    }
}
// pass B layout: A columns from the per-k table (i = idx / kKC), dense B rows
template <typename PC>
__device__ __forceinline__ void fetch_packed(ChunkRegs& cr, const void* const* kt_a, const int* kt_lda,
                                             const int* kt_lk, const size_t* kt_brow, int row0, int m, int j0,
                                             int s, const float2* __restrict__ B) {
#pragma unroll
    for (int e = 0; e < kPerThread; ++e) {
        const int idx = threadIdx.x + e * kGemmThreads;
        const int i = idx / kKC, k = idx % kKC;
        float2 v = make_float2(0.f, 0.f);
        const PC* A = (const PC*)kt_a[k];
        if (A && row0 + i < m) v = ld_c(A, (size_t)(row0 + i) * kt_lda[k] + kt_lk[k]);
        cr.a[e] = v;
        const int j = threadIdx.x % kNC, kb = (threadIdx.x / kNC) * kPerThread + e;
        float2 w = make_float2(0.f, 0.f);
        if (kt_a[kb] && j0 + j < s) w = B[kt_brow[kb] + j0 + j];
        cr.b[e] = w;
    }
}
template <typename Lds>
__device__ __forceinline__ void store_packed(Lds& L, const ChunkRegs& cr) {
    const int j = threadIdx.x % kNC, kg = (threadIdx.x / kNC) * kPerThread;
#pragma unroll
    for (int e = 0; e < kPerThread; ++e) {
        const int idx = threadIdx.x + e * kGemmThreads;
        stage_a(L, idx / kKC, idx % kKC, cr.a[e]);  // A rows: lanes over k (row-major sources)
    }
    if constexpr (std::is_same<Lds, GemmLdsB>::value) {
        stage8(L.bh, L.bl, j, kg, cr.b);
    } else {
#pragma unroll
        for (int e = 0; e < kPerThread; ++e) stage_b(L, kg + e, j, cr.b[e]);
    }
}

// Work item w of a persistent launch, mapped so that each XCD takes one
// contiguous range of items (workgroups are dealt to the 8 XCDs round-robin):
// consecutive tiles of one problem then share an L2.
struct XcdItems {
    int first, end, stride;
    __device__ XcdItems(int total) {
        const int G = gridDim.x;
        if (G < 8) {  // some XCD would have no workgroup: plain striding
            first = blockIdx.x;
            end = total;
            stride = G;
            return;
        }
        const int x = blockIdx.x % 8, j = blockIdx.x / 8;
        const int gx = G / 8 + (x < G % 8 ? 1 : 0);  // workgroups on this XCD (>= 1)
        const int per = (total + 7) / 8;
        first = x * per + j;
        end = min(total, (x + 1) * per);
        stride = gx;
    }
};

// kRowSeg: pass A layout (a problem's rows are stacked contributions with
// their own A, one gathered K range); otherwise pass B layout (one row range,
// K concatenated over segments with their own A and dense B).
// Three workgroups per CU (at most 168 VGPRs + AGPRs): the BF16x3 form would
// otherwise take ~200 and run two; the few registers spilled (<= 112 bytes)
// cost less than the lost overlap (9.4 ms against 10.5 ms for 256
// right-hand sides on the 10242-vertex sphere). Four per CU gains nothing for
// the FP32 forms.
template <typename PC, int kMode, bool kRowSeg>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_mrhs_gemm(const MrhsProb* __restrict__ probs, int ntiles_total,
                                                   const int32_t* __restrict__ tile_prob,
                                                   const int32_t* __restrict__ tile_row0,
                                                   const MrhsSeg* __restrict__ segs,
                                                   const MrhsRowSeg* __restrict__ rsegs, int nchunks,
                                                   const float2* __restrict__ B, int s, float2* __restrict__ Out) {
    using Lds = typename std::conditional<kMode == kMrhsMfmaBf16x3, GemmLdsB, GemmLds>::type;
    __shared__ Lds L;
    __shared__ const void* rt_a[64];
    __shared__ int rt_lda[64], rt_lrow[64];
    __shared__ const void* kt_a[2][kKC];  // per-k tables of the current and the next chunk
    __shared__ int kt_lda[2][kKC], kt_lk[2][kKC];
    __shared__ size_t kt_brow[2][kKC];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const XcdItems it(ntiles_total * nchunks);
    for (int w = it.first; w < it.end; w += it.stride) {
        const int tile = w / nchunks, j0 = (w % nchunks) * kNC;
        const MrhsProb& G = probs[tile_prob[tile]];
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
        auto compute = [&]() {
                if constexpr (kMode == kMrhsValu) {
                    const int ib = (tid / 16) * 4, jb = (tid % 16) * 4;
                    for (int k = 0; k < kKC; ++k) {
                        float ar[4], ai[4], br[4], bi[4];
#pragma unroll
                        for (int a = 0; a < 4; ++a) {
                            ar[a] = L.ar[ib + a][k];
                            ai[a] = L.ai[ib + a][k];
                            br[a] = L.br[k][jb + a];
                            bi[a] = L.bi[k][jb + a];
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
                            a_r[t] = L.ar[row][k4 + lk];
                            a_i[t] = L.ai[row][k4 + lk];
                            const int col = 32 * (wv & 1) + 16 * t + li;
                            b_r[t] = L.br[k4 + lk][col];
                            b_i[t] = L.bi[k4 + lk][col];
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
                    // v_mfma_f32_16x16x32_bf16: lane supplies A[l%16][8*(l/16) + 0..7], B[8*(l/16) + 0..7][l%16],
                    // read as 16-byte vectors of the staged hi / lo planes
                    const int li = lane & 15, lk = (lane >> 4) * 8;
                    // B operands of both column tiles, then one row tile's A operands at a time
                    bf16x8 brh[2], brl[2], bih[2], bil[2];
#pragma unroll
                    for (int t = 0; t < 2; ++t) {
                        const int col = 32 * (wv & 1) + 16 * t + li;
                        brh[t] = *(const bf16x8*)&L.bh[0][col][lk];
                        brl[t] = *(const bf16x8*)&L.bl[0][col][lk];
                        bih[t] = *(const bf16x8*)&L.bh[1][col][lk];
                        bil[t] = *(const bf16x8*)&L.bl[1][col][lk];
                    }
#pragma unroll
                    for (int rt = 0; rt < 2; ++rt) {
                        const int row = 32 * (wv >> 1) + 16 * rt + li;
                        const bf16x8 arh = *(const bf16x8*)&L.ah[0][row][lk];
                        const bf16x8 arl = *(const bf16x8*)&L.al[0][row][lk];
                        const bf16x8 aih = *(const bf16x8*)&L.ah[1][row][lk];
                        const bf16x8 ail = *(const bf16x8*)&L.al[1][row][lk];
                        const bf16x8 naih = neg_bf16x8(aih), nail = neg_bf16x8(ail);
#pragma unroll
                        for (int ct = 0; ct < 2; ++ct) {
#define FARCUR_BF3(acc, AH, AL, BH, BL)                                                \
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(AH, BH[ct], acc, 0, 0, 0);           \
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(AH, BL[ct], acc, 0, 0, 0);           \
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(AL, BH[ct], acc, 0, 0, 0);
                            FARCUR_BF3(accr[rt][ct], arh, arl, brh, brl)
                            FARCUR_BF3(accr[rt][ct], naih, nail, bih, bil)
                            FARCUR_BF3(acci[rt][ct], arh, arl, bih, bil)
                            FARCUR_BF3(acci[rt][ct], aih, ail, brh, brl)
#undef FARCUR_BF3
                        }
                    }
                }
        };
        if constexpr (kRowSeg) {
            __syncthreads();  // the row table may still be read by the previous item
            if (tid < 64) {
                const int row = row0 + tid;
                const void* pa = nullptr;
                int lda = 0, lrow = 0;
                if (row < m) {
                    int lo = G.seg_begin, hi = G.seg_end - 1;  // last segment with row0 <= row
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (rsegs[mid].row0 <= row) lo = mid; else hi = mid - 1;
                    }
                    pa = rsegs[lo].a;
                    lda = rsegs[lo].lda;
                    lrow = row - rsegs[lo].row0;
                }
                rt_a[tid] = pa;
                rt_lda[tid] = lda;
                rt_lrow[tid] = lrow;
            }
            __syncthreads();
            ChunkRegs cr;
            fetch_rows<PC>(cr, rt_a, rt_lda, rt_lrow, G.b_ids, G.k, j0, s, 0, B);
            for (int k0 = 0; k0 < G.k; k0 += kKC) {
                __syncthreads();  // the previous chunk's products are done with L
                store_rows(L, cr);
                __syncthreads();
                if (k0 + kKC < G.k) fetch_rows<PC>(cr, rt_a, rt_lda, rt_lrow, G.b_ids, G.k, j0, s, k0 + kKC, B);
                compute();
            }
        } else {
            // K packed across segments (segment g covers [g.kbase, g.kbase + g.k)):
            // each 32-wide chunk gets a per-k table of (A column, B row), built
            // one chunk ahead in the other table buffer.
            auto build_table = [&](int k0, int t) {
                if (tid < kKC) {
                    const int kk = k0 + tid;
                    const void* pa = nullptr;
                    int lda = 0, lk = 0;
                    size_t brow = 0;
                    if (kk < G.k) {
                        int lo = G.seg_begin, hi = G.seg_end - 1;  // last segment with kbase <= kk
                        while (lo < hi) {
                            const int mid = (lo + hi + 1) >> 1;
                            if (segs[mid].kbase <= kk) lo = mid; else hi = mid - 1;
                        }
                        const MrhsSeg& g = segs[lo];
                        pa = g.a;
                        lda = g.lda;
                        lk = kk - g.kbase;
                        brow = g.b_off + (size_t)lk * s;
                    }
                    kt_a[t][tid] = pa;
                    kt_lda[t][tid] = lda;
                    kt_lk[t][tid] = lk;
                    kt_brow[t][tid] = brow;
                }
            };
            __syncthreads();  // the previous item may still read the tables
            build_table(0, 0);
            __syncthreads();
            ChunkRegs cr;
            fetch_packed<PC>(cr, kt_a[0], kt_lda[0], kt_lk[0], kt_brow[0], row0, m, j0, s, B);
            int c = 0;
            for (int k0 = 0; k0 < G.k; k0 += kKC, ++c) {
                __syncthreads();  // the previous chunk's products are done with L
                store_packed(L, cr);
                const bool next = k0 + kKC < G.k;
                if (next) build_table(k0 + kKC, (c + 1) & 1);
                __syncthreads();
                if (next) {
                    const int t = (c + 1) & 1;
                    fetch_packed<PC>(cr, kt_a[t], kt_lda[t], kt_lk[t], kt_brow[t], row0, m, j0, s, B);
                }
                compute();
            }
        }
        // store
        if constexpr (kMode == kMrhsValu) {
            const int ib = (tid / 16) * 4, jb = (tid % 16) * 4;
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const int i = row0 + ib + a;
                if (i >= m) continue;
                const size_t orow = G.out_ids ? (size_t)G.out_ids[i] * s : G.out_off + (size_t)i * s;
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const int j = j0 + jb + b;
                    if (j < s) Out[orow + j] = make_float2(vr[a][b], vi[a][b]);
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
                        if (i < m && j < s) {
                            const size_t orow = G.out_ids ? (size_t)G.out_ids[i] * s : G.out_off + (size_t)i * s;
                            Out[orow + j] = make_float2(accr[rt][ct][v], acci[rt][ct][v]);
                        }
                    }
        }
    }
}

// ---------------------------------------------------------------- one RHS
// y[dst] += A2^T (A1 x[src]) per contribution, one workgroup each: the pass-A
// vector t (r values) stays in LDS, so each panel is read exactly once and
// nothing intermediate goes to HBM. t = A1 x: a thread owns one k and one
// slice of the columns (consecutive threads read consecutive entries of a
// column), the slices are summed in LDS. y: a thread owns one output row.
template <typename PC>
__global__ __launch_bounds__(256) void k_matvec1(const Mv1Job* __restrict__ jobs, const float2* __restrict__ x,
                                                 float* __restrict__ y) {
    extern __shared__ float2 lds1[];
    const Mv1Job J = jobs[blockIdx.x];
    const PC* A1 = (const PC*)J.a1;
    const PC* A2 = (const PC*)J.a2;
    const int r = J.r, K = J.K, M = J.M, tid = threadIdx.x;
    float2* t = lds1;
    float2* part = lds1 + r;
    if (r <= 256) {
        const int ns = 256 / r, k = tid % r, sl = tid / r;
        float ar = 0, ai = 0;
        if (sl < ns)
            for (int j = sl; j < K; j += ns) {
                const float2 a = ld_c(A1, (size_t)j * r + k), v = x[J.src[j]];
                ar = fmaf(a.x, v.x, fmaf(-a.y, v.y, ar));
                ai = fmaf(a.x, v.y, fmaf(a.y, v.x, ai));
            }
        part[tid] = make_float2(ar, ai);
        __syncthreads();
        if (tid < r) {
            float sr = 0, si = 0;
            for (int q = 0; q < ns; ++q) {
                sr += part[q * r + tid].x;
                si += part[q * r + tid].y;
            }
            t[tid] = make_float2(sr, si);
        }
    } else {
        for (int k = tid; k < r; k += 256) {
            float ar = 0, ai = 0;
            for (int j = 0; j < K; ++j) {
                const float2 a = ld_c(A1, (size_t)j * r + k), v = x[J.src[j]];
                ar = fmaf(a.x, v.x, fmaf(-a.y, v.y, ar));
                ai = fmaf(a.x, v.y, fmaf(a.y, v.x, ai));
            }
            t[k] = make_float2(ar, ai);
        }
    }
    __syncthreads();
    for (int i = tid; i < M; i += 256) {
        const PC* row = A2 + (size_t)i * r;
        float ar = 0, ai = 0;
        for (int k = 0; k < r; ++k) {
            const float2 a = ld_c(row, k), v = t[k];
            ar = fmaf(a.x, v.x, fmaf(-a.y, v.y, ar));
            ai = fmaf(a.x, v.y, fmaf(a.y, v.x, ai));
        }
        const int d = J.dst[i];
        atomicAdd(&y[2 * d], ar);
        atomicAdd(&y[2 * d + 1], ai);
    }
}

}  // namespace

size_t matvec1_lds_bytes(int rmax) { return ((size_t)rmax + 256) * sizeof(float2); }

void launch_matvec1(const Mv1Job* jobs, int njobs, int rmax, int f32, const float2* x, float2* y, hipStream_t st) {
    if (njobs <= 0) return;
    const size_t lds = matvec1_lds_bytes(rmax);
    if (f32)
        hipLaunchKernelGGL(k_matvec1<c64>, dim3(njobs), dim3(256), lds, st, jobs, x, (float*)y);
    else
        hipLaunchKernelGGL(k_matvec1<c128>, dim3(njobs), dim3(256), lds, st, jobs, x, (float*)y);
}

int mrhs_chunk() { return kNC; }

void launch_mrhs_interp(const MrhsInterp* jobs, int nb, int f32, hipStream_t st) {
    if (nb <= 0) return;
    const int grid = nb < 16384 ? nb : 16384;
    if (f32)
        hipLaunchKernelGGL(k_mrhs_interp<c64>, dim3(grid), dim3(256), 0, st, jobs, nb);
    else
        hipLaunchKernelGGL(k_mrhs_interp<c128>, dim3(grid), dim3(256), 0, st, jobs, nb);
}

void launch_mrhs_gemm(const MrhsProb* probs, int ntiles, const int32_t* tile_prob, const int32_t* tile_row0,
                      const MrhsSeg* segs, const MrhsRowSeg* rsegs, int f32, int mode, const float2* B, int s,
                      float2* Out, hipStream_t st) {
    if (ntiles <= 0) return;
    const int nchunks = (s + kNC - 1) / kNC;
    const int work = ntiles * nchunks;
    const int grid = work < 65536 ? work : 65536;
#define FARCUR_LAUNCH(PC, MODE, RS)                                                                            \
    hipLaunchKernelGGL((k_mrhs_gemm<PC, MODE, RS>), dim3(grid), dim3(256), 0, st, probs, ntiles, tile_prob,     \
                       tile_row0, segs, rsegs, nchunks, B, s, Out)
#define FARCUR_LAUNCH_M(PC, RS)                                                                                \
    if (mode == kMrhsValu) FARCUR_LAUNCH(PC, kMrhsValu, RS);                                                  \
    else if (mode == kMrhsMfmaF32) FARCUR_LAUNCH(PC, kMrhsMfmaF32, RS);                                       \
    else FARCUR_LAUNCH(PC, kMrhsMfmaBf16x3, RS);
    if (rsegs) {
        if (f32) { FARCUR_LAUNCH_M(c64, true) } else { FARCUR_LAUNCH_M(c128, true) }
    } else {
        if (f32) { FARCUR_LAUNCH_M(c64, false) } else { FARCUR_LAUNCH_M(c128, false) }
    }
#undef FARCUR_LAUNCH_M
#undef FARCUR_LAUNCH
}

}  // namespace farcur
