// Far-field CUR kernels for gfx950: sampling, pivoting, panel fill, core pinv
// and the CUR residual, all driven by per-round BlockJob lists so that one
// launch of each serves one block (the per-block drop-in) or every active block
// of a batch (persistent workgroups striding over the work).
#include <hip/hip_runtime.h>

#include <type_traits>

#include "common.hpp"
#include "curmv.hpp"
#include "element.hpp"
#include "jacobi.hpp"
#include "kernels.hpp"
#include "philox.hpp"

namespace farcur {

// ---------------------------------------------------------------- sampling
// Thread (job, kind): kind 0 draws r of the n block columns, kind 1 r of the m
// block rows, by selection sampling; output is sorted by position in the block
// and mapped to RWG ids (compres_single_block_spheres_for_sweep2.jl:492-497:
// sample(1:NN, r, replace=false), sort, r2[samples]).
// Words of stream (block, round, purpose) in groups of four.
__device__ __forceinline__ U4 stream4(uint64_t seed, uint32_t block, uint32_t round, uint32_t purpose, uint32_t t4) {
    U4 c;
    c.x = t4 >> 2;
    c.y = block;
    c.z = (round << 8) | purpose;
    c.w = 0x46435552u;
    return philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// Held-out rows (h of m, stream kStreamHeldout) and r sample rows among the
// other m - h positions (stream kStreamRows, word index = candidate index),
// in one pass over the positions: both are selection scans in position order,
// so interleaving them makes the same decisions as running them one after the
// other. Positions are recorded first and mapped to RWG ids afterwards, so the
// scan carries no dependent global loads.
__device__ void sample_rows_heldout(const BlockJob& J, uint64_t seed) {
    // the job's fields in registers: the scan's stores could otherwise alias
    // the job record and force reloads every iteration
    const uint32_t m = J.m, h = J.h, r = J.r, N = m - h, b = J.b, round = J.round;
    int32_t* const hpos = J.hpos;
    int32_t* const rpos = J.rpos;
    uint32_t goth = 0, got = 0, cand = 0;
    U4 oh{}, orow{};
    for (uint32_t t = 0; t < m && (goth < h || got < r); ++t) {
        if ((t & 3u) == 0 && goth < h) oh = stream4(seed, b, round, kStreamHeldout, t);
        if (goth < h) {
            const uint32_t l = t & 3u;
            const uint32_t u = l == 0 ? oh.x : (l == 1 ? oh.y : (l == 2 ? oh.z : oh.w));
            if (select_step(u, m - t, h - goth)) {
                hpos[goth++] = t;
                continue;
            }
        }
        if (got < r) {
            if ((cand & 3u) == 0) orow = stream4(seed, b, round, kStreamRows, cand);
            const uint32_t l = cand & 3u;
            const uint32_t u = l == 0 ? orow.x : (l == 1 ? orow.y : (l == 2 ? orow.z : orow.w));
            if (select_step(u, N - cand, r - got)) rpos[got++] = t;
        }
        ++cand;
    }
    const int32_t* const rows = J.rows;
    int32_t* const hs = J.hs;
    int32_t* const rs = J.rs;
#pragma unroll 8
    for (uint32_t k = 0; k < goth; ++k) hs[k] = rows[hpos[k]];
#pragma unroll 8
    for (uint32_t k = 0; k < got; ++k) rs[k] = rows[rpos[k]];
}

__global__ void k_sample(const BlockJob* __restrict__ jobs, int njobs, uint64_t seed) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = t >> 1, kind = t & 1;
    if (j >= njobs) return;
    const BlockJob& J = jobs[j];
    if (J.sample_mode == 1) return;
    if (kind == 1 && J.sample_mode == 2) return;
    if (kind == 1 && J.sample_mode == 3) {
        sample_rows_heldout(J, seed);
        return;
    }
    const uint32_t N = kind ? J.m : J.n;
    const int32_t* ids = kind ? J.rows : J.cols;
    int32_t* out = kind ? J.rs : J.cs;
    // positions first (into rpos for rows when the job keeps them, else into
    // the output), RWG ids after the scan
    int32_t* pos = kind == 1 && J.rpos ? J.rpos : out;
    const uint32_t purpose = kind ? kStreamRows : kStreamCols;
    const uint32_t need = J.r, b = J.b, round = J.round;
    uint32_t got = 0;
    for (uint32_t t4 = 0; t4 < N && got < need; t4 += 4) {
        U4 c;
        c.x = t4 >> 2;
        c.y = b;
        c.z = (round << 8) | purpose;
        c.w = 0x46435552u;
        const U4 o = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
        const uint32_t w[4] = {o.x, o.y, o.z, o.w};
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            const uint32_t tt = t4 + l;
            if (tt < N && got < need && select_step(w[l], N - tt, need - got)) pos[got++] = tt;
        }
    }
#pragma unroll 8
    for (uint32_t k = 0; k < got; ++k) out[k] = ids[pos[k]];
}

// The same scans with a wave per (job, kind), for launches of few jobs. A
// scan takes candidate t of N when q_t = (u_t (N - t)) >> 32 is below the
// number still needed (select_step). The lanes hold 64 consecutive candidates
// and compute their q_t in parallel; the decisions, a chain through the count
// taken so far, run over the chunk's candidate lanes in order on scalar
// registers (one scalar unit per CU: many such waves contend for it, so a
// launch of many jobs keeps the thread-per-scan kernel). The held-out
// decisions do not depend on the sample rows, so a chunk's held-out chain runs
// before its sample-row chain, as in the interleaved scan.
__device__ __forceinline__ uint64_t select_chunk(uint32_t q, uint64_t cands, uint32_t& got, uint32_t need) {
    uint64_t taken = 0;
    while (cands && got < need) {
        const int l = __builtin_ctzll(cands);
        cands &= cands - 1;
        if ((uint32_t)__builtin_amdgcn_readlane((int)q, l) < need - got) {
            taken |= 1ull << l;
            ++got;
        }
    }
    return taken;
}
// set bits of mask below this lane
__device__ __forceinline__ uint32_t lane_rank(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}
__global__ __launch_bounds__(256) void k_sample_wave(const BlockJob* __restrict__ jobs, int njobs, uint64_t seed) {
    // the wave's index, made visibly uniform: the job's fields and the chains
    // then live in scalar registers
    const int wid = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) / kWave));
    const int lane = threadIdx.x & (kWave - 1);
    const int j = wid >> 1, kind = wid & 1;
    if (j >= njobs) return;
    const BlockJob& J = jobs[j];
    if (J.sample_mode == 1) return;
    if (kind == 1 && J.sample_mode == 2) return;
    const uint32_t b = J.b, round = J.round;
    if (kind == 1 && J.sample_mode == 3) {
        const uint32_t m = J.m, h = J.h, r = J.r, N = m - h;
        uint32_t goth = 0, got = 0, cand0 = 0;
        for (uint32_t base = 0; base < m && (goth < h || got < r); base += kWave) {
            const uint32_t t = base + lane;
            const bool in = t < m;
            const uint64_t valid = __ballot(in);
            uint64_t held = 0;
            if (goth < h) {
                uint32_t q = 0;
                if (in) q = (uint32_t)(((uint64_t)stream_word(seed, b, round, kStreamHeldout, t) * (m - t)) >> 32);
                const uint32_t g0 = goth;
                held = select_chunk(q, valid, goth, h);
                if ((held >> lane) & 1) {
                    const uint32_t k = g0 + lane_rank(held);
                    J.hpos[k] = t;
                    J.hs[k] = J.rows[t];
                }
            }
            const uint64_t cands = valid & ~held;
            if (got < r) {
                const uint32_t ci = cand0 + lane_rank(cands);  // this position's candidate index
                uint32_t q = 0;
                if ((cands >> lane) & 1)
                    q = (uint32_t)(((uint64_t)stream_word(seed, b, round, kStreamRows, ci) * (N - ci)) >> 32);
                const uint32_t g0 = got;
                const uint64_t taken = select_chunk(q, cands, got, r);
                if ((taken >> lane) & 1) {
                    const uint32_t k = g0 + lane_rank(taken);
                    J.rpos[k] = t;
                    J.rs[k] = J.rows[t];
                }
            }
            cand0 += __popcll(cands);
        }
        return;
    }
    const uint32_t N = kind ? J.m : J.n;
    const int32_t* ids = kind ? J.rows : J.cols;
    int32_t* out = kind ? J.rs : J.cs;
    int32_t* pos = kind == 1 ? J.rpos : nullptr;  // sample rows keep their positions when the job does
    const uint32_t purpose = kind ? kStreamRows : kStreamCols;
    const uint32_t need = J.r;
    uint32_t got = 0;
    for (uint32_t base = 0; base < N && got < need; base += kWave) {
        const uint32_t t = base + lane;
        uint32_t q = 0;
        if (t < N) q = (uint32_t)(((uint64_t)stream_word(seed, b, round, purpose, t) * (N - t)) >> 32);
        const uint32_t g0 = got;
        const uint64_t taken = select_chunk(q, __ballot(t < N), got, need);
        if ((taken >> lane) & 1) {
            const uint32_t k = g0 + lane_rank(taken);
            out[k] = ids[t];
            if (pos) pos[k] = t;
        }
    }
}

// Launches of at most this many jobs (the learned mode's later rounds, single
// blocks) take the wave-per-scan kernel.
constexpr int kSampleWaveMaxJobs = 4096;

void launch_sample(const BlockJob* jobs, int njobs, uint64_t seed, hipStream_t s) {
    if (njobs <= 0) return;
    const int threads = 256;
    if (njobs <= kSampleWaveMaxJobs) {
        const int grid = (2 * njobs * kWave + threads - 1) / threads;  // four (job, kind) waves per workgroup
        hipLaunchKernelGGL(k_sample_wave, dim3(grid), dim3(threads), 0, s, jobs, njobs, seed);
        return;
    }
    const int grid = (2 * njobs + threads - 1) / threads;
    hipLaunchKernelGGL(k_sample, dim3(grid), dim3(threads), 0, s, jobs, njobs, seed);
}

// Rows by pivoting: for every sampled column, the block row of largest |C|
// (first one on ties), as cur_smart_pinv2
// (codes/GPU_implementation/seed/compres_single_block_spheres.jl:2870-2887).
template <typename PC>
__global__ __launch_bounds__(256) void k_pivot_rows(const BlockJob* __restrict__ jobs, int njobs) {
    for (int jb = blockIdx.x; jb < njobs; jb += gridDim.x) {
        const BlockJob& J = jobs[jb];
        if (J.sample_mode != 2) continue;
        const PC* C = (const PC*)J.C;
        const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave, nw = blockDim.x / kWave;
        for (int col = w; col < J.r; col += nw) {
            double best = -1.0;
            int besti = 0x7fffffff;
            for (int i = lane; i < J.m; i += kWave) {
                const c128 a = to_c128(C[(size_t)col * J.m + i]);
                const double v = a.re * a.re + a.im * a.im;
                if (v > best) {
                    best = v;
                    besti = i;
                }
            }
            for (int o = kWave / 2; o > 0; o >>= 1) {
                const double ob = __shfl_xor(best, o);
                const int oi = __shfl_xor(besti, o);
                if (ob > best || (ob == best && oi < besti)) {
                    best = ob;
                    besti = oi;
                }
            }
            if (besti >= J.m) besti = 0;  // all entries NaN: keep a valid row
            if (lane == 0) {
                J.rs[col] = J.rows[besti];
                if (J.rpos) J.rpos[col] = besti;
            }
        }
    }
}

void launch_pivot_rows(const BlockJob* jobs, int njobs, int f32, hipStream_t s) {
    if (njobs <= 0) return;
    const int grid = njobs < 4096 ? njobs : 4096;
    if (f32)
        hipLaunchKernelGGL(k_pivot_rows<c64>, dim3(grid), dim3(256), 0, s, jobs, njobs);
    else
        hipLaunchKernelGGL(k_pivot_rows<c128>, dim3(grid), dim3(256), 0, s, jobs, njobs);
}

// ---------------------------------------------------------------- fill
__host__ __device__ static inline int cdiv(int a, int b) { return (a + b - 1) / b; }

// C and W: lanes over rows; R and H (few rows, many columns): lanes over
// columns, R's r sample rows and H's h held-out rows in the same tiles.
int fill_tiles(int m, int n, int r, int h, int panels) {
    int t = 0;
    const int hh = (panels & kPanelH) && h > 0 ? h : 0;
    if (panels & kPanelC) t += cdiv(m, kTileRows) * cdiv(r, kTileCols);
    if (panels & kPanelW) t += cdiv(r, kTileRows) * cdiv(r, kTileCols);
    if (panels & kPanelR)
        t += cdiv(n, kTileRows) * cdiv(r + hh, kTileCols);
    else if (hh > 0)
        t += cdiv(n, kTileRows) * cdiv(hh, kTileCols);
    return t;
}

// Fill constants for the wavenumber k and impedance eta.
__device__ inline F32Consts fill_consts_f32(double k, double eta) {
    const double il = k / (2.0 * M_PI);
    return F32Consts{(float)il, (float)(4.0 / (k * k)), (float)(k * eta / (144.0 * M_PI)), il};
}

constexpr LdsTail kNoTail{nullptr, 0, nullptr, 0};

// A tile of a panel in the fill's precision: FP64 per entry, FP32 through the
// tile's distinct triangles.
template <typename T>
struct FillLdsOf;
template <>
struct FillLdsOf<double> {
    using type = FillTileLds;
};
template <>
struct FillLdsOf<float> {
    using type = FillTriLds;
};

template <typename T, typename OutC, bool kLanesAreRows>
__device__ __forceinline__ void fill_any(OutC* out, int ld, const int* rows, int nr, const int* cols, int nc,
                                         int row0, int col0, const int32_t* lane_perm, const T* rec,
                                         const double* cen, const TriGeom& tg, double k, double eta,
                                         typename FillLdsOf<T>::type& lds, LdsTail tail) {
    if constexpr (sizeof(T) == 8) {
        const FillConsts fc = fill_consts_f64(k, eta);
        fill_tile<OutC, kLanesAreRows>(out, ld, rows, nr, cols, nc, row0, col0, rec, fc.k, fc.c4k2, fc.scale, lds,
                                       tail);
    } else {
        fill_tile_tri<OutC, kLanesAreRows>(out, ld, rows, nr, cols, nc, row0, col0, lane_perm, rec, cen, tg,
                                           fill_consts_f32(k, eta), lds, tail);
    }
}

// Persistent fill over the tiles of all jobs. T = double: C, W, R in c128.
// T = float: C, R in c64 and W widened to c128 for the FP64 pinv.
template <typename T, typename PC>
__global__ __launch_bounds__(kFillThreads) void k_fill(const BlockJob* __restrict__ jobs, int njobs,
                                                       const int32_t* __restrict__ prefix, int total, int panels,
                                                       const T* __restrict__ rec, const double* __restrict__ cen,
                                                       TriGeom tg, double k, double eta) {
    __shared__ typename FillLdsOf<T>::type lds;
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        int lo = 0, hi = njobs - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (prefix[mid] <= t) lo = mid; else hi = mid - 1;
        }
        const BlockJob& J = jobs[lo];
        int local = t - prefix[lo];
        const int m = J.m, n = J.n, r = J.r;
        const int panels_j = J.rpos ? (panels & ~kPanelW) : panels;  // W is gathered from C
        if (panels_j & kPanelC) {
            const int mt = cdiv(m, kTileRows), nC = mt * cdiv(r, kTileCols);
            if (local < nC) {
                fill_any<T, PC, true>((PC*)J.C, m, J.rows, m, J.cs, r, (local % mt) * kTileRows,
                                      (local / mt) * kTileCols, J.row_perm, rec, cen, tg, k, eta, lds, kNoTail);
                continue;
            }
            local -= nC;
        }
        const int nt = cdiv(n, kTileRows);
        if (panels_j & kPanelW) {
            const int rt = cdiv(r, kTileRows);
            const int nW = rt * cdiv(r, kTileCols);
            if (local < nW) {
                fill_any<T, c128, true>(J.W, r, J.rs, r, J.cs, r, (local % rt) * kTileRows, (local / rt) * kTileCols,
                                        nullptr, rec, cen, tg, k, eta, lds, kNoTail);
                continue;
            }
            local -= nW;
        }
        if (panels_j & kPanelR) {
            // R's sample rows and H's held-out rows share the lane set (the
            // block columns): one tile range over the r + h rows
            const int hh = (panels_j & kPanelH) && J.h > 0 ? J.h : 0;
            const int nR = nt * cdiv(r + hh, kTileCols);
            if (local < nR) {
                fill_any<T, PC, false>((PC*)J.R, r, J.rs, r, J.cols, n, (local / nt) * kTileCols,
                                       (local % nt) * kTileRows, J.col_perm, rec, cen, tg, k, eta, lds,
                                       LdsTail{J.hs, hh, J.H, J.h});
                continue;
            }
            continue;  // no H-only tiles follow
        }
        fill_any<T, PC, false>((PC*)J.H, J.h, J.hs, J.h, J.cols, n, (local / nt) * kTileCols, (local % nt) * kTileRows,
                               J.col_perm, rec, cen, tg, k, eta, lds, LdsTail{J.hs, 0, J.H, J.h});
    }
}

void launch_fill(const BlockJob* jobs, int njobs, const int32_t* tile_prefix, int total_tiles, int panels,
                 const FillParams& fp, int grid, hipStream_t s) {
    if (njobs <= 0 || total_tiles <= 0) return;
    if (grid <= 0 || grid > total_tiles) grid = total_tiles;
    if (fp.f32)
        hipLaunchKernelGGL((k_fill<float, c64>), dim3(grid), dim3(kFillThreads), 0, s, jobs, njobs, tile_prefix,
                           total_tiles, panels, (const float*)fp.rec, fp.cen, fp.tri, fp.k, fp.eta);
    else
        hipLaunchKernelGGL((k_fill<double, c128>), dim3(grid), dim3(kFillThreads), 0, s, jobs, njobs, tile_prefix,
                           total_tiles, panels, (const double*)fp.rec, fp.cen, fp.tri, fp.k, fp.eta);
}

template <typename T, typename PC>
__global__ __launch_bounds__(kFillThreads) void k_fill_panel(PC* __restrict__ Z, int ld,
                                                             const int32_t* __restrict__ rows, int nr,
                                                             const int32_t* __restrict__ cols, int nc,
                                                             const T* __restrict__ rec,
                                                             const double* __restrict__ cen, TriGeom tg, double k,
                                                             double eta) {
    __shared__ typename FillLdsOf<T>::type lds;
    fill_any<T, PC, true>(Z, ld, rows, nr, cols, nc, blockIdx.x * kTileRows, blockIdx.y * kTileCols, nullptr, rec,
                          cen, tg, k, eta, lds, kNoTail);
}

void launch_fill_panel(void* Z, int ld, const int32_t* rows, int nr, const int32_t* cols, int nc,
                       const FillParams& fp, hipStream_t s) {
    if (nr <= 0 || nc <= 0) return;
    const dim3 grid(cdiv(nr, kTileRows), cdiv(nc, kTileCols));
    if (fp.f32)
        hipLaunchKernelGGL((k_fill_panel<float, c64>), grid, dim3(kFillThreads), 0, s, (c64*)Z, ld, rows, nr, cols,
                           nc, (const float*)fp.rec, fp.cen, fp.tri, fp.k, fp.eta);
    else
        hipLaunchKernelGGL((k_fill_panel<double, c128>), grid, dim3(kFillThreads), 0, s, (c128*)Z, ld, rows, nr,
                           cols, nc, (const double*)fp.rec, fp.cen, fp.tri, fp.k, fp.eta);
}

// ---------------------------------------------------------------- pinv
struct HeldErr {
    double plain;  // ||H - C[hpos, :] U R||_F / ||H||_F
    double est;    // ratio estimate of ||Z - C U R||_F / ||Z||_F over the whole block
};
template <typename PC>
__device__ __forceinline__ HeldErr heldout_error(const BlockJob& J, double* sP, int capd, double* s_red,
                                                 int psam_known);

// Element idx (column-major, r x r) of the core W: the W panel, or the rows
// rpos of C (the same entries, filled once).
__device__ __forceinline__ c128 load_core(const BlockJob& J, int idx, int f32) {
    if (!J.rpos) return J.W[idx];
    const int i = idx % J.r, j = idx / J.r;
    const size_t at = (size_t)j * J.m + J.rpos[i];
    return f32 ? to_c128(((const c64*)J.C)[at]) : ((const c128*)J.C)[at];
}

// LDS variant: W (then X = R^H), the accumulated rotations and the small
// per-column arrays of one problem in dynamic LDS sized for the largest r of
// the launch (r <= kPinvLdsMaxR); one workgroup per problem. The pinv is
// preconditioned by QR with column pivoting (pinv_qr_lds); the reflectors go to
// the problem's W panel in HBM. Team width kT and cached elements kE follow
// from that r and the launch's thread count (pinv_lds_shape).
struct PinvLds {
    c128 *a, *b, *alpha;
    double *w, *norm, *tau;
    int *phys, *flag;
    __device__ PinvLds(char* smem, int rmax, bool one_buf) {
        a = (c128*)smem;
        b = one_buf ? nullptr : a + (size_t)rmax * rmax;
        alpha = a + (size_t)rmax * rmax * (one_buf ? 1 : 2);
        w = (double*)(alpha + rmax);
        norm = w + rmax;     // 2 x rmax
        tau = norm + 2 * rmax;
        phys = (int*)(tau + rmax);  // 2 x rmax
        flag = phys + 2 * rmax;
    }
};

// Two (one_buf: one) r x r complex buffers and the small arrays.
size_t pinv_lds_bytes(int rmax, bool one_buf) {
    return (size_t)(one_buf ? 16 : 32) * rmax * rmax + (size_t)16 * rmax + (size_t)8 * 4 * rmax +
           (size_t)4 * 2 * rmax + 16;
}

// PC: the panels' precision (one kernel per precision, so that each carries
// one held-out check: fewer registers live across the kernel).
template <int kT, int kE, typename PC>
// heldout: the workgroup also computes the held-out error of each job with
// h > 0 (it has just written U; R, H and C are filled), reusing the core's
// LDS buffer; the residual launch then skips those jobs. qr_core: jobs with
// h > 0 in the one-buffer form take the truncated QR pseudo-inverse (2: every
// job of the one-buffer form).
__global__ void k_pinv_lds(const BlockJob* __restrict__ jobs, const int32_t* __restrict__ list, int nlist, int rmax,
                           double atol, double rtol, double label_eps, int one_buf, int heldout, int qr_core,
                           JobOut* __restrict__ out) {
    constexpr int f32 = std::is_same<PC, c64>::value ? 1 : 0;
    extern __shared__ __align__(16) char smem[];
    const PinvLds L(smem, rmax, one_buf != 0);
    for (int li = blockIdx.x; li < nlist; li += gridDim.x) {
        const int jb = list[li];
        const BlockJob& J = jobs[jb];
        const int r = J.r, ld = r;
        for (int idx = threadIdx.x; idx < r * r; idx += blockDim.x) L.a[idx] = load_core(J, idx, f32);
        int kept, reps, sweeps;
        double smax;
        // Relative thresholds of 1e-8 and above keep the weak components'
        // error of the V-free form (eps sigma_max / sigma_j) negligible.
        // one_buf (set by the host only with rtol >= 1e-8 and every core
        // gathered from C) implies the V-free form.
        const bool vfree = rtol >= 1e-8 || one_buf;
        // element (k, c) of W: row position rpos[k] of C's column c, or the W panel
        const void* Cg = J.C;
        const int32_t* rp = J.rpos;
        const c128* Wp = J.W;
        const int m = J.m;
        auto wload = [=](int k, int c) -> c128 {
            if (!rp) return Wp[(size_t)c * r + k];
            const size_t at = (size_t)c * m + rp[k];
            return f32 ? to_c128(((const c64*)Cg)[at]) : ((const c128*)Cg)[at];
        };
        if (qr_core && one_buf && (J.h > 0 || qr_core == 2)) {
            // learned rounds (qr_core 2: every core, the autotuner's probes):
            // the truncated QR pseudo-inverse (no SVD)
            pinv_qr_trunc<kT>(L.a, r, ld, atol, rtol, J.U, r, J.sigma, L.phys, L.norm, L.tau, L.alpha, &kept, &smax,
                              label_eps, &reps);
            sweeps = 0;
        } else {
            pinv_qr_lds<kT, kE>(L.a, L.b, r, ld, atol, rtol, vfree, wload, J.U, r, J.sigma, J.W, L.flag, L.w,
                                L.phys, L.norm, L.tau, L.alpha, &kept, &smax, label_eps, &reps, &sweeps);
        }
        if (threadIdx.x == 0) {
            out[jb].kept = kept;
            out[jb].sigma_max = smax;
            out[jb].rank_eps = reps;
            out[jb].sweeps = sweeps;
        }
        if (heldout && J.h > 0) {
            // the core buffer (rmax^2 complex, rmax >= 16) holds the 16 x 2r
            // interpolation tile of the held-out rows
            const int capd = 2 * rmax * rmax;
            // with the QR core, W U = Q_k Q_k^H: the sampled rows' interpolation
            // norms sum to k
            const int pk = qr_core && one_buf ? kept : -1;  // (jobs with h > 0 took the QR core)
            const HeldErr e = heldout_error<PC>(J, (double*)L.a, capd, L.norm, pk);
            if (threadIdx.x == 0) {
                out[jb].err = e.plain;
                out[jb].err_est = e.est;
            }
        }
        __syncthreads();
    }
}

// Global-memory variant (any r): A = the W panel, V = the job's scratch; the
// weights stay in LDS up to r = 1024, beyond that after V (pinv_v_elems).
template <int kT>
__device__ void pinv_global_one(const BlockJob& J, int jb, double atol, double rtol, double label_eps, double* s_w,
                                int* s_flag, JobOut* __restrict__ out) {
    const int r = J.r;
    double* w = r <= 1024 ? s_w : (double*)(J.V + (size_t)r * r);
    int kept, reps, sweeps;
    double smax;
    jacobi_pinv_t<kT, 0>(J.W, J.V, r, atol, rtol, J.U, r, J.sigma, s_flag, w, &kept, &smax, label_eps, &reps,
                         &sweeps);
    if (threadIdx.x == 0) {
        out[jb].kept = kept;
        out[jb].sigma_max = smax;
        out[jb].rank_eps = reps;
        out[jb].sweeps = sweeps;
    }
}

__global__ void k_pinv_global(const BlockJob* __restrict__ jobs, const int32_t* __restrict__ list, int nlist,
                              double atol, double rtol, double label_eps, int f32, JobOut* __restrict__ out) {
    __shared__ double s_w[1024];
    __shared__ int s_flag;
    for (int li = blockIdx.x; li < nlist; li += gridDim.x) {
        const int jb = list[li];
        const BlockJob& J = jobs[jb];
        if (J.rpos) {  // the Jacobi works in place on the W panel
            for (int idx = threadIdx.x; idx < J.r * J.r; idx += blockDim.x) J.W[idx] = load_core(J, idx, f32);
            __syncthreads();
        }
        switch (jacobi_team(J.r, blockDim.x)) {
            case 1: pinv_global_one<1>(J, jb, atol, rtol, label_eps, s_w, &s_flag, out); break;
            case 2: pinv_global_one<2>(J, jb, atol, rtol, label_eps, s_w, &s_flag, out); break;
            case 4: pinv_global_one<4>(J, jb, atol, rtol, label_eps, s_w, &s_flag, out); break;
            case 8: pinv_global_one<8>(J, jb, atol, rtol, label_eps, s_w, &s_flag, out); break;
            case 16: pinv_global_one<16>(J, jb, atol, rtol, label_eps, s_w, &s_flag, out); break;
            case 32: pinv_global_one<32>(J, jb, atol, rtol, label_eps, s_w, &s_flag, out); break;
            default: pinv_global_one<64>(J, jb, atol, rtol, label_eps, s_w, &s_flag, out); break;
        }
        __syncthreads();
    }
}

// (T, E) of the LDS variant for class size rmax on `threads` threads: T the
// team width (at most 32), E the elements per lane.
void pinv_lds_shape(int rmax, int threads, int* T, int* E) {
    *T = std::min(jacobi_team(rmax, threads), 32);
    *E = (rmax + *T - 1) / *T;
}

namespace {

template <int kT, int kE, typename PC>
void launch_lds_pc(int grid, int threads, size_t lds, hipStream_t s, const BlockJob* jobs, const int32_t* list,
                   int nlist, int rmax, double atol, double rtol, double label_eps, int one_buf, int heldout,
                   int qr_core, JobOut* out) {
    static bool attr_set = false;
    if (!attr_set) {
        (void)hipFuncSetAttribute(
            (const void*)k_pinv_lds<kT, kE, PC>, hipFuncAttributeMaxDynamicSharedMemorySize,
            (int)std::max(pinv_lds_bytes(kPinvLdsMaxR, false), pinv_lds_bytes(kPinvLdsOneBufMaxR, true)));
        attr_set = true;
    }
    hipLaunchKernelGGL((k_pinv_lds<kT, kE, PC>), dim3(grid), dim3(threads), lds, s, jobs, list, nlist, rmax, atol,
                       rtol, label_eps, one_buf, heldout, qr_core, out);
}
template <int kT, int kE>
void launch_lds(int grid, int threads, size_t lds, hipStream_t s, const BlockJob* jobs, const int32_t* list,
                int nlist, int rmax, double atol, double rtol, double label_eps, int f32, int one_buf, int heldout,
                int qr_core, JobOut* out) {
    if (f32)
        launch_lds_pc<kT, kE, c64>(grid, threads, lds, s, jobs, list, nlist, rmax, atol, rtol, label_eps, one_buf,
                                   heldout, qr_core, out);
    else
        launch_lds_pc<kT, kE, c128>(grid, threads, lds, s, jobs, list, nlist, rmax, atol, rtol, label_eps, one_buf,
                                    heldout, qr_core, out);
}

// E rounded up to a compiled register count (1, 2, 3, 4, 6, 8, 16); past 16 elements the team
// streams its columns from LDS (kE = 0).
template <int kT>
void launch_lds_t(int E, int grid, int threads, size_t lds, hipStream_t s, const BlockJob* jobs,
                  const int32_t* list, int nlist, int rmax, double atol, double rtol, double label_eps, int f32,
                  int one_buf, int heldout, int qr_core, JobOut* out) {
    if (E <= 1) launch_lds<kT, 1>(grid, threads, lds, s, jobs, list, nlist, rmax, atol, rtol, label_eps, f32, one_buf, heldout, qr_core, out);
    else if (E <= 2) launch_lds<kT, 2>(grid, threads, lds, s, jobs, list, nlist, rmax, atol, rtol, label_eps, f32, one_buf, heldout, qr_core, out);
    else if (E <= 3) launch_lds<kT, 3>(grid, threads, lds, s, jobs, list, nlist, rmax, atol, rtol, label_eps, f32, one_buf, heldout, qr_core, out);
    else if (E <= 4) launch_lds<kT, 4>(grid, threads, lds, s, jobs, list, nlist, rmax, atol, rtol, label_eps, f32, one_buf, heldout, qr_core, out);
    else if (E <= 6) launch_lds<kT, 6>(grid, threads, lds, s, jobs, list, nlist, rmax, atol, rtol, label_eps, f32, one_buf, heldout, qr_core, out);
    else if (E <= 8) launch_lds<kT, 8>(grid, threads, lds, s, jobs, list, nlist, rmax, atol, rtol, label_eps, f32, one_buf, heldout, qr_core, out);
    else if (E <= 16) launch_lds<kT, 16>(grid, threads, lds, s, jobs, list, nlist, rmax, atol, rtol, label_eps, f32, one_buf, heldout, qr_core, out);
    else launch_lds<kT, 0>(grid, threads, lds, s, jobs, list, nlist, rmax, atol, rtol, label_eps, f32, one_buf, heldout, qr_core, out);
}

}  // namespace

void launch_pinv(const BlockJob* jobs, const int32_t* list, int nlist, double atol, double rtol, double label_eps,
                 int rmax, int threads, int f32, int one_buf, int heldout, int qr_core, JobOut* out, hipStream_t s) {
    if (nlist <= 0) return;
    const int grid = nlist < 16384 ? nlist : 16384;
    if (rmax > 0 && rmax <= pinv_lds_limit(one_buf, rtol)) {
        int T, E;
        pinv_lds_shape(rmax, threads, &T, &E);
        one_buf = one_buf && rtol >= 1e-8;
        const size_t lds = pinv_lds_bytes(rmax, one_buf != 0);
        switch (T) {
            case 1:
            case 2: launch_lds_t<2>(E, grid, threads, lds, s, jobs, list, nlist, rmax, atol, rtol, label_eps, f32, one_buf, heldout, qr_core, out); break;
            case 4: launch_lds_t<4>(E, grid, threads, lds, s, jobs, list, nlist, rmax, atol, rtol, label_eps, f32, one_buf, heldout, qr_core, out); break;
            case 8: launch_lds_t<8>(E, grid, threads, lds, s, jobs, list, nlist, rmax, atol, rtol, label_eps, f32, one_buf, heldout, qr_core, out); break;
            case 16: launch_lds_t<16>(E, grid, threads, lds, s, jobs, list, nlist, rmax, atol, rtol, label_eps, f32, one_buf, heldout, qr_core, out); break;
            default: launch_lds_t<32>(E, grid, threads, lds, s, jobs, list, nlist, rmax, atol, rtol, label_eps, f32, one_buf, heldout, qr_core, out); break;
        }
    } else {
        hipLaunchKernelGGL(k_pinv_global, dim3(grid), dim3(threads), 0, s, jobs, list, nlist, atol, rtol, label_eps,
                           f32, out);
    }
}

// ---------------------------------------------------------------- residual
constexpr int kResidMaxR = 512;

typedef double dbl4 __attribute__((ext_vector_type(4)));

// Complex product X = A B (M x K times K x N) as the real product
//   [X_re X_im] = [A_re A_im] [[B_re B_im], [-B_im B_re]]
// on the FP64 matrix cores (v_mfma_f64_16x16x4f64: D(16 x 16) += A(16 x 4) B(4 x 16),
// A[i][k] in lane i + 16 k, B[k][j] in lane j + 16 k, D[i][j] in lane j + 16 (i % 4),
// element i / 4). Real inner and column indices interleave the parts,
// kk = 2 l + p and jj = 2 j + q (p, q: 0 real, 1 imaginary):
//   A''[i][2l + p] = p ? Im A[i][l] : Re A[i][l]
//   B''[2l + p][2j + q] = Re B[l][j] (p = q), Im B[l][j] (p = 0, q = 1), -Im B[l][j] (p = 1, q = 0).
// The waves of the workgroup take the 16 x 16 output tiles in turn and hand
// each to epi(i0, jj0, acc): lane t, element v holds X''[i0 + t / 16 + 4 v][jj0 + t % 16].
// Operands outside M x K and K x N read as zero, so padded outputs are zero.
// fa(i, l), fb(l, j): the complex entries. Every thread of the workgroup calls.
template <typename FA, typename FB, typename Epi>
__device__ __forceinline__ void cgemm_mfma(int M, int N, int K, const FA& fa, const FB& fb, const Epi& epi) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
    const int mt = (M + 15) >> 4, ntl = (2 * N + 15) >> 4;
    const int li = lane & 15, lk = lane >> 4;
    for (int t = wave; t < mt * ntl; t += nwaves) {
        const int i0 = (t / ntl) << 4, jj0 = (t % ntl) << 4;
        const int i = i0 + li, jj = jj0 + li, j = jj >> 1, q = jj & 1;
        dbl4 acc = {0.0, 0.0, 0.0, 0.0};
        // four k-steps per iteration: their operand loads are all in flight
        // before the first product
        for (int k0 = 0; k0 < 2 * K; k0 += 16) {
            double a[4], b[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int kk = k0 + 4 * u + lk, l = kk >> 1, p = kk & 1;
                a[u] = b[u] = 0.0;
                if (l < K) {
                    if (i < M) {
                        const c128 x = fa(i, l);
                        a[u] = p ? x.im : x.re;
                    }
                    if (j < N) {
                        const c128 y = fb(l, j);
                        b[u] = p == q ? y.re : (q ? y.im : -y.im);
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b[u], acc, 0, 0, 0);
        }
        epi(i0, jj0, acc);
    }
}

// sum over cnt rows pos(i) of |C[pos(i), :] U|^2 on the FP64 matrix cores, in
// the real form of cgemm_mfma. U is read from sUt (r x r, row-major: sUt[l r + j]
// = U[l][j], LDS or global). A wave takes a 16-row tile and loads its C
// operands once per k-step for up to four column tiles (64 real columns).
template <typename PC, typename Pos>
__device__ __forceinline__ double pnorm_mfma(const PC* __restrict__ C, int m, int r, int cnt, const Pos& pos,
                                             const c128* sUt) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int ncg = (2 * r + 31) >> 5;  // column groups of two 16-column tiles
    const double* ud = (const double*)sUt;
    double s = 0;
    for (int t = wave; t < ((cnt + 15) >> 4) * ncg; t += nwaves) {
        const int rt = t / ncg, cg = t - rt * ncg;
        const int i = (rt << 4) + li;
        const int row = i < cnt ? pos(i) : -1;
        dbl4 acc[2];
#pragma unroll
        for (int c = 0; c < 2; ++c) acc[c] = dbl4{0.0, 0.0, 0.0, 0.0};
        // the C entries of the next k-step are loaded while this one's products
        // run; a lane's real index kk = k0 + 4 u + lk has the parity of lk, so
        // it needs one part of each entry
        using Real = typename ComplexTraits<PC>::real;
        const Real* Cp = (const Real*)C + (lk & 1);
        Real xn[4];
        auto fetch = [&](int k0) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int l = (k0 + 4 * u + lk) >> 1;
                xn[u] = 0;
                if (row >= 0 && l < r) xn[u] = Cp[2 * ((size_t)l * m + row)];
            }
        };
        fetch(0);
        for (int k0 = 0; k0 < 2 * r; k0 += 16) {
            double a[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) a[u] = (double)xn[u];
            if (k0 + 16 < 2 * r) fetch(k0 + 16);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int kk = k0 + 4 * u + lk, l = kk >> 1, p = kk & 1;
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    const int jj = (cg << 5) + (c << 4) + li, j = jj >> 1, q = jj & 1;
                    double b = 0.0;
                    if (l < r && j < r) {
                        // component p ^ q of U[l][j], negated for (p, q) = (1, 0)
                        b = ud[2 * ((size_t)l * r + j) + (p ^ q)];
                        if (p && !q) b = -b;
                    }
                    acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b, acc[c], 0, 0, 0);
                }
            }
        }
#pragma unroll
        for (int c = 0; c < 2; ++c) s += acc[c][0] * acc[c][0] + acc[c][1] * acc[c][1] + acc[c][2] * acc[c][2] +
                                         acc[c][3] * acc[c][3];
    }
    return s;
}

// Held-out comparison of g rows: Y R (Y = sY, g x r complex, real-interleaved
// with row stride ldp) against H[c0 .. c0 + g, :], accumulating |H - Y R|^2
// and |H|^2 in the real form of cgemm_mfma. With sRw (kZSlot doubles per
// wave in LDS) a wave stages each 8 x 8 complex block of R (eight 64-byte
// column segments, one entry per lane) once for its four k-steps, instead of
// every lane loading its own entry of R for every product; without, R is read
// from the panel. A wave owns its slot (LDS is in order within a wave).
constexpr int kZSlot = 2 * 8 * 9;  // 8 columns x 8 rows of c128, column stride 9 (bank spread)
template <typename PC>
__device__ __forceinline__ void zpass_mfma(int g, int n, int r, const double* sY, int ldp, const PC* __restrict__ R,
                                           const PC* __restrict__ H, int h, int c0, double* sRw, double& num,
                                           double& den) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int ntl = (2 * n + 15) >> 4;
    c128* slot = sRw ? (c128*)(sRw + (size_t)wave * kZSlot) : nullptr;
    for (int t = wave; t < ntl; t += nwaves) {
        const int jj0 = t << 4, j0 = jj0 >> 1;
        const int jl = li >> 1, q = li & 1, j = j0 + jl;
        dbl4 acc = {0.0, 0.0, 0.0, 0.0};
        // the next block of R is loaded while this one's products run
        const int sj = lane >> 3, sl = lane & 7;
        PC xn{};
        if (slot && j0 + sj < n && sl < r) xn = R[(size_t)(j0 + sj) * r + sl];
        for (int l0 = 0; l0 < r; l0 += 8) {
            if (slot) {
                slot[sj * 9 + sl] = to_c128(xn);
                xn = PC{};
                if (j0 + sj < n && l0 + 8 + sl < r) xn = R[(size_t)(j0 + sj) * r + l0 + 8 + sl];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int kk = 4 * u + lk, ll = kk >> 1, p = kk & 1, l = l0 + ll;
                double a = 0.0, b = 0.0;
                if (l < r) {
                    if (li < g) a = sY[li * ldp + 2 * l + p];
                    if (j < n) {
                        const c128 y = slot ? slot[jl * 9 + ll] : to_c128(R[(size_t)j * r + l]);
                        b = p == q ? y.re : (q ? y.im : -y.im);
                    }
                }
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
            }
        }
        if (j < n) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = lk + 4 * v;
                if (i < g) {
                    const c128 hv = to_c128(H[(size_t)j * h + c0 + i]);
                    const double x = q ? hv.im : hv.re, d = x - acc[v];
                    num += d * d;
                    den += x * x;
                }
            }
        }
    }
}

// Relative Frobenius error of the CUR on the held-out rows,
// ||H - C[hpos, :] U R||_F / ||H||_F, and a ratio estimate of the whole
// block's error, all products on the FP64 matrix cores (cgemm_mfma). The
// held-out rows go 16 at a time (one tile row): Y = C[rows, :] U (their
// interpolation coefficients P_i) into sP (16 x 2r doubles, capacity capd),
// then Y R against H[rows, :]. The buffer first holds U row-major (2 r^2
// doubles) for the interpolation norms and, when both fit, for Y.
//
// The estimate: a row's error |Z_i - P_i R| grows with |P_i| (the
// interpolation from the sampled rows amplifies what they miss), so the
// held-out rows' errors are scaled by the interpolation norms of every
// unsampled row (a ratio estimator with auxiliary variable |P_i|^2):
//   est^2 = (sum_h |E_h|^2 / sum_h |P_h|^2) sum_{i unsampled} |P_i|^2
//           / (|R|_F^2 + (m - r) / h |H|_F^2).
// Uniform held-out rows alone miss blocks whose error sits in a few rows;
// those rows have the largest |P_i| and enter the estimate through it.
// psam_known >= 0: the sampled rows' interpolation norms sum to it (the QR
// core's W U is a projector of that rank), and their pass is skipped.
template <typename PC>
__device__ __forceinline__ HeldErr heldout_error(const BlockJob& J, double* sP, int capd, double* s_red,
                                                 int psam_known) {
    const PC* C = (const PC*)J.C;
    const PC* R = (const PC*)J.R;
    const PC* H = (const PC*)J.H;
    const c128* U = J.U;
    const int m = J.m, n = J.n, r = J.r, h = J.h, tid = threadIdx.x, nt = blockDim.x;
    const int lane = tid & 63;
    // tile row stride: odd in doubles (rows on distinct LDS banks) when it fits
    const int ldp = 16 * (2 * r + 1) <= capd ? 2 * r + 1 : 2 * r;
    auto sumsq = [](const dbl4& a) { return a[0] * a[0] + a[1] * a[1] + a[2] * a[2] + a[3] * a[3]; };
    // U row-major into the buffer
    c128* sUt = (c128*)sP;
    for (int idx = tid; idx < r * r; idx += nt) {
        const int j = idx / r, l = idx - j * r;
        sUt[(size_t)l * r + j] = U[idx];
    }
    __syncthreads();
    // interpolation norms of every row and of the sampled rows (their
    // difference: the unsampled rows)
    const double pall = pnorm_mfma(C, m, r, m, [](int i) { return i; }, sUt);
    const double psam = psam_known >= 0 ? (threadIdx.x == 0 ? (double)psam_known : 0.0)
                                        : pnorm_mfma(C, m, r, r, [&](int i) { return J.rpos[i]; }, sUt);
    // Y's tile after U when both fit, else over it (U then read from the panel)
    const bool keep_u = 2 * r * r + 16 * ldp <= capd;
    double* sY = keep_u ? sP + 2 * r * r : sP;
    if (!keep_u) __syncthreads();
    auto fu = [&](int l, int j) { return keep_u ? sUt[(size_t)l * r + j] : U[(size_t)j * r + l]; };
    // R staging slots after Y's tile, when they fit
    const size_t zoff = (size_t)(sY - sP) + 16 * (size_t)ldp;
    double* sRw = zoff + (size_t)(blockDim.x >> 6) * kZSlot <= (size_t)capd ? sP + zoff : nullptr;
    double num = 0, den = 0, ph = 0;
    for (int c0 = 0; c0 < h; c0 += 16) {
        const int g = min(16, h - c0);
        cgemm_mfma(g, r, r, [&](int i, int l) { return to_c128(C[(size_t)l * m + J.hpos[c0 + i]]); }, fu,
                   [&](int i0, int jj0, const dbl4& a) {
                       const int jj = jj0 + (lane & 15);
#pragma unroll
                       for (int v = 0; v < 4; ++v) {
                           const int i = i0 + (lane >> 4) + 4 * v;
                           if (i < g && jj < 2 * r) sY[i * ldp + jj] = a[v];
                       }
                       ph += sumsq(a);
                   });
        __syncthreads();
        zpass_mfma<PC>(g, n, r, sY, ldp, R, H, h, c0, sRw, num, den);
        __syncthreads();
    }
    // |R|_F^2: the sampled rows' norms
    double rr = 0;
    for (int idx = tid; idx < r * n; idx += nt) {
        const c128 b = to_c128(R[idx]);
        rr += b.re * b.re + b.im * b.im;
    }
    num = block_sum(num, s_red);
    den = block_sum(den, s_red);
    ph = block_sum(ph, s_red);
    const double pun = fmax(block_sum(pall, s_red) - block_sum(psam, s_red), 0.0);
    rr = block_sum(rr, s_red);
    HeldErr e;
    e.plain = den > 0 ? sqrt(num / den) : (num > 0 ? INFINITY : 0.0);
    const double ztot = rr + (double)(m - r) / h * den;
    if (num == 0)
        e.est = 0.0;
    else if (ph > 0 && ztot > 0)
        e.est = sqrt(num / ph * pun / ztot);
    else
        e.est = INFINITY;
    return e;
}

// skip_heldout_rmax: the pinv launches computed the held-out error of the jobs
// with h > 0 and r <= skip_heldout_rmax (their LDS variant); those are skipped here.
template <typename PC>
__global__ void k_residual(const BlockJob* __restrict__ jobs, int njobs, int skip_heldout_rmax,
                           JobOut* __restrict__ out) {
    __shared__ c128 s_t1[kResidMaxR];
    __shared__ c128 s_t2[kResidMaxR];
    __shared__ c128 s_part[1024];
    __shared__ double s_red[16];
    for (int jb = blockIdx.x; jb < njobs; jb += gridDim.x) {
        const BlockJob& J = jobs[jb];
        if (J.h > 0 && J.r <= skip_heldout_rmax) continue;
        if (J.h == 0 && J.sample_mode == 3) continue;  // learned, every row sampled: stored exactly
        if (J.h > 0) {
            // The 16 x 2r interpolation tile goes to s_t1 while it fits, else
            // to the W panel (2 r^2 doubles >= 16 (2r + 1) for r >= 17), which
            // the pinv has already consumed.
            const HeldErr e = 16 * (2 * J.r + 1) <= 2 * kResidMaxR
                                  ? heldout_error<PC>(J, (double*)s_t1, 2 * kResidMaxR, s_red, -1)
                                  : heldout_error<PC>(J, (double*)J.W, 2 * J.r * J.r, s_red, -1);
            if (threadIdx.x == 0) {
                out[jb].err = e.plain;
                out[jb].err_est = e.est;
            }
            __syncthreads();
            continue;
        }
        // Beyond kResidMaxR the two r-vectors use the W panel (r*r >= 2r),
        // which the pinv has already consumed.
        const double err =
            J.r <= kResidMaxR
                ? cur_residual<PC>((const PC*)J.C, J.m, J.r, J.U, (const PC*)J.R, J.n, J.v, J.y_old, J.y_new, s_t1,
                                   s_t2, s_part, s_red)
                : cur_residual<PC>((const PC*)J.C, J.m, J.r, J.U, (const PC*)J.R, J.n, J.v, J.y_old, J.y_new, J.W,
                                   J.W + J.r, s_part, s_red);
        if (threadIdx.x == 0) out[jb].err = err;
        __syncthreads();
    }
}

void launch_residual(const BlockJob* jobs, int njobs, int f32, int threads, int skip_heldout_rmax, JobOut* out,
                     hipStream_t s) {
    if (njobs <= 0) return;
    const int grid = njobs < 65536 ? njobs : 65536;  // one workgroup per block up to 65536
    if (f32)
        hipLaunchKernelGGL(k_residual<c64>, dim3(grid), dim3(threads), 0, s, jobs, njobs, skip_heldout_rmax, out);
    else
        hipLaunchKernelGGL(k_residual<c128>, dim3(grid), dim3(threads), 0, s, jobs, njobs, skip_heldout_rmax, out);
}

// ---------------------------------------------------------------- ACA
// Partially pivoted adaptive cross approximation of one block per workgroup
// (ACA.m:40-128, the reference's comparator), with the FP64 element filling
// one row or one column of the block per step.

// Workgroup argmax of v over (value, index) pairs: largest value, smallest
// index among ties (ACA.m's find(abs(R) == max(abs(R))) then the first).
__device__ void block_argmax(double v, int i, double* s_v, int* s_i, double& bv, int& bi) {
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const double ov = __shfl_xor(v, o);
        const int oi = __shfl_xor(i, o);
        if (ov > v || (ov == v && oi < i)) {
            v = ov;
            i = oi;
        }
    }
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    __syncthreads();
    if (lane == 0) {
        s_v[w] = v;
        s_i[w] = i;
    }
    __syncthreads();
    bv = s_v[0];
    bi = s_i[0];
    for (int k = 1; k < (int)blockDim.x / kWave; ++k)
        if (s_v[k] > bv || (s_v[k] == bv && s_i[k] < bi)) {
            bv = s_v[k];
            bi = s_i[k];
        }
}

__global__ __launch_bounds__(kFillThreads) void k_aca(const AcaJob* __restrict__ jobs, double tol,
                                                      const double* __restrict__ rec, double k, double eta) {
    __shared__ double s_v[kFillThreads / kWave], s_red[kFillThreads / kWave];
    __shared__ int s_i[kFillThreads / kWave];
    const AcaJob J = jobs[blockIdx.x];
    const int M = J.m, N = J.n, tid = threadIdx.x, nt = blockDim.x;
    const int lane = tid & (kWave - 1), w = tid / kWave, nw = nt / kWave;
    const FillConsts fc = fill_consts_f64(k, eta);
    int32_t* used_r = J.used;
    int32_t* used_c = J.used + M;
    for (int i = tid; i < M + N; i += nt) J.used[i] = 0;
    __syncthreads();
    // one entry per thread, the row RWG as the test function (as the tiles'
    // lane records): a vector keeps every thread busy without LDS staging
    auto fill_row = [&](int i, c128* out) {  // out[j] = Z[i, j]
        RwgRec<double> a;
        load_rec(rec, J.rows[i], a);
        for (int j = tid; j < N; j += nt) {
            double zr, zi;
            entry_f64(a, rec + (size_t)J.cols[j] * kRecStride, fc.k, fc.c4k2, fc.scale, zr, zi);
            out[j] = c128{zr, zi};
        }
    };
    auto fill_col = [&](int j, c128* out) {  // out[i] = Z[i, j]
        const double* b = rec + (size_t)J.cols[j] * kRecStride;
        for (int i = tid; i < M; i += nt) {
            RwgRec<double> a;
            load_rec(rec, J.rows[i], a);
            double zr, zi;
            entry_f64(a, b, fc.k, fc.c4k2, fc.scale, zr, zi);
            out[i] = c128{zr, zi};
        }
    };
    if (M == 1 || N == 1) {  // a vector: nothing to compress (ACA.m:47)
        if (N == 1) {
            fill_col(0, J.U);
            if (tid == 0) J.Vt[0] = c128{1.0, 0.0};
        } else {
            fill_row(0, J.Vt);
            if (tid == 0) J.U[0] = c128{1.0, 0.0};
        }
        if (tid == 0) {
            J.I[0] = 0;
            J.J[0] = 0;
            *J.k_out = 1;
            *J.status = 1;
        }
        return;
    }
    const int kend = min(min(M, N), J.kmax);
    int Ik = 0, kk = 0, status = 0;
    if (tid == 0) used_r[0] = 1;
    double normZ = 0;
    while (true) {
        // residual row Ik
        fill_row(Ik, J.row);
        __syncthreads();
        for (int j = tid; j < N; j += nt) {
            c128 x = J.row[j];
            for (int l = 0; l < kk; ++l) {
                const c128 u = J.U[Ik + (size_t)l * M], v = J.Vt[j + (size_t)l * N];
                x.re -= u.re * v.re - u.im * v.im;
                x.im -= u.re * v.im + u.im * v.re;
            }
            J.row[j] = x;
        }
        __syncthreads();
        double bv = -1.0;
        int Jk = 0x7fffffff;
        for (int j = tid; j < N; j += nt)
            if (!used_c[j]) {
                const double a = J.row[j].re * J.row[j].re + J.row[j].im * J.row[j].im;
                if (a > bv) {
                    bv = a;
                    Jk = j;
                }
            }
        block_argmax(bv, Jk, s_v, s_i, bv, Jk);
        if (Jk >= N) {  // no finite residual entry left in the row: stop, status 3
            status = 3;
            break;
        }
        const c128 piv = J.row[Jk];
        if (piv.re == 0.0 && piv.im == 0.0) {  // R(Ik, Jk) == 0 (ACA.m:91)
            status = 1;
            break;
        }
        __syncthreads();
        if (tid == 0) {
            used_c[Jk] = 1;
            J.I[kk] = Ik;
            J.J[kk] = Jk;
        }
        const double ip = 1.0 / (piv.re * piv.re + piv.im * piv.im);
        for (int j = tid; j < N; j += nt) {  // V_k = R_row / R(Ik, Jk)
            const c128 x = J.row[j];
            J.Vt[j + (size_t)kk * N] = c128{(x.re * piv.re + x.im * piv.im) * ip, (x.im * piv.re - x.re * piv.im) * ip};
        }
        // residual column Jk into U_k
        c128* Uk = J.U + (size_t)kk * M;
        fill_col(Jk, Uk);
        __syncthreads();
        for (int i = tid; i < M; i += nt) {
            c128 x = Uk[i];
            for (int l = 0; l < kk; ++l) {
                const c128 u = J.U[i + (size_t)l * M], v = J.Vt[Jk + (size_t)l * N];
                x.re -= u.re * v.re - u.im * v.im;
                x.im -= u.re * v.im + u.im * v.re;
            }
            Uk[i] = x;
        }
        __syncthreads();
        // |U_k|^2 |V_k|^2 and the cross terms sum_l |(U_l^H U_k)(V_k V_l^H)|
        double nu = 0, nv = 0;
        for (int i = tid; i < M; i += nt) nu += Uk[i].re * Uk[i].re + Uk[i].im * Uk[i].im;
        const c128* Vk = J.Vt + (size_t)kk * N;
        for (int j = tid; j < N; j += nt) nv += Vk[j].re * Vk[j].re + Vk[j].im * Vk[j].im;
        nu = block_sum(nu, s_red);
        nv = block_sum(nv, s_red);
        double cross = 0;
        for (int l = w; l < kk; l += nw) {
            const c128* Ul = J.U + (size_t)l * M;
            const c128* Vl = J.Vt + (size_t)l * N;
            double ar = 0, ai = 0, br = 0, bi = 0;
            for (int i = lane; i < M; i += kWave) {  // U_l^H U_k
                ar += Ul[i].re * Uk[i].re + Ul[i].im * Uk[i].im;
                ai += Ul[i].re * Uk[i].im - Ul[i].im * Uk[i].re;
            }
            for (int j = lane; j < N; j += kWave) {  // V_k V_l^H
                br += Vk[j].re * Vl[j].re + Vk[j].im * Vl[j].im;
                bi += Vk[j].im * Vl[j].re - Vk[j].re * Vl[j].im;
            }
            ar = team_sum(ar, kWave);
            ai = team_sum(ai, kWave);
            br = team_sum(br, kWave);
            bi = team_sum(bi, kWave);
            if (lane == 0) cross += sqrt(ar * ar + ai * ai) * sqrt(br * br + bi * bi);
        }
        cross = block_sum(cross, s_red);
        normZ += 2.0 * cross + nu * nv;
        ++kk;
        if (!(nu * nv < INFINITY) || !(normZ < INFINITY)) {  // NaN or Inf in the cross: status 3
            status = 3;
            break;
        }
        if (kk > 1 && sqrt(nu * nv) <= tol * sqrt(normZ)) {  // ACA.m:114
            status = 1;
            break;
        }
        if (kk == min(M, N)) {  // ACA.m:118
            status = 1;
            break;
        }
        if (kk == kend) {  // storage bound reached before convergence
            status = 2;
            break;
        }
        // next row: largest |U_k| among the rows not yet taken (ACA.m:123)
        bv = -1.0;
        int In = 0x7fffffff;
        for (int i = tid; i < M; i += nt)
            if (!used_r[i]) {
                const double a = Uk[i].re * Uk[i].re + Uk[i].im * Uk[i].im;
                if (a > bv) {
                    bv = a;
                    In = i;
                }
            }
        block_argmax(bv, In, s_v, s_i, bv, In);
        if (In >= M) {  // no finite entry left in the column: stop, status 3
            status = 3;
            break;
        }
        __syncthreads();
        Ik = In;
        if (tid == 0) used_r[Ik] = 1;
        __syncthreads();
    }
    if (tid == 0) {
        *J.k_out = kk;
        *J.status = status;
    }
}

void launch_aca(const AcaJob* jobs, int njobs, double tol, const FillParams& fp, hipStream_t s) {
    if (njobs <= 0) return;
    hipLaunchKernelGGL(k_aca, dim3(njobs), dim3(kFillThreads), 0, s, jobs, tol, (const double*)fp.rec, fp.k, fp.eta);
}

// ---------------------------------------------------------------- exact blocks
template <typename PC>
__global__ void k_exact(const ExactJob* __restrict__ jobs, int njobs) {
    for (int b = blockIdx.x; b < njobs; b += gridDim.x) {
        const ExactJob J = jobs[b];
        const int r = J.r;
        for (int idx = threadIdx.x; idx < r * r; idx += blockDim.x)
            J.U[idx] = c128{(idx % r) == (idx / r) ? 1.0 : 0.0, 0.0};
        // side 0: C is m x r with r = m; side 1: R is r x n with r = n
        PC* X = (PC*)(J.side == 0 ? J.C : J.R);
        for (int idx = threadIdx.x; idx < r * r; idx += blockDim.x)
            store_c(&X[idx], (idx % r) == (idx / r) ? 1.0 : 0.0, 0.0);
    }
}

void launch_exact(const ExactJob* jobs, int njobs, int f32, hipStream_t s) {
    if (njobs <= 0) return;
    const int grid = njobs < 4096 ? njobs : 4096;
    if (f32)
        hipLaunchKernelGGL(k_exact<c64>, dim3(grid), dim3(256), 0, s, jobs, njobs);
    else
        hipLaunchKernelGGL(k_exact<c128>, dim3(grid), dim3(256), 0, s, jobs, njobs);
}

// ---------------------------------------------------------------- test vectors
__global__ void k_vtest(c128* __restrict__ v, const int32_t* __restrict__ offsets, int nblocks, uint64_t seed) {
    for (int b = blockIdx.x; b < nblocks; b += gridDim.x) {
        const int o = offsets[b], n = offsets[b + 1] - o;
        for (int j = threadIdx.x; j < n; j += blockDim.x) {
            double re, im;
            vtest_entry(seed, (uint32_t)b, (uint32_t)j, &re, &im);
            v[o + j] = c128{re, im};
        }
    }
}

void launch_vtest(c128* v, const int32_t* offsets, int nblocks, uint64_t seed, hipStream_t s) {
    if (nblocks <= 0) return;
    const int grid = nblocks < 4096 ? nblocks : 4096;
    hipLaunchKernelGGL(k_vtest, dim3(grid), dim3(256), 0, s, v, offsets, nblocks, seed);
}

}  // namespace farcur
