// EFIE / RWG impedance element, evaluated by 3-point Gauss quadrature on the
// test and source triangles, and the LDS-staged panel tile fill built on it.
//
// Replaces the reference's binary-only impedance_matrix_single_element2
// (declared at codes/GPU_implementation/seed/user_impedance_for_julia_cuda.cu:178-214,
// 3 Gauss points per CU:34) as called by impedance_matrix_compute_elbyel3
// (CU:353-439). The formula is the Galerkin EFIE with RWG functions
//   f = s (l / 2A) (r - v_free),  div f = s l / A,  s = +1 on T+, -1 on T-,
//   Z_mn = -j k eta ∫∫ [f_m . f_n - (1/k^2) div f_m div' f_n] e^{-jkR} / (4 pi R),
// which with the rule (2/3,1/6,1/6) x 3, weights 1/3, collapses to
//   Z_mn = -(k eta l_m l_n / (144 pi)) * sum_{a,b,i,j} T_abij (sin kR + j cos kR) / R,
//   T_abij = rho_ai . rho_bj - s_a s_b 4 / k^2,  rho = s (q - v_free),
// the areas cancelling. Sign convention: Z J = -V for V_m = <f_m, E_inc>
// (the reference solves J = -Z \ Ei, codes/GPU_implementation/seed/compres_single_block_spheres.jl:3773).
#pragma once
#include "common.hpp"

namespace farcur {

// Per-RWG record, structure of arrays: field f of RWG e at rec[f * ne + e].
//   f = h*9 + i*3 + c        quadrature point i (xyz c) of triangle T+ (h=0) / T- (h=1)
//   f = 18 + h*9 + i*3 + c   rho_hi = s_h (q_hi - v_h)
//   f = 36                   edge length l
constexpr int kRecFields = 37;

template <typename T>
struct RwgRec {
    T q[2][3][3];
    T p[2][3][3];
    T l;
};

template <typename T>
__device__ __forceinline__ void load_rec(const T* __restrict__ rec, int ne, int e, RwgRec<T>& r) {
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                r.q[h][i][c] = rec[(size_t)(h * 9 + i * 3 + c) * ne + e];
                r.p[h][i][c] = rec[(size_t)(18 + h * 9 + i * 3 + c) * ne + e];
            }
    r.l = rec[(size_t)36 * ne + e];
}

// Green-kernel sample: accumulates T (sin kR, cos kR) / R.
__device__ __forceinline__ void green_acc(double dx, double dy, double dz, double term, double k, double& as,
                                          double& ac) {
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double r = sqrt(r2);
    double s, c;
    sincos(k * r, &s, &c);
    const double w = term / r;
    as += w * s;
    ac += w * c;
}

// FP32: 1/R from the hardware reciprocal square root, and the phase in
// revolutions (R / lambda) reduced to [-1/2, 1/2] so that the hardware sine and
// cosine (which take revolutions) see a small argument.
__device__ __forceinline__ void green_acc(float dx, float dy, float dz, float term, float inv_lambda, float& as,
                                          float& ac) {
    const float r2 = dx * dx + dy * dy + dz * dz;
    const float ir = __builtin_amdgcn_rsqf(r2);
    const float t = r2 * ir * inv_lambda;
    const float f = t - __builtin_rintf(t);
    const float s = __builtin_amdgcn_sinf(f);
    const float c = __builtin_amdgcn_cosf(f);
    const float w = term * ir;
    as += w * s;
    ac += w * c;
}

// Source point (hb, j) against the 6 test points of `a`.
template <typename T>
__device__ __forceinline__ void source_point_sums(const RwgRec<T>& a, const T* __restrict__ b, int hb, int j, T kw,
                                                  T c4k2, T& as, T& ac) {
    const T bx = b[hb * 9 + j * 3 + 0], by = b[hb * 9 + j * 3 + 1], bz = b[hb * 9 + j * 3 + 2];
    const T px = b[18 + hb * 9 + j * 3 + 0], py = b[18 + hb * 9 + j * 3 + 1], pz = b[18 + hb * 9 + j * 3 + 2];
#pragma unroll
    for (int ha = 0; ha < 2; ++ha) {
        const T ss = (ha == hb) ? c4k2 : -c4k2;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const T term = a.p[ha][i][0] * px + a.p[ha][i][1] * py + a.p[ha][i][2] * pz - ss;
            green_acc(a.q[ha][i][0] - bx, a.q[ha][i][1] - by, a.q[ha][i][2] - bz, term, kw, as, ac);
        }
    }
}

// Sum over the 4 triangle pairs x 9 point pairs. `kw` is k for FP64 and
// 1/lambda for FP32 (see green_acc). Returns sum T sin/R and sum T cos/R.
// FP64 keeps the 6 source points as a loop (the FP64 sine/cosine is long and
// a full unroll exhausts the register file); FP32 unrolls completely.
template <typename T>
__device__ __forceinline__ void entry_sums(const RwgRec<T>& a, const T* __restrict__ b, T kw, T c4k2, T& as,
                                           T& ac) {
    as = 0;
    ac = 0;
    if constexpr (sizeof(T) == 8) {
#pragma unroll 1
        for (int s = 0; s < 6; ++s) source_point_sums(a, b, s / 3, s % 3, kw, c4k2, as, ac);
    } else {
#pragma unroll
        for (int s = 0; s < 6; ++s) source_point_sums(a, b, s / 3, s % 3, kw, c4k2, as, ac);
    }
}

// Constants of one fill: k, k*eta/(144 pi), 4/k^2, and the phase factor kw.
template <typename T>
struct FillConsts {
    T kw;     // k (FP64) or 1/lambda = k / (2 pi) (FP32)
    T c4k2;   // 4 / k^2
    T scale;  // k eta / (144 pi)
};

// Tile geometry of the panel fill: a workgroup of 4 waves covers 64 rows x
// (4 * kTileColsPerWave) columns; lane = row, wave = group of columns. The
// column (source) records of the tile are staged in LDS and read as broadcasts.
constexpr int kTileRows = 64;
constexpr int kTileColsPerWave = 4;
constexpr int kTileCols = 4 * kTileColsPerWave;
constexpr int kFillThreads = 256;

// Fills rows [row0, row0+64) x cols [col0, col0+16) of the panel
// out[i + j*ld] = Z(rows[i], cols[j]). Must be called by all 256 threads.
template <typename T, typename OutC>
__device__ __forceinline__ void fill_tile(OutC* __restrict__ out, int ld, const int* __restrict__ rows, int nr,
                                          const int* __restrict__ cols, int nc, int row0, int col0,
                                          const T* __restrict__ rec, int ne, const FillConsts<T>& fc,
                                          T (*s_col)[kRecFields]) {
    const int tid = threadIdx.x;
    __syncthreads();  // s_col may still be read by a previous tile
    for (int idx = tid; idx < kTileCols * kRecFields; idx += kFillThreads) {
        const int c = idx / kRecFields, f = idx - c * kRecFields;
        const int j = col0 + c;
        s_col[c][f] = (j < nc) ? rec[(size_t)f * ne + cols[j]] : T(0);
    }
    __syncthreads();
    const int lane = tid & (kWave - 1);
    const int w = tid / kWave;
    const int i = row0 + lane;
    if (i >= nr) return;
    RwgRec<T> a;
    load_rec(rec, ne, rows[i], a);
#pragma unroll 1
    for (int cc = 0; cc < kTileColsPerWave; ++cc) {
        const int c = w * kTileColsPerWave + cc;
        const int j = col0 + c;
        if (j >= nc) break;
        T as, ac;
        entry_sums<T>(a, s_col[c], fc.kw, fc.c4k2, as, ac);
        const T kap = fc.scale * a.l * s_col[c][36];
        store_c(&out[(size_t)j * ld + i], -kap * as, -kap * ac);
    }
}

}  // namespace farcur
