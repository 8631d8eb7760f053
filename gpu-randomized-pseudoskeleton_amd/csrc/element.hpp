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

// Per-RWG record, array of structures: field f of RWG e at
// rec[e * kRecStride + f] (panel rows and columns are scattered RWG ids, so a
// lane's whole record is a few contiguous cache lines read by vector loads).
// FP64 records:
//   f = h*9 + i*3 + c        quadrature point q_hi (xyz c) of triangle T+ (h=0) / T- (h=1)
//   f = 18 + h*9 + i*3 + c   rho_hi = s_h (q_hi - v_h)
//   f = 36                   edge length l
//   f = 37..39               padding (zero)
// FP32 records hold u_hi = q_hi - c_e (offsets from the RWG centre c_e, the
// mean of its six quadrature points) in place of q_hi; c_e itself is kept in
// FP64 (cen[e * kCenStride + c]) so that centre-to-centre distances keep full
// accuracy.
constexpr int kRecFields = 37;
constexpr int kRecStride = 40;
constexpr int kCenStride = 4;

template <typename T>
struct RwgRec {
    T q[2][3][3];  // q (FP64) or u (FP32)
    T p[2][3][3];
    T l;
};

template <typename T>
__device__ __forceinline__ void load_rec(const T* __restrict__ rec, int e, RwgRec<T>& r) {
    const T* src = (const T*)__builtin_assume_aligned(rec + (size_t)e * kRecStride, 16);
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                r.q[h][i][c] = src[h * 9 + i * 3 + c];
                r.p[h][i][c] = src[18 + h * 9 + i * 3 + c];
            }
    r.l = src[36];
}

// ---------------------------------------------------------------- FP64
// Accumulates T (sin kR, cos kR) / R.
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

// Source point (hb, j) against the 6 test points of `a`.
__device__ __forceinline__ void source_point_sums(const RwgRec<double>& a, const double* __restrict__ b, int hb, int j,
                                                  double k, double c4k2, double& as, double& ac) {
    const double bx = b[hb * 9 + j * 3 + 0], by = b[hb * 9 + j * 3 + 1], bz = b[hb * 9 + j * 3 + 2];
    const double px = b[18 + hb * 9 + j * 3 + 0], py = b[18 + hb * 9 + j * 3 + 1], pz = b[18 + hb * 9 + j * 3 + 2];
#pragma unroll
    for (int ha = 0; ha < 2; ++ha) {
        const double ss = (ha == hb) ? c4k2 : -c4k2;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const double term = a.p[ha][i][0] * px + a.p[ha][i][1] * py + a.p[ha][i][2] * pz - ss;
            green_acc(a.q[ha][i][0] - bx, a.q[ha][i][1] - by, a.q[ha][i][2] - bz, term, k, as, ac);
        }
    }
}

// Z_mn in FP64. The 6 source points stay a loop: the FP64 sine/cosine is long
// and a full unroll exhausts the register file.
__device__ __forceinline__ void entry_f64(const RwgRec<double>& a, const double* __restrict__ b, double k,
                                          double c4k2, double scale, double& zr, double& zi) {
    double as = 0, ac = 0;
#pragma unroll 1
    for (int s = 0; s < 6; ++s) source_point_sums(a, b, s / 3, s % 3, k, c4k2, as, ac);
    const double kap = scale * a.l * b[36];
    zr = -kap * as;
    zi = -kap * ac;
}

// ---------------------------------------------------------------- FP32
// With R0 = |c_m - c_n| and delta = R - R0 for each point pair,
//   e^{-jkR}/R = e^{-jkR0} g,   g = e^{-jk delta} / R,
// and because sum_{a,b} s_a s_b = 0 the charge (div-div) term may use
//   g - 1/R0 = [(cos k delta - 1) - j sin k delta] / R - delta / (R R0)
// instead of g: the near-cancelling second difference of the scalar potential
// is then formed from small quantities only. delta = X / (R + R0) with
// X = R^2 - R0^2 = 2 d.D + |D|^2 (d = c_m - c_n, D = u_a - u_b) is exact to
// rounding. Sines take revolutions (hardware sine), arguments |delta / lambda| < 1.
//
// The point-pair loop is VALU-issue bound, so everything that depends on one
// point only is hoisted to the entry: in half-wavelength units (ihl = 1/(2 lambda)),
//   X ihl = alpha_i + beta_j - 2 u_ai . u'_bj,
//   alpha_i = (2 d.u_ai + |u_ai|^2) ihl,  beta_j = (|u_bj|^2 - 2 d.u_bj) ihl,  u'_bj = u_bj ihl,
// and with th = delta / (2 lambda) (so k delta = 2 pi (2 th) in revolutions):
//   cos(k delta) / R = (1 - 2 sin^2 th) / R,   sin(k delta) / R,
//   -Re(g - 1/R0) = 2 sin^2 th / R + th (2 lambda / R0) / R.
struct F32Consts {
    float inv_lambda;  // k / (2 pi)
    float c4k2;        // 4 / k^2
    float scale;       // k eta / (144 pi)
    double inv_lambda64;
};

__device__ __forceinline__ void entry_f32(const RwgRec<float>& a, const double* __restrict__ ac3,
                                          const float* __restrict__ b, const double* __restrict__ bc3,
                                          const F32Consts& fc, float& zr, float& zi) {
    const double dx64 = ac3[0] - bc3[0], dy64 = ac3[1] - bc3[1], dz64 = ac3[2] - bc3[2];
    const double r064 = sqrt(dx64 * dx64 + dy64 * dy64 + dz64 * dz64);
    const double ph = r064 * fc.inv_lambda64;
    const float f0 = (float)(ph - rint(ph));
    const float s0 = __builtin_amdgcn_sinf(f0), c0 = __builtin_amdgcn_cosf(f0);
    const float dx = (float)dx64, dy = (float)dy64, dz = (float)dz64;
    const float r0 = (float)r064, r0sq = r0 * r0;
    const float ihl = 0.5f * fc.inv_lambda;  // 1 / (2 lambda)
    const float twol = 1.0f / ihl;           // 2 lambda
    const float c2 = twol / r0;              // 2 lambda / R0
    float alpha[2][3], beta[2][3], ub[2][3][3];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const float* q = a.q[h][i];
            const float du = dx * q[0] + dy * q[1] + dz * q[2];
            const float uu = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
            alpha[h][i] = (2.0f * du + uu) * ihl;
            const float* u = b + h * 9 + i * 3;
            const float dv = dx * u[0] + dy * u[1] + dz * u[2];
            const float vv = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
            beta[h][i] = (vv - 2.0f * dv) * ihl;
            ub[h][i][0] = u[0] * ihl;
            ub[h][i][1] = u[1] * ihl;
            ub[h][i][2] = u[2] * ihl;
        }
    float ar = 0, ai = 0;  // ai accumulates -Im
#pragma unroll
    for (int hb = 0; hb < 2; ++hb) {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const float px = b[18 + hb * 9 + j * 3 + 0], py = b[18 + hb * 9 + j * 3 + 1];
            const float pz = b[18 + hb * 9 + j * 3 + 2];
#pragma unroll
            for (int ha = 0; ha < 2; ++ha) {
                const float ss = (ha == hb) ? fc.c4k2 : -fc.c4k2;
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    const float w = a.q[ha][i][0] * ub[hb][j][0] + a.q[ha][i][1] * ub[hb][j][1] +
                                    a.q[ha][i][2] * ub[hb][j][2];
                    const float xl = fmaf(-2.0f, w, alpha[ha][i] + beta[hb][j]);  // X / (2 lambda)
                    const float r2 = fmaf(xl, twol, r0sq);
                    const float ir = __builtin_amdgcn_rsqf(r2);
                    const float th = xl * __builtin_amdgcn_rcpf(fmaf(r2, ir, r0));  // delta / (2 lambda)
                    const float sh = __builtin_amdgcn_sinf(th);
                    const float sn = __builtin_amdgcn_sinf(th + th);
                    const float v = sh * sh * ir;
                    const float gr = fmaf(-2.0f, v, ir);                  // Re g
                    const float hneg = fmaf(th * ir, c2, v + v);          // -Re(g - 1/R0)
                    const float gin = sn * ir;                            // -Im g
                    const float term = a.p[ha][i][0] * px + a.p[ha][i][1] * py + a.p[ha][i][2] * pz;
                    // term * g - ss * (g - 1/R0)
                    ar = fmaf(ss, hneg, fmaf(term, gr, ar));
                    ai = fmaf(term - ss, gin, ai);
                }
            }
        }
    }
    ai = -ai;
    // Z = -j kappa e^{-jkR0} (ar + j ai)
    const float pr = c0 * ar + s0 * ai, pi = c0 * ai - s0 * ar;
    const float kap = fc.scale * a.l * b[36];
    zr = kap * pi;
    zi = -kap * pr;
}

// Tile geometry of the panel fill: a workgroup of 4 waves covers 64 rows x up
// to 64 columns; lane = row, and the waves take the tile's columns round-robin
// (wave w: columns w, w + 4, ...). Sampled panels have r of a few tens of
// columns, so one tile spans all of them and the waves stay balanced (a 16-
// column tile left whole waves idle on the last partial column group). The
// column (source) records of the tile are staged in LDS and read as broadcasts.
constexpr int kTileRows = 64;
constexpr int kTileCols = 64;
constexpr int kFillWaves = 4;
constexpr int kFillThreads = kFillWaves * 64;

template <typename T>
struct FillTileLds {
    T col[kTileCols][kRecFields];
    double cen[kTileCols][3];
};

// One tile of a panel out[i + j*ld] = Z(rows[i], cols[j]). Lanes run over
// 64 consecutive entries of one index set (the lane set), up to 64 entries of
// the other set (the LDS set) are staged in LDS and read as broadcasts.
// kLanesAreRows: lanes over rows (tall panels, C and W); otherwise lanes over
// columns (wide panels R and H), using reciprocity Z(m, n) = Z(n, m) of the
// symmetric Galerkin element. Must be called by all 256 threads.
template <typename T, typename OutC, typename Consts, bool kLanesAreRows>
__device__ __forceinline__ void fill_tile(OutC* __restrict__ out, int ld, const int* __restrict__ rows, int nr,
                                          const int* __restrict__ cols, int nc, int row0, int col0,
                                          const T* __restrict__ rec, const double* __restrict__ cen, int ne,
                                          const Consts& fc, FillTileLds<T>& lds) {
    const int* lane_ids = kLanesAreRows ? rows : cols;
    const int* lds_ids = kLanesAreRows ? cols : rows;
    const int nlane = kLanesAreRows ? nr : nc;
    const int nlds = kLanesAreRows ? nc : nr;
    const int lane0 = kLanesAreRows ? row0 : col0;
    const int lds0 = kLanesAreRows ? col0 : row0;
    const int tid = threadIdx.x;
    const int ncol = min(kTileCols, nlds - lds0);
    __syncthreads();  // the LDS tile may still be read by a previous tile
    for (int idx = tid; idx < ncol * kRecFields; idx += kFillThreads) {
        const int c = idx / kRecFields, f = idx - c * kRecFields;
        lds.col[c][f] = rec[(size_t)lds_ids[lds0 + c] * kRecStride + f];
    }
    if constexpr (sizeof(T) == 4) {
        for (int idx = tid; idx < ncol * 3; idx += kFillThreads) {
            const int c = idx / 3, x = idx - 3 * c;
            lds.cen[c][x] = cen[(size_t)lds_ids[lds0 + c] * kCenStride + x];
        }
    }
    __syncthreads();
    const int lane = tid & (kWave - 1);
    const int w = tid / kWave;
    const int i = lane0 + lane;
    if (i >= nlane) return;
    const int e = lane_ids[i];
    RwgRec<T> a;
    load_rec(rec, e, a);
    double ac3[3] = {0, 0, 0};
    if constexpr (sizeof(T) == 4) {
        ac3[0] = cen[(size_t)e * kCenStride];
        ac3[1] = cen[(size_t)e * kCenStride + 1];
        ac3[2] = cen[(size_t)e * kCenStride + 2];
    }
#pragma unroll 1
    for (int c = w; c < ncol; c += kFillWaves) {
        const int j = lds0 + c;
        T zr, zi;
        if constexpr (sizeof(T) == 8)
            entry_f64(a, lds.col[c], fc.k, fc.c4k2, fc.scale, zr, zi);
        else
            entry_f32(a, ac3, lds.col[c], lds.cen[c], fc, zr, zi);
        if constexpr (kLanesAreRows)
            store_c(&out[(size_t)j * ld + i], zr, zi);
        else
            store_c(&out[(size_t)i * ld + j], zr, zi);
    }
}

// Constants of an FP64 fill.
struct FillConsts {
    double k;      // wavenumber
    double c4k2;   // 4 / k^2
    double scale;  // k eta / (144 pi)
};

}  // namespace farcur
