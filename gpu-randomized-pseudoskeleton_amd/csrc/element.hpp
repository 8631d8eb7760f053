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
// The FP32 fill shares work between RWGs that share a triangle. For a test
// triangle a (of test RWGs m, sign s_a, free vertex v_ma) and a source RWG n
// (six points q_j, rho_nj = s_b (q_j - v_nb) with signs), write the test
// points q_ai = c_a + u_ai (c_a the centroid) and w_ma = v_ma - c_a. Then
//   Z_mn = -j kappa_mn sum_{a in {a+, a-}} s_a [A(a,n) - w_ma . B(a,n) - (4/k^2) C(a,n)],
//   T_i = sum_j rho_nj G_ij,  A = sum_i u_ai . T_i,  B = sum_i T_i,
//   C = sum_{i,j} s_b(j) (G_ij - e^{-jkR0} / R0),
// with kappa_mn = k eta l_m l_n / (144 pi), G_ij = e^{-jkR_ij} / R_ij (the
// areas cancel against the 1/3 weights as in the FP64 element), and R0 =
// |c_a - c_n| (c_n the RWG centre). The subtracted term cancels in C because
// sum_j s_b(j) = 0; it keeps the source-side difference of the charge term
// formed from small quantities. A, B, C of one (triangle, RWG) item serve
// every test RWG on that triangle, so a tile of 64 test RWGs evaluates its
// distinct triangles (about 0.9 per RWG on a mesh patch) instead of four
// triangle pairs per entry.
//
// Per point pair, with delta = R - R0 = X / (R + R0), X = R^2 - R0^2 =
// 2 d.D + |D|^2 (d = c_a - c_n, D = u_ai - u_j, u_j = q_j - c_n) exact to
// rounding, and G = e^{-jkR0} g, g = e^{-jk delta} / R:
//   X ihl = alpha_i + beta_j - 2 ihl u_ai . u_j  (ihl = 1/(2 lambda)),
//   alpha_i = (2 d.u_ai + |u_ai|^2) ihl,  beta_j = (|u_j|^2 - 2 d.u_j) ihl,
//   th = delta / (2 lambda) (sines take revolutions, so k delta = 2 pi (2 th)),
//   Re g = (1 - 2 sin^2 th) / R,  -Im g = sin(2 pi 2 th) / R,
//   -Re(g - 1/R0) = 2 sin^2 th / R + th (2 lambda / R0) / R.
struct F32Consts {
    float inv_lambda;  // k / (2 pi)
    float c4k2;        // 4 / k^2
    float scale;       // k eta / (144 pi)
    double inv_lambda64;
};

// Triangle tables (built with the RWG records):
//   tri[t * kTriStride + i * 3 + c]  u_ti = q_ti - c_t (quadrature points of
//                                    triangle t relative to its centroid)
//   tri[t * kTriStride + 9 + i]      |u_ti|^2
//   tcen[t * kCenStride + c]         c_t (FP64)
//   lrec[e * kLaneStride + h * 3 + c] w_eh = v_eh - c_{t_eh} (free vertex of
//                                    RWG e in its half h, from that centroid)
//   lrec[e * kLaneStride + 6]        l_e
//   rtri[2 e + h]                    t_eh, the triangle of half h (0: T+, 1: T-)
constexpr int kTriStride = 12;
constexpr int kLaneStride = 8;
struct TriGeom {
    const float* tri;
    const double* tcen;
    const float* lrec;
    const int32_t* rtri;
};

// Tile geometry of the panel fill: a workgroup of 4 waves covers 64 RWGs of
// one index set (the lane set: rows for the tall panels C and W, columns for
// the wide panels R and H, using reciprocity Z(m, n) = Z(n, m) of the
// symmetric Galerkin element) times up to 64 RWGs of the other set (the LDS
// set), whose records are staged in LDS and read as broadcasts.
constexpr int kTileRows = 64;
constexpr int kTileCols = 64;
constexpr int kFillWaves = 4;
constexpr int kFillThreads = kFillWaves * 64;

// A second list appended to the LDS set of a wide-panel tile (lanes over
// columns): the held-out rows after the sample rows, so that R and H, which
// share the lane set, fill in one tile. Entry j >= n1 of the LDS set is RWG
// ids[j - n1] and goes to out[i * ld + (j - n1)].
struct LdsTail {
    const int* ids;
    int n;
    void* out;
    int ld;
};

// ---------------------------------------------------------------- FP64 tile
struct FillTileLds {
    double col[kTileCols][kRecFields];
};

// One tile of a panel out[i + j*ld] = Z(rows[i], cols[j]) in FP64: lane =
// entry of the lane set, the waves take the LDS set round-robin. Must be
// called by all 256 threads.
template <typename OutC, bool kLanesAreRows>
__device__ __forceinline__ void fill_tile(OutC* __restrict__ out, int ld, const int* __restrict__ rows, int nr,
                                          const int* __restrict__ cols, int nc, int row0, int col0,
                                          const double* __restrict__ rec, double k, double c4k2, double scale,
                                          FillTileLds& lds, LdsTail tail) {
    const int* lane_ids = kLanesAreRows ? rows : cols;
    const int* lds_ids = kLanesAreRows ? cols : rows;
    const int nlane = kLanesAreRows ? nr : nc;
    const int nl1 = kLanesAreRows ? nc : nr;
    const int nlds = kLanesAreRows ? nl1 : nl1 + tail.n;
    const int lane0 = kLanesAreRows ? row0 : col0;
    const int lds0 = kLanesAreRows ? col0 : row0;
    const int tid = threadIdx.x;
    const int ncol = min(kTileCols, nlds - lds0);
    auto lds_id = [&](int j) {
        if constexpr (kLanesAreRows)
            return lds_ids[j];
        else
            return j < nl1 ? lds_ids[j] : tail.ids[j - nl1];
    };
    __syncthreads();  // the LDS tile may still be read by a previous tile
    for (int idx = tid; idx < ncol * kRecFields; idx += kFillThreads) {
        const int c = idx / kRecFields, f = idx - c * kRecFields;
        lds.col[c][f] = rec[(size_t)lds_id(lds0 + c) * kRecStride + f];
    }
    __syncthreads();
    const int lane = tid & (kWave - 1);
    const int w = tid / kWave;
    const int i = lane0 + lane;
    if (i >= nlane) return;
    RwgRec<double> a;
    load_rec(rec, lane_ids[i], a);
#pragma unroll 1
    for (int c = w; c < ncol; c += kFillWaves) {
        const int j = lds0 + c;
        double zr, zi;
        entry_f64(a, lds.col[c], k, c4k2, scale, zr, zi);
        if constexpr (kLanesAreRows)
            store_c(&out[(size_t)j * ld + i], zr, zi);
        else if (j < nl1)
            store_c(&out[(size_t)i * ld + j], zr, zi);
        else
            store_c(&((OutC*)tail.out)[(size_t)i * tail.ld + (j - nl1)], zr, zi);
    }
}

// ---------------------------------------------------------------- FP32 tile
// The tile's distinct lane-set triangles (at most two per RWG) are found with
// an LDS hash table. Items (triangle a, LDS-set RWG c), numbered it = c nt + a
// over the tile's nt triangles, are evaluated 256 per round (one per thread)
// into a ring of kTriRing moment slots; after each round every column whose
// items are all done is assembled, each entry gathering the items of its two
// triangles. A column's unassembled items start less than nt <= kTriMax items
// before the round's end, so a ring of 256 + kTriMax slots is never
// overwritten before it is read.
constexpr int kTriMax = 2 * kTileRows;
constexpr int kTriHash = 256;                    // open addressing, load factor <= 1/2
constexpr int kTriRing = kFillThreads + kTriMax;  // moment slots
constexpr int kSrcFields = 44;   // u_j (18), rho_j (18), l (1), |u_j|^2 (6), pad
constexpr int kMoments = 10;     // A, B (3), C: complex

struct FillTriLds {
    float src[kTileCols][kSrcFields];
    double srcc[kTileCols][3];
    float tu[9][kTriMax];   // u_ti of the distinct triangles (structure of arrays)
    float tuu[3][kTriMax];  // |u_ti|^2
    double tc[3][kTriMax];  // centroids
    int hkey[kTriHash];
    int hval[kTriHash];
    int tid_of[kTriMax];    // triangle id of each local triangle
    int loc[kTriMax];       // local triangle of (lane RWG i, half h) at 2 i + h
    int ntri;
    float mom[kMoments][kTriRing];
};

// Moments A, B, C of item (local triangle a, LDS-set RWG c), already rotated
// by e^{-jkR0}, into L.mom[.][it].
__device__ __forceinline__ void tri_item(FillTriLds& L, int a, int c, int it, const F32Consts& fc) {
    const float* S = L.src[c];
    const double dx64 = L.tc[0][a] - L.srcc[c][0], dy64 = L.tc[1][a] - L.srcc[c][1],
                 dz64 = L.tc[2][a] - L.srcc[c][2];
    const double r064 = sqrt(dx64 * dx64 + dy64 * dy64 + dz64 * dz64);
    const double ph = r064 * fc.inv_lambda64;
    const float f0 = (float)(ph - rint(ph));
    const float s0 = __builtin_amdgcn_sinf(f0), c0 = __builtin_amdgcn_cosf(f0);
    const float dx = (float)dx64, dy = (float)dy64, dz = (float)dz64;
    const float r0 = (float)r064, r0sq = r0 * r0;
    const float ihl = 0.5f * fc.inv_lambda;  // 1 / (2 lambda)
    const float twol = 1.0f / ihl;           // 2 lambda
    const float c2 = twol / r0;              // 2 lambda / R0
    const float m2ihl = -2.0f * ihl;
    float ua[3][3], alpha[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        ua[i][0] = L.tu[i * 3 + 0][a];
        ua[i][1] = L.tu[i * 3 + 1][a];
        ua[i][2] = L.tu[i * 3 + 2][a];
        const float du = dx * ua[i][0] + dy * ua[i][1] + dz * ua[i][2];
        alpha[i] = (2.0f * du + L.tuu[i][a]) * ihl;
    }
    float tr[3][3] = {}, tn[3][3] = {};  // T_i = tr - j tn
    float chn = 0, cgn = 0;              // C = -(chn + j cgn)
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        const float ux = S[j * 3 + 0], uy = S[j * 3 + 1], uz = S[j * 3 + 2];
        const float px = S[18 + j * 3 + 0], py = S[18 + j * 3 + 1], pz = S[18 + j * 3 + 2];
        const float beta = fmaf(m2ihl, dx * ux + dy * uy + dz * uz, S[37 + j] * ihl);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const float wd = ua[i][0] * ux + ua[i][1] * uy + ua[i][2] * uz;
            const float xl = fmaf(m2ihl, wd, alpha[i] + beta);  // X / (2 lambda)
            const float r2 = fmaf(xl, twol, r0sq);
            const float ir = __builtin_amdgcn_rsqf(r2);
            const float th = xl * __builtin_amdgcn_rcpf(fmaf(r2, ir, r0));  // delta / (2 lambda)
            const float sh = __builtin_amdgcn_sinf(th);
            const float sn = __builtin_amdgcn_sinf(th + th);
            const float v = sh * sh * ir;
            const float gr = fmaf(-2.0f, v, ir);          // Re g
            const float hneg = fmaf(th * ir, c2, v + v);  // -Re(g - 1/R0)
            const float gin = sn * ir;                    // -Im g
            tr[i][0] = fmaf(px, gr, tr[i][0]);
            tr[i][1] = fmaf(py, gr, tr[i][1]);
            tr[i][2] = fmaf(pz, gr, tr[i][2]);
            tn[i][0] = fmaf(px, gin, tn[i][0]);
            tn[i][1] = fmaf(py, gin, tn[i][1]);
            tn[i][2] = fmaf(pz, gin, tn[i][2]);
            if (j < 3) {
                chn += hneg;
                cgn += gin;
            } else {
                chn -= hneg;
                cgn -= gin;
            }
        }
    }
    // A = sum_i u_ai . T_i, B = sum_i T_i, C; then times e^{-jkR0} = c0 - j s0.
    float ar = 0, an = 0, br[3], bn[3];
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        br[x] = tr[0][x] + tr[1][x] + tr[2][x];
        bn[x] = tn[0][x] + tn[1][x] + tn[2][x];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            ar = fmaf(ua[i][x], tr[i][x], ar);
            an = fmaf(ua[i][x], tn[i][x], an);
        }
    }
    // (xr - j xn)(c0 - j s0) = (xr c0 - xn s0) - j (xn c0 + xr s0)
    auto rot = [&](float xr, float xn, int f) {
        L.mom[f][it] = xr * c0 - xn * s0;
        L.mom[f + 1][it] = -(xn * c0 + xr * s0);
    };
    rot(ar, an, 0);
    rot(br[0], bn[0], 2);
    rot(br[1], bn[1], 4);
    rot(br[2], bn[2], 6);
    rot(-chn, cgn, 8);  // C = -chn - j cgn
}

// One tile of a panel, FP32 (panels out of OutC = c64). Same tile geometry
// and arguments as fill_tile (including the LDS-set tail); lane_perm (or NULL for the list order) orders the
// lane set so that a tile's 64 RWGs form a compact patch sharing triangles:
// tile lane i takes position lane_perm[lane0 + i] of the lane set. Must be
// called by all 256 threads.
template <typename OutC, bool kLanesAreRows>
__device__ __forceinline__ void fill_tile_tri(OutC* __restrict__ out, int ld, const int* __restrict__ rows, int nr,
                                              const int* __restrict__ cols, int nc, int row0, int col0,
                                              const int32_t* __restrict__ lane_perm, const float* __restrict__ rec,
                                              const double* __restrict__ cen, const TriGeom& tg,
                                              const F32Consts& fc, FillTriLds& L, LdsTail tail) {
    const int* lane_ids = kLanesAreRows ? rows : cols;
    const int* lds_ids = kLanesAreRows ? cols : rows;
    const int nlane = kLanesAreRows ? nr : nc;
    const int nl1 = kLanesAreRows ? nc : nr;
    const int nlds = kLanesAreRows ? nl1 : nl1 + tail.n;
    auto lds_id = [&](int j) {
        if constexpr (kLanesAreRows)
            return lds_ids[j];
        else
            return j < nl1 ? lds_ids[j] : tail.ids[j - nl1];
    };
    const int lane0 = kLanesAreRows ? row0 : col0;
    const int lds0 = kLanesAreRows ? col0 : row0;
    const int tid = threadIdx.x;
    const int nl = min(kTileRows, nlane - lane0);
    const int ncol = min(kTileCols, nlds - lds0);
    __syncthreads();  // the LDS of a previous tile may still be read
    for (int idx = tid; idx < ncol * kRecFields; idx += kFillThreads) {
        const int c = idx / kRecFields, f = idx - c * kRecFields;
        L.src[c][f] = rec[(size_t)lds_id(lds0 + c) * kRecStride + f];
    }
    for (int idx = tid; idx < ncol * 6; idx += kFillThreads) {
        const int c = idx / 6, j = idx - 6 * c;
        const float* u = rec + (size_t)lds_id(lds0 + c) * kRecStride + j * 3;
        L.src[c][37 + j] = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    }
    for (int idx = tid; idx < ncol * 3; idx += kFillThreads) {
        const int c = idx / 3, x = idx - 3 * c;
        L.srcc[c][x] = cen[(size_t)lds_id(lds0 + c) * kCenStride + x];
    }
    for (int idx = tid; idx < kTriHash; idx += kFillThreads) L.hkey[idx] = -1;
    if (tid == 0) L.ntri = 0;
    __syncthreads();
    // Distinct triangles of the lane set: slot 2 i + h inserts the triangle of
    // half h of lane RWG i; the first inserter numbers it. The numbering
    // depends on timing, the results do not (every item is evaluated alike).
    int hslot = 0;
    auto lane_pos = [&](int i) { return lane_perm ? lane_perm[lane0 + i] : lane0 + i; };
    if (tid < 2 * nl) {
        const int t = tg.rtri[2 * (size_t)lane_ids[lane_pos(tid >> 1)] + (tid & 1)];
        unsigned h = ((unsigned)t * 2654435761u) >> 24;
        while (true) {
            const int old = atomicCAS(&L.hkey[h], -1, t);
            if (old == -1) {
                const int k = atomicAdd(&L.ntri, 1);
                L.hval[h] = k;
                L.tid_of[k] = t;
                break;
            }
            if (old == t) break;
            h = (h + 1) & (kTriHash - 1);
        }
        hslot = (int)h;
    }
    __syncthreads();
    if (tid < 2 * nl) L.loc[tid] = L.hval[hslot];
    const int nt = L.ntri;
    for (int idx = tid; idx < nt * kTriStride; idx += kFillThreads) {
        const int a = idx / kTriStride, f = idx - a * kTriStride;
        const float v = tg.tri[(size_t)L.tid_of[a] * kTriStride + f];
        if (f < 9)
            L.tu[f][a] = v;
        else
            L.tuu[f - 9][a] = v;
    }
    for (int idx = tid; idx < nt * 3; idx += kFillThreads) {
        const int a = idx / 3, x = idx - 3 * a;
        L.tc[x][a] = tg.tcen[(size_t)L.tid_of[a] * kCenStride + x];
    }
    __syncthreads();
    const int lane = tid & (kWave - 1);
    const int w = tid / kWave;
    const bool has_lane = lane < nl;
    int la[2] = {0, 0}, i = 0;
    float wv[2][3] = {}, lm = 0;
    if (has_lane) {
        la[0] = L.loc[2 * lane];
        la[1] = L.loc[2 * lane + 1];
        i = lane_pos(lane);
        const float* lr = tg.lrec + (size_t)lane_ids[i] * kLaneStride;
#pragma unroll
        for (int x = 0; x < 6; ++x) wv[x / 3][x % 3] = lr[x];
        lm = lr[6];
    }
    const int nitems = nt * ncol;
    int cdone = 0;
    for (int i0 = 0; i0 < nitems; i0 += kFillThreads) {
        const int it = i0 + tid;
        if (it < nitems) {
            const int c = it / nt;
            tri_item(L, it - c * nt, c, it % kTriRing, fc);
        }
        __syncthreads();
        const int cnew = min(ncol, (i0 + kFillThreads) / nt);
        if (has_lane) {
            for (int c = cdone + w; c < cnew; c += kFillWaves) {
                float zr = 0, zi = 0;
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int q = (c * nt + la[h]) % kTriRing;
                    const float vr = L.mom[0][q] - wv[h][0] * L.mom[2][q] - wv[h][1] * L.mom[4][q] -
                                     wv[h][2] * L.mom[6][q] - fc.c4k2 * L.mom[8][q];
                    const float vi = L.mom[1][q] - wv[h][0] * L.mom[3][q] - wv[h][1] * L.mom[5][q] -
                                     wv[h][2] * L.mom[7][q] - fc.c4k2 * L.mom[9][q];
                    if (h == 0) {
                        zr += vr;
                        zi += vi;
                    } else {
                        zr -= vr;
                        zi -= vi;
                    }
                }
                // Z = -j kappa (zr + j zi)
                const float kap = fc.scale * lm * L.src[c][36];
                const int j = lds0 + c;
                if constexpr (kLanesAreRows)
                    store_c(&out[(size_t)j * ld + i], kap * zi, -kap * zr);
                else if (j < nl1)
                    store_c(&out[(size_t)i * ld + j], kap * zi, -kap * zr);
                else
                    store_c(&((OutC*)tail.out)[(size_t)i * tail.ld + (j - nl1)], kap * zi, -kap * zr);
            }
        }
        cdone = cnew;
        __syncthreads();
    }
}

// Constants of an FP64 fill.
struct FillConsts {
    double k;      // wavenumber
    double c4k2;   // 4 / k^2
    double scale;  // k eta / (144 pi)
};

// Fill constants for the wavenumber k and impedance eta.
__host__ __device__ inline FillConsts fill_consts_f64(double k, double eta) {
    return FillConsts{k, 4.0 / (k * k), k * eta / (144.0 * M_PI)};
}

}  // namespace farcur
