// Host side of the C ABI: contexts, geometry upload, the per-block adaptive
// loop (the reference's cuda_cur_compress3 driving cuda_complete_compression),
// the batched driver and the drop-in entry point.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "farcur.h"
#include "element.hpp"
#include "geometry.hpp"
#include "kernels.hpp"
#include "mlp.hpp"
#include "mrhs.hpp"
#include "noalloc.hpp"
#include "verify.hpp"
#include "near.hpp"
#include <complex>
#include "philox.hpp"

using namespace farcur;

struct farcur_mesh {
    Mesh m;
    bool rwg = false;
};
struct farcur_partition {
    BlockPartition p;
};

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPTRY(expr)                                                                            \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) return fail(FARCUR_EHIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

#define TRY(expr)                   \
    do {                            \
        int rc_ = (expr);           \
        if (rc_ != FARCUR_OK) return rc_; \
    } while (0)

// Grow-only device buffer.
struct DBuf {
    void* p = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes) {
        if (bytes <= cap && p) return FARCUR_OK;
        if (p) hipFree(p);
        p = nullptr;
        cap = 0;
        const size_t want = std::max<size_t>(bytes, 256);
        if (hipMalloc(&p, want) != hipSuccess) {
            p = nullptr;
            return fail(FARCUR_ENOMEM, "hipMalloc of %zu bytes failed", want);
        }
        cap = want;
        return FARCUR_OK;
    }
    void release() {
        if (p) hipFree(p);
        p = nullptr;
        cap = 0;
    }
    template <typename T>
    T* as() const {
        return (T*)p;
    }
};

// Bump allocator over device chunks; reset() keeps the chunks.
struct Arena {
    std::vector<std::pair<char*, size_t>> chunks;
    size_t cur = 0, off = 0;
    size_t chunk_bytes = (size_t)1 << 30;
    void reset() {
        cur = 0;
        off = 0;
    }
    int alloc(size_t bytes, void** out) {
        bytes = (bytes + 255) & ~(size_t)255;
        while (true) {
            if (cur < chunks.size() && off + bytes <= chunks[cur].second) {
                *out = chunks[cur].first + off;
                off += bytes;
                return FARCUR_OK;
            }
            if (cur + 1 < chunks.size()) {
                ++cur;
                off = 0;
                continue;
            }
            const size_t sz = std::max(chunk_bytes, bytes);
            char* p = nullptr;
            if (hipMalloc((void**)&p, sz) != hipSuccess) return fail(FARCUR_ENOMEM, "arena hipMalloc %zu failed", sz);
            chunks.emplace_back(p, sz);
            cur = chunks.size() - 1;
            off = 0;
        }
    }
    void release() {
        for (auto& c : chunks) hipFree(c.first);
        chunks.clear();
        reset();
    }
};

__global__ void k_set_ones(c128* y, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        y[i] = c128{1.0, 0.0};
}

void launch_ones(c128* y, size_t n, hipStream_t s) {
    if (n == 0) return;
    const int grid = (int)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_set_ones, dim3(grid), dim3(256), 0, s, y, n);
}

inline int r_start_rule(int D) { return D < 10 ? 1 : D / 10; }

// splitmix64 for host-side shuffles and initialisation.
struct SplitMixHost {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

// Per-kernel-class device timing by event pairs.
struct KTimer {
    // sample, fill, pinv (all), residual, then pinv per r class (4 + c)
    static constexpr int kClasses = 4 + kPinvClasses;
    std::vector<hipEvent_t> pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> marks;
    size_t used = 0;
    double ms[kClasses] = {};
    int launches[kClasses] = {};
    bool on = false;
    hipEvent_t get() {
        if (used == pool.size()) {
            hipEvent_t e;
            hipEventCreate(&e);
            pool.push_back(e);
        }
        return pool[used++];
    }
    void begin(int cls, hipStream_t s, hipEvent_t* a) {
        if (!on) return;
        *a = get();
        hipEventRecord(*a, s);
        (void)cls;
    }
    void end(int cls, hipStream_t s, hipEvent_t a) {
        if (!on) return;
        hipEvent_t b = get();
        hipEventRecord(b, s);
        marks.push_back({cls, {a, b}});
    }
    // After a stream sync: accumulate and recycle.
    void collect() {
        for (auto& mk : marks) {
            float t = 0;
            hipEventElapsedTime(&t, mk.second.first, mk.second.second);
            ms[mk.first] += t;
            launches[mk.first] += 1;
            if (mk.first >= 4) {  // a pinv class also counts towards "pinv"
                ms[2] += t;
                launches[2] += 1;
            }
        }
        marks.clear();
        used = 0;
    }
    void clear() {
        for (int i = 0; i < kClasses; ++i) {
            ms[i] = 0;
            launches[i] = 0;
        }
    }
    ~KTimer() {
        for (auto e : pool) hipEventDestroy(e);
    }
};

}  // namespace

// Launch shapes, chosen by the autotuner (farcur_autotune) or defaults.
struct Tuning {
    // the autotuner's choice on an MI355X for the truncated QR core of the
    // learned rounds (narrow teams: small cores take few threads)
    int pinv_threads[kPinvClasses] = {64, 64, 128, 128, 128, 256, 256, 256, 512, 512, 512};
    // fill workgroups per launch, capped at the launch's tile count: one per
    // tile balances tiles of unequal cost best (sphere: fill 5.40 ms with a
    // persistent grid of 4096, 5.10 with one workgroup per tile)
    int fill_grid = 1 << 20;
    int resid_threads = 256;
};

struct farcur_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    double k = 2.0 * M_PI;
    double eta = 120.0 * M_PI;
    int ne = 0, nt = 0;
    std::vector<int32_t> tri_of_rwg;  // 2 * ne: T+, T- (host, for accounting)
    std::vector<double> h_vertex;     // host copies for block features
    std::vector<int32_t> h_topol, h_edges;
    std::vector<double> h_ln;
    DBuf rec64, rec32, cen;
    DBuf tri32, tcen, lrec32, rtri;  // triangle tables of the FP32 fill
    std::vector<double> h_cen;       // RWG centres (kCenStride), for the fill's tile order
    uint64_t geom_hash = 0;
    // single-block scratch
    DBuf d_rows, d_cols, d_cs, d_rs, d_rpos, d_C, d_R, d_W, d_V, d_U, d_sigma, d_v, d_y0, d_y1, d_plan, d_out;
    int last_m = 0, last_n = 0, last_r = 0, last_f32 = 0, last_y = 0;
    bool has_result = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // Second stream of the batched rounds' two-stream pipeline (default
    // priority, like the first) and the events joining the two.
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_pfork = nullptr, ev_pjoin = nullptr;
    Tuning tune;

    int select() {
        HIPTRY(hipSetDevice(device));
        return FARCUR_OK;
    }
    FillParams fill_params(int f32) const {
        FillParams fp;
        fp.rec = f32 ? rec32.p : rec64.p;
        fp.cen = cen.as<double>();
        fp.tri = TriGeom{tri32.as<float>(), tcen.as<double>(), lrec32.as<float>(), rtri.as<int32_t>()};
        fp.ne = ne;
        fp.f32 = f32;
        fp.k = k;
        fp.eta = eta;
        return fp;
    }
};

namespace {

// RWG records (kRecFields x ne SoA) from the reference tables, 0-based ids.
void build_records(const std::vector<double>& vertex, const std::vector<int32_t>& topol,
                   const std::vector<int32_t>& edges, const std::vector<double>& ln, std::vector<double>& rec,
                   std::vector<float>& rec32, std::vector<double>& cen) {
    const int ne = (int)ln.size();
    rec.assign((size_t)kRecStride * ne, 0.0);
    rec32.assign((size_t)kRecStride * ne, 0.0f);
    cen.assign((size_t)kCenStride * ne, 0.0);
    const double b[3][3] = {{2.0 / 3, 1.0 / 6, 1.0 / 6}, {1.0 / 6, 2.0 / 3, 1.0 / 6}, {1.0 / 6, 1.0 / 6, 2.0 / 3}};
    for (int e = 0; e < ne; ++e) {
        double* R = &rec[(size_t)e * kRecStride];
        for (int h = 0; h < 2; ++h) {
            const int t = edges[4 * e + h];
            const int vf = edges[4 * e + 2 + h];
            const double s = h == 0 ? 1.0 : -1.0;
            const double* V0 = &vertex[3 * topol[3 * t]];
            const double* V1 = &vertex[3 * topol[3 * t + 1]];
            const double* V2 = &vertex[3 * topol[3 * t + 2]];
            for (int i = 0; i < 3; ++i) {
                for (int c = 0; c < 3; ++c) {
                    const double q = b[i][0] * V0[c] + b[i][1] * V1[c] + b[i][2] * V2[c];
                    R[h * 9 + i * 3 + c] = q;
                    R[18 + h * 9 + i * 3 + c] = s * (q - vertex[3 * vf + c]);
                }
            }
        }
        R[36] = ln[e];
        // FP32 record: quadrature points as offsets from the RWG centre (mean of
        // the six points), centre kept in FP64.
        float* R32 = &rec32[(size_t)e * kRecStride];
        for (int f = 18; f < kRecFields; ++f) R32[f] = (float)R[f];
        for (int c = 0; c < 3; ++c) {
            double m = 0;
            for (int h = 0; h < 2; ++h)
                for (int i = 0; i < 3; ++i) m += R[h * 9 + i * 3 + c];
            m /= 6.0;
            cen[(size_t)e * kCenStride + c] = m;
            for (int h = 0; h < 2; ++h)
                for (int i = 0; i < 3; ++i) R32[h * 9 + i * 3 + c] = (float)(R[h * 9 + i * 3 + c] - m);
        }
    }
}

// Triangle tables of the FP32 fill: per triangle the quadrature points as
// offsets from its centroid and their squared norms (FP32, kTriStride) and the
// centroid (FP64, kCenStride); per RWG the free vertex of each half relative
// to that triangle's centroid and the edge length (FP32, kLaneStride), and the
// two triangle ids.
void build_triangles(const std::vector<double>& vertex, const std::vector<int32_t>& topol,
                     const std::vector<int32_t>& edges, const std::vector<double>& ln, std::vector<float>& tri,
                     std::vector<double>& tcen, std::vector<float>& lrec, std::vector<int32_t>& rtri) {
    const int nt = (int)(topol.size() / 3), ne = (int)ln.size();
    const double b[3][3] = {{2.0 / 3, 1.0 / 6, 1.0 / 6}, {1.0 / 6, 2.0 / 3, 1.0 / 6}, {1.0 / 6, 1.0 / 6, 2.0 / 3}};
    tri.assign((size_t)kTriStride * nt, 0.0f);
    tcen.assign((size_t)kCenStride * nt, 0.0);
    for (int t = 0; t < nt; ++t) {
        double q[3][3], c[3] = {0, 0, 0};
        for (int i = 0; i < 3; ++i)
            for (int x = 0; x < 3; ++x) {
                q[i][x] = b[i][0] * vertex[3 * topol[3 * t] + x] + b[i][1] * vertex[3 * topol[3 * t + 1] + x] +
                          b[i][2] * vertex[3 * topol[3 * t + 2] + x];
                c[x] += q[i][x] / 3.0;
            }
        float* T = &tri[(size_t)t * kTriStride];
        for (int i = 0; i < 3; ++i) {
            double uu = 0;
            for (int x = 0; x < 3; ++x) {
                const double u = q[i][x] - c[x];
                T[i * 3 + x] = (float)u;
                uu += u * u;
            }
            T[9 + i] = (float)uu;
        }
        for (int x = 0; x < 3; ++x) tcen[(size_t)t * kCenStride + x] = c[x];
    }
    lrec.assign((size_t)kLaneStride * ne, 0.0f);
    rtri.assign(2 * (size_t)ne, 0);
    for (int e = 0; e < ne; ++e) {
        float* L = &lrec[(size_t)e * kLaneStride];
        for (int h = 0; h < 2; ++h) {
            const int t = edges[4 * e + h], vf = edges[4 * e + 2 + h];
            rtri[2 * (size_t)e + h] = t;
            for (int x = 0; x < 3; ++x) L[h * 3 + x] = (float)(vertex[3 * vf + x] - tcen[(size_t)t * kCenStride + x]);
        }
        L[6] = (float)ln[e];
    }
}

// Tile order of an RWG list for the FP32 fill: positions 0..n-1 arranged so
// that each run of kTileRows consecutive entries is a compact patch (the
// tile's RWGs then share most of their triangles). Recursive bisection along
// the longest extent of the RWG centres, split at the multiple of kTileRows
// nearest the middle; runs of at most kTileRows keep the list order.
void tile_order(const int32_t* ids, int n, const double* cen, int32_t* perm) {
    std::vector<int32_t> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    struct Range {
        int lo, hi;
    };
    std::vector<Range> stack{{0, n}};
    while (!stack.empty()) {
        const Range rg = stack.back();
        stack.pop_back();
        const int len = rg.hi - rg.lo;
        if (len <= kTileRows) continue;
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int i = rg.lo; i < rg.hi; ++i)
            for (int x = 0; x < 3; ++x) {
                const double v = cen[(size_t)ids[idx[i]] * kCenStride + x];
                lo[x] = std::min(lo[x], v);
                hi[x] = std::max(hi[x], v);
            }
        int ax = 0;
        for (int x = 1; x < 3; ++x)
            if (hi[x] - lo[x] > hi[ax] - lo[ax]) ax = x;
        std::stable_sort(idx.begin() + rg.lo, idx.begin() + rg.hi, [&](int32_t a, int32_t b) {
            return cen[(size_t)ids[a] * kCenStride + ax] < cen[(size_t)ids[b] * kCenStride + ax];
        });
        const int k = std::max(kTileRows, (len / 2 + kTileRows / 2) / kTileRows * kTileRows);
        stack.push_back({rg.lo, rg.lo + k});
        stack.push_back({rg.lo + k, rg.hi});
    }
    std::copy(idx.begin(), idx.end(), perm);
}

uint64_t fnv1a(uint64_t h, const void* data, size_t n) {
    const unsigned char* p = (const unsigned char*)data;
    // 8 bytes at a time for speed; still position-sensitive.
    size_t i = 0;
    for (; i + 8 <= n; i += 8) {
        uint64_t w;
        std::memcpy(&w, p + i, 8);
        h ^= w;
        h *= 0x100000001B3ull;
        h ^= h >> 29;
    }
    for (; i < n; ++i) {
        h ^= p[i];
        h *= 0x100000001B3ull;
    }
    return h;
}

int set_geometry_impl(farcur_ctx* ctx, const std::vector<double>& vertex, const std::vector<int32_t>& topol,
                      const std::vector<int32_t>& edges, const std::vector<double>& ln) {
    const int nv = (int)(vertex.size() / 3), nt = (int)(topol.size() / 3), ne = (int)ln.size();
    if (nv <= 0 || nt <= 0 || ne <= 0) return fail(FARCUR_EINVAL, "empty geometry (nv=%d nt=%d ne=%d)", nv, nt, ne);
    for (size_t i = 0; i < topol.size(); ++i)
        if (topol[i] < 0 || topol[i] >= nv) return fail(FARCUR_EINVAL, "topol entry %zu out of range", i);
    for (int e = 0; e < ne; ++e) {
        for (int h = 0; h < 2; ++h) {
            if (edges[4 * e + h] < 0 || edges[4 * e + h] >= nt)
                return fail(FARCUR_EINVAL, "edges(%d,%d) triangle out of range", h + 1, e + 1);
            if (edges[4 * e + 2 + h] < 0 || edges[4 * e + 2 + h] >= nv)
                return fail(FARCUR_EINVAL, "edges(%d,%d) vertex out of range", h + 3, e + 1);
        }
        if (!(ln[e] > 0) || !std::isfinite(ln[e])) return fail(FARCUR_EINVAL, "ln(%d) not positive finite", e + 1);
    }
    for (double x : vertex)
        if (!std::isfinite(x)) return fail(FARCUR_EINVAL, "non-finite vertex coordinate");
    std::vector<double> rec, cen;
    std::vector<float> rec32;
    build_records(vertex, topol, edges, ln, rec, rec32, cen);
    std::vector<float> tri32, lrec32;
    std::vector<double> tcen;
    std::vector<int32_t> rtri;
    build_triangles(vertex, topol, edges, ln, tri32, tcen, lrec32, rtri);
    TRY(ctx->select());
    TRY(ctx->tri32.ensure(tri32.size() * sizeof(float)));
    TRY(ctx->tcen.ensure(tcen.size() * sizeof(double)));
    TRY(ctx->lrec32.ensure(lrec32.size() * sizeof(float)));
    TRY(ctx->rtri.ensure(rtri.size() * sizeof(int32_t)));
    HIPTRY(hipMemcpyAsync(ctx->tri32.p, tri32.data(), tri32.size() * sizeof(float), hipMemcpyHostToDevice,
                          ctx->stream));
    HIPTRY(hipMemcpyAsync(ctx->tcen.p, tcen.data(), tcen.size() * sizeof(double), hipMemcpyHostToDevice,
                          ctx->stream));
    HIPTRY(hipMemcpyAsync(ctx->lrec32.p, lrec32.data(), lrec32.size() * sizeof(float), hipMemcpyHostToDevice,
                          ctx->stream));
    HIPTRY(hipMemcpyAsync(ctx->rtri.p, rtri.data(), rtri.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                          ctx->stream));
    TRY(ctx->rec64.ensure(rec.size() * sizeof(double)));
    TRY(ctx->rec32.ensure(rec32.size() * sizeof(float)));
    TRY(ctx->cen.ensure(cen.size() * sizeof(double)));
    HIPTRY(hipMemcpyAsync(ctx->cen.p, cen.data(), cen.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIPTRY(hipMemcpyAsync(ctx->rec64.p, rec.data(), rec.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIPTRY(hipMemcpyAsync(ctx->rec32.p, rec32.data(), rec32.size() * sizeof(float), hipMemcpyHostToDevice,
                          ctx->stream));
    HIPTRY(hipStreamSynchronize(ctx->stream));
    ctx->ne = ne;
    ctx->nt = nt;
    ctx->h_cen = cen;
    ctx->h_vertex = vertex;
    ctx->h_topol = topol;
    ctx->h_edges = edges;
    ctx->h_ln = ln;
    ctx->tri_of_rwg.resize(2 * (size_t)ne);
    for (int e = 0; e < ne; ++e) {
        ctx->tri_of_rwg[2 * e] = edges[4 * e];
        ctx->tri_of_rwg[2 * e + 1] = edges[4 * e + 1];
    }
    ctx->has_result = false;
    return FARCUR_OK;
}

int check_ids(const int* ids, int n, int ne, const char* what) {
    for (int i = 0; i < n; ++i)
        if (ids[i] < 0 || ids[i] >= ne) return fail(FARCUR_EINVAL, "%s[%d] = %d out of range [0, %d)", what, i, ids[i], ne);
    return FARCUR_OK;
}

// Far-field precondition of a block rows x cols. The fill integrates every
// triangle pair with the 3-point rule, which is only valid when the test and
// source triangles are apart: coincident triangles put quadrature points on
// top of each other (R = 0, a non-finite entry) and triangles that share a
// vertex or an edge get a wrong value with no sign of it. The reference
// switches such pairs, and every pair of triangle centroids closer than its
// Rinteg_s / Ranal_s radii, to singular integration (options passed at
// codes/GPU_implementation/seed/compres_single_block_spheres_for_sweep2.jl:1095-1098
// into the element, user_impedance_for_julia_cuda.cu:178-214), which a
// compressed far block never needs. So a block whose row and column triangles
// share a vertex, or (rmin > 0) have two centroids closer than rmin, is
// refused with FARCUR_EINVAL. `stamp` holds one int per vertex (all below
// `tag` on entry); the caller passes a fresh tag per block.
struct FarCheck {
    std::vector<int32_t> stamp;
    std::vector<int32_t> tri_a, tri_b;
    int32_t tag = 0;
};

int check_far_block(const farcur_ctx* ctx, const int32_t* rows, int m, const int32_t* cols, int n, double rmin,
                    FarCheck& fc, const char* what, int block) {
    const int nv = (int)(ctx->h_vertex.size() / 3);
    if ((int)fc.stamp.size() != nv) {
        fc.stamp.assign(nv, 0);
        fc.tag = 0;
    }
    if (fc.tag == INT32_MAX) {
        std::fill(fc.stamp.begin(), fc.stamp.end(), 0);
        fc.tag = 0;
    }
    const int32_t tag = ++fc.tag;
    for (int i = 0; i < m; ++i)
        for (int h = 0; h < 2; ++h) {
            const int t = ctx->tri_of_rwg[2 * (size_t)rows[i] + h];
            for (int k = 0; k < 3; ++k) fc.stamp[ctx->h_topol[3 * t + k]] = tag;
        }
    for (int j = 0; j < n; ++j)
        for (int h = 0; h < 2; ++h) {
            const int t = ctx->tri_of_rwg[2 * (size_t)cols[j] + h];
            for (int k = 0; k < 3; ++k)
                if (fc.stamp[ctx->h_topol[3 * t + k]] == tag)
                    return fail(FARCUR_EINVAL,
                                "%s %d: column RWG %d touches a row triangle (shared vertex %d): a near-field pair, "
                                "which the far-field quadrature cannot integrate",
                                what, block, cols[j], ctx->h_topol[3 * t + k]);
        }
    if (!(rmin > 0)) return FARCUR_OK;
    // triangle centroids of both sides; a box test first, pairs only if the boxes come closer than rmin
    auto gather = [&](const int32_t* ids, int cnt, std::vector<int32_t>& out) {
        out.clear();
        for (int i = 0; i < cnt; ++i)
            for (int h = 0; h < 2; ++h) out.push_back(ctx->tri_of_rwg[2 * (size_t)ids[i] + h]);
        std::sort(out.begin(), out.end());
        out.erase(std::unique(out.begin(), out.end()), out.end());
    };
    gather(rows, m, fc.tri_a);
    gather(cols, n, fc.tri_b);
    auto centroid = [&](int t, double* c) {
        for (int x = 0; x < 3; ++x)
            c[x] = (ctx->h_vertex[3 * ctx->h_topol[3 * t] + x] + ctx->h_vertex[3 * ctx->h_topol[3 * t + 1] + x] +
                    ctx->h_vertex[3 * ctx->h_topol[3 * t + 2] + x]) / 3.0;
    };
    double lo[2][3], hi[2][3];
    for (int s = 0; s < 2; ++s)
        for (int x = 0; x < 3; ++x) {
            lo[s][x] = 1e300;
            hi[s][x] = -1e300;
        }
    for (int s = 0; s < 2; ++s)
        for (int t : (s == 0 ? fc.tri_a : fc.tri_b)) {
            double c[3];
            centroid(t, c);
            for (int x = 0; x < 3; ++x) {
                lo[s][x] = std::min(lo[s][x], c[x]);
                hi[s][x] = std::max(hi[s][x], c[x]);
            }
        }
    double gap2 = 0;
    for (int x = 0; x < 3; ++x) {
        const double g = std::max({0.0, lo[1][x] - hi[0][x], lo[0][x] - hi[1][x]});
        gap2 += g * g;
    }
    if (gap2 >= rmin * rmin) return FARCUR_OK;
    std::vector<double> cb(3 * fc.tri_b.size());
    for (size_t q = 0; q < fc.tri_b.size(); ++q) centroid(fc.tri_b[q], &cb[3 * q]);
    for (int ta : fc.tri_a) {
        double ca[3];
        centroid(ta, ca);
        for (size_t q = 0; q < fc.tri_b.size(); ++q) {
            const double dx = ca[0] - cb[3 * q], dy = ca[1] - cb[3 * q + 1], dz = ca[2] - cb[3 * q + 2];
            if (dx * dx + dy * dy + dz * dz < rmin * rmin)
                return fail(FARCUR_EINVAL,
                            "%s %d: triangles %d and %d are %.3g apart, closer than the integration radius %g: "
                            "a near-field pair, which the far-field quadrature cannot integrate",
                            what, block, ta, fc.tri_b[q], std::sqrt(dx * dx + dy * dy + dz * dz), rmin);
        }
    }
    return FARCUR_OK;
}

// Panel element size.
inline size_t psize(int f32) { return f32 ? sizeof(c64) : sizeof(c128); }

// Runs the kernels of one round for the uploaded job list.
struct RoundPlan {
    int njobs = 0;
    int f32 = 0;
    bool pivot = false;
    int fill_grid = 0;
    int pinv_threads[kPinvClasses] = {64, 64, 128, 128, 128, 256, 256, 256, 512, 512, 512};
    int resid_threads = 256;
    double atol = 1e-10, rtol = 0, label_eps = 1e-4;
    int one_buf = 0;  // every job gathers its core from C (one-buffer pinv allowed)
    int qr_core = 0;  // learned jobs take the truncated QR pseudo-inverse
    int n_noheld = 0, rmax_held = 0;  // as PackedRound: whether the residual launch has work
    uint64_t seed = 0;
    // device pointers
    const BlockJob* jobs = nullptr;
    const int32_t* prefix_all = nullptr;  // fill prefix: all panels (or C only when pivoting)
    const int32_t* prefix_wr = nullptr;   // fill prefix: W + R (pivot second pass)
    int total_all = 0, total_wr = 0;
    const int32_t* list[kPinvClasses] = {};
    int nlist[kPinvClasses] = {};
    JobOut* out = nullptr;
    void tune(const Tuning& t) {
        for (int c = 0; c < kPinvClasses; ++c) pinv_threads[c] = t.pinv_threads[c];
        fill_grid = t.fill_grid;
        resid_threads = t.resid_threads;
    }
};

// side: a second stream for the pinv classes (nullptr: everything on s). The
// classes are independent, so with a side stream the large-r classes (few cores,
// each long: the tail of a launch) run beside the many small ones.
void run_round(const farcur_ctx* ctx, const RoundPlan& P, KTimer& kt, hipStream_t s, hipStream_t side = nullptr) {
    const FillParams fp = ctx->fill_params(P.f32);
    hipEvent_t a;
    kt.begin(0, s, &a);
    launch_sample(P.jobs, P.njobs, P.seed, s);
    kt.end(0, s, a);
    if (P.pivot) {
        kt.begin(1, s, &a);
        launch_fill(P.jobs, P.njobs, P.prefix_all, P.total_all, kPanelC, fp, P.fill_grid, s);
        kt.end(1, s, a);
        kt.begin(0, s, &a);
        launch_pivot_rows(P.jobs, P.njobs, P.f32, s);
        kt.end(0, s, a);
        kt.begin(1, s, &a);
        launch_fill(P.jobs, P.njobs, P.prefix_wr, P.total_wr, kPanelW | kPanelR | kPanelH, fp, P.fill_grid, s);
        kt.end(1, s, a);
    } else {
        kt.begin(1, s, &a);
        launch_fill(P.jobs, P.njobs, P.prefix_all, P.total_all, kPanelC | kPanelW | kPanelR | kPanelH, fp,
                    P.fill_grid, s);
        kt.end(1, s, a);
    }
    if (side) {
        (void)hipEventRecord(ctx->ev_pfork, s);
        (void)hipStreamWaitEvent(side, ctx->ev_pfork, 0);
    }
    // largest classes first; with a side stream they alternate between the two
    int k = 0;
    for (int c = kPinvClasses - 1; c >= 0; --c) {
        if (P.nlist[c] == 0) continue;
        hipStream_t sc = side && (k++ & 1) == 0 ? side : s;
        kt.begin(4 + c, sc, &a);
        launch_pinv(P.jobs, P.list[c], P.nlist[c], P.atol, P.rtol, P.label_eps, kPinvClassMax[c], P.pinv_threads[c],
                    P.f32, P.one_buf, 1, P.qr_core, P.out, sc);
        kt.end(4 + c, sc, a);
    }
    if (side) {
        (void)hipEventRecord(ctx->ev_pjoin, side);
        (void)hipStreamWaitEvent(s, ctx->ev_pjoin, 0);
    }
    // held-out errors of the LDS-variant jobs came with their pinv launch; the
    // residual launch runs when some job is left to it
    const int skip = pinv_lds_limit(P.one_buf, P.rtol);
    if (P.n_noheld > 0 || P.rmax_held > skip) {
        kt.begin(3, s, &a);
        launch_residual(P.jobs, P.njobs, P.f32, P.resid_threads, skip, P.out, s);
        kt.end(3, s, a);
    }
}

// Host-side packing of a round: jobs, fill prefixes, pinv lists by r class,
// into one device upload. pack_layout computes the sections (prefix sums and
// class lists into `scratch`), pack_write writes them at `dst` with the
// padding zeroed, so equal plans are equal bytes.
struct PackedRound {
    size_t off_jobs = 0, off_pa = 0, off_pwr = 0, off_out = 0, size = 0;
    size_t off_list[kPinvClasses] = {};
    int total_all = 0, total_wr = 0;
    int all_rpos = 1;  // every job gathers its core from C
    int nlist[kPinvClasses] = {};
    // jobs of the plain loops (successive residual) and the largest core with
    // held-out rows (held-out check in the pinv launch up to its LDS limit)
    int n_noheld = 0, rmax_held = 0;
};

struct PackScratch {
    std::vector<int32_t> pa, pwr;
    std::vector<int32_t> lists[kPinvClasses];
};

// What the layout needs of a job: its sizes and whether it has row positions.
struct JobDims {
    int32_t m, n, r, h, rpos;
    int32_t learned;  // a learned job: without held-out rows it is stored exactly
};

inline JobDims dims_of(const BlockJob& J) {
    return JobDims{J.m, J.n, J.r, J.h, J.rpos != nullptr, J.sample_mode == 3};
}

template <class Dims>
void pack_layout(int nj, bool pivot, PackedRound& pk, PackScratch& sc, const Dims& dims) {
    sc.pa.assign(nj + 1, 0);
    sc.pwr.assign(nj + 1, 0);
    for (auto& l : sc.lists) l.clear();
    pk.all_rpos = 1;
    pk.n_noheld = pk.rmax_held = 0;
    for (int j = 0; j < nj; ++j) {
        const JobDims J = dims(j);
        if (!J.rpos) pk.all_rpos = 0;
        const int pw = J.rpos ? 0 : kPanelW;  // with row positions W is gathered from C, not filled
        const int all = pivot ? fill_tiles(J.m, J.n, J.r, J.h, kPanelC)
                              : fill_tiles(J.m, J.n, J.r, J.h, kPanelC | pw | kPanelR | kPanelH);
        sc.pa[j + 1] = sc.pa[j] + all;
        sc.pwr[j + 1] = sc.pwr[j] + (pivot ? fill_tiles(J.m, J.n, J.r, J.h, pw | kPanelR | kPanelH) : 0);
        sc.lists[pinv_class(J.r)].push_back(j);
        if (J.h > 0)
            pk.rmax_held = std::max(pk.rmax_held, J.r);
        else if (!J.learned)
            ++pk.n_noheld;
    }
    auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
    size_t o = 0;
    pk.off_jobs = o;
    o = align(o + nj * sizeof(BlockJob));
    pk.off_pa = o;
    o = align(o + sc.pa.size() * 4);
    pk.off_pwr = o;
    o = align(o + sc.pwr.size() * 4);
    for (int c = 0; c < kPinvClasses; ++c) {
        pk.off_list[c] = o;
        o = align(o + std::max<size_t>(sc.lists[c].size(), 1) * 4);
        pk.nlist[c] = (int)sc.lists[c].size();
    }
    pk.off_out = o;
    o = align(o + nj * sizeof(JobOut));
    pk.size = o;
    pk.total_all = sc.pa[nj];
    pk.total_wr = sc.pwr[nj];
}

// Every section after the job records (which the caller writes at
// dst + pk.off_jobs = dst): prefixes, class lists, zeroed padding and JobOut slots.
void pack_write_tail(int nj, const PackedRound& pk, const PackScratch& sc, char* dst) {
    size_t at = pk.off_jobs + (size_t)nj * sizeof(BlockJob);
    auto put = [&](size_t off, const void* src, size_t bytes) {
        if (off > at) std::memset(dst + at, 0, off - at);  // padding
        if (bytes) std::memcpy(dst + off, src, bytes);
        at = off + bytes;
    };
    put(pk.off_pa, sc.pa.data(), sc.pa.size() * 4);
    put(pk.off_pwr, sc.pwr.data(), sc.pwr.size() * 4);
    for (int c = 0; c < kPinvClasses; ++c) put(pk.off_list[c], sc.lists[c].data(), sc.lists[c].size() * 4);
    put(pk.off_out, nullptr, 0);
    std::memset(dst + pk.off_out, 0, pk.size - pk.off_out);  // JobOut slots (written by the kernels)
}

void pack_write(const BlockJob* jobs, int nj, const PackedRound& pk, const PackScratch& sc, char* dst) {
    std::memcpy(dst + pk.off_jobs, jobs, (size_t)nj * sizeof(BlockJob));
    pack_write_tail(nj, pk, sc, dst);
}

// Single-job rounds (per-block loop, drop-in, standalone calls).
void pack_round(const std::vector<BlockJob>& jobs, bool pivot, PackedRound& pk, std::vector<char>& bytes) {
    PackScratch sc;
    pack_layout((int)jobs.size(), pivot, pk, sc, [&](int j) { return dims_of(jobs[j]); });
    bytes.resize(pk.size);
    pack_write(jobs.data(), (int)jobs.size(), pk, sc, bytes.data());
}

RoundPlan plan_from(const PackedRound& pk, void* dev, int njobs) {
    char* d = (char*)dev;
    RoundPlan P;
    P.njobs = njobs;
    P.jobs = (const BlockJob*)(d + pk.off_jobs);
    P.prefix_all = (const int32_t*)(d + pk.off_pa);
    P.prefix_wr = (const int32_t*)(d + pk.off_pwr);
    P.total_all = pk.total_all;
    P.total_wr = pk.total_wr;
    for (int c = 0; c < kPinvClasses; ++c) {
        P.list[c] = (const int32_t*)(d + pk.off_list[c]);
        P.nlist[c] = pk.nlist[c];
    }
    P.out = (JobOut*)(d + pk.off_out);
    P.one_buf = pk.all_rpos;
    P.n_noheld = pk.n_noheld;
    P.rmax_held = pk.rmax_held;
    return P;
}

int validate_opts(const farcur_opts* o) {
    if (!o) return fail(FARCUR_EINVAL, "opts is NULL");
    if (!(o->tol >= 0)) return fail(FARCUR_EINVAL, "tol must be >= 0");
    if (!(o->pinv_tol >= 0) || !(o->pinv_rtol >= 0)) return fail(FARCUR_EINVAL, "pinv thresholds must be >= 0");
    if (o->precision != FARCUR_F64 && o->precision != FARCUR_F32) return fail(FARCUR_EINVAL, "bad precision");
    if (o->row_mode != FARCUR_ROWS_UNIFORM && o->row_mode != FARCUR_ROWS_PIVOT) return fail(FARCUR_EINVAL, "bad row_mode");
    if (o->learned_core != FARCUR_CORE_QR && o->learned_core != FARCUR_CORE_SVD)
        return fail(FARCUR_EINVAL, "bad learned_core");
    if (o->r_start < 0 || o->max_rank < 0) return fail(FARCUR_EINVAL, "r_start and max_rank must be >= 0");
    if (o->heldout < 0 || o->heldout > 256) return fail(FARCUR_EINVAL, "heldout must be in [0, 256]");
    if (o->oversample < 0) return fail(FARCUR_EINVAL, "oversample must be >= 0");
    if (!(o->label_eps >= 0)) return fail(FARCUR_EINVAL, "label_eps must be >= 0");
    if (!(o->heldout_safety > 0) || o->heldout_safety > 1) return fail(FARCUR_EINVAL, "heldout_safety must be in (0, 1]");
    if (!(o->growth >= 1.1) || o->growth > 16) return fail(FARCUR_EINVAL, "growth must be in [1.1, 16]");
    return FARCUR_OK;
}

// Next sample count after a round, following cuda_cur_compress3
// (compres_single_block_spheres_for_sweep2.jl:910-914 after the first call,
// :955-967 inside the loop). Returns false when the loop stops at the cap.
bool next_r(int r, int cap, int round, int* rn) {
    if (round == 0) {
        *rn = (2 * r < cap) ? 2 * r : cap;
        return true;
    }
    if (r >= cap) return false;
    *rn = (2 * r <= cap) ? 2 * r : cap;
    return true;
}

}  // namespace

// =================================================================== C ABI
extern "C" {

const char* farcur_last_error(void) { return g_err.c_str(); }
const char* farcur_version(void) { return "farcur 0.1 (gfx950)"; }

void farcur_opts_default(farcur_opts* o) {
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->tol = 1e-4;
    o->pinv_tol = 1e-10;
    o->pinv_rtol = 0.0;
    o->seed = 12345;
    o->precision = FARCUR_F64;
    o->row_mode = FARCUR_ROWS_UNIFORM;
    o->heldout = 16;
    o->oversample = 18;
    o->label_eps = 1e-4;
    o->heldout_safety = 0.3;
    o->growth = 1.5;
    o->learned_core = FARCUR_CORE_QR;
}

// ------------------------------------------------------------------ meshes
int farcur_mesh_plate(double lx, double ly, int nx, int ny, double x0, double y0, double z0, farcur_mesh** out) {
    if (!out || nx <= 0 || ny <= 0 || !(lx > 0) || !(ly > 0)) return fail(FARCUR_EINVAL, "bad plate arguments");
    *out = new farcur_mesh{mesh_plate(lx, ly, nx, ny, x0, y0, z0)};
    return FARCUR_OK;
}

int farcur_mesh_octasphere(double radius, int n_levels, farcur_mesh** out) {
    if (!out || n_levels < 0 || n_levels > 8 || !(radius > 0)) return fail(FARCUR_EINVAL, "bad octasphere arguments");
    *out = new farcur_mesh{mesh_octasphere(radius, n_levels)};
    return FARCUR_OK;
}

int farcur_mesh_icosphere(double ax, double ay, double az, int level, double jitter, uint64_t seed,
                          farcur_mesh** out) {
    if (!out || level < 0 || level > 8 || !(ax > 0) || !(ay > 0) || !(az > 0) || !(jitter >= 0) || jitter >= 0.5)
        return fail(FARCUR_EINVAL, "bad icosphere arguments");
    *out = new farcur_mesh{mesh_icosphere(ax, ay, az, level, jitter, seed)};
    return FARCUR_OK;
}

int farcur_mesh_append(farcur_mesh* dst, const farcur_mesh* src, double dx, double dy, double dz) {
    if (!dst || !src) return fail(FARCUR_EINVAL, "NULL mesh");
    mesh_append(dst->m, src->m, dx, dy, dz);
    dst->rwg = false;
    return FARCUR_OK;
}

int farcur_mesh_build_rwg(farcur_mesh* mesh) {
    if (!mesh) return fail(FARCUR_EINVAL, "NULL mesh");
    std::string err;
    if (!build_rwg(mesh->m, err)) return fail(FARCUR_EINVAL, "%s", err.c_str());
    mesh->rwg = true;
    return FARCUR_OK;
}

int farcur_mesh_sizes(const farcur_mesh* mesh, int* nv, int* nt, int* ne) {
    if (!mesh) return fail(FARCUR_EINVAL, "NULL mesh");
    if (nv) *nv = mesh->m.nv();
    if (nt) *nt = mesh->m.nt();
    if (ne) *ne = mesh->rwg ? mesh->m.ne() : 0;
    return FARCUR_OK;
}

int farcur_mesh_tables(const farcur_mesh* mesh, double* vertex, double* topol, double* trian, double* edges,
                       double* un, double* ds, double* ln, double* cent) {
    if (!mesh) return fail(FARCUR_EINVAL, "NULL mesh");
    const Mesh& m = mesh->m;
    if (vertex) std::copy(m.vertex.begin(), m.vertex.end(), vertex);
    if (topol)
        for (size_t i = 0; i < m.topol.size(); ++i) topol[i] = m.topol[i] + 1;
    if ((trian || edges || un || ds || ln || cent) && !mesh->rwg)
        return fail(FARCUR_ESTATE, "RWG tables not built (call farcur_mesh_build_rwg)");
    if (trian)
        for (size_t i = 0; i < m.trian.size(); ++i) trian[i] = m.trian[i];
    if (edges)
        for (size_t i = 0; i < m.edges.size(); ++i) edges[i] = m.edges[i] + 1;
    if (un) std::copy(m.un.begin(), m.un.end(), un);
    if (ds) std::copy(m.ds.begin(), m.ds.end(), ds);
    if (ln) std::copy(m.ln.begin(), m.ln.end(), ln);
    if (cent) std::copy(m.cent.begin(), m.cent.end(), cent);
    return FARCUR_OK;
}

void farcur_mesh_destroy(farcur_mesh* mesh) { delete mesh; }

int farcur_partition_build(const farcur_mesh* mesh, double leaf_size, double eta_adm, farcur_partition** out) {
    if (!mesh || !out) return fail(FARCUR_EINVAL, "NULL argument");
    if (!mesh->rwg) return fail(FARCUR_ESTATE, "RWG tables not built");
    if (!(leaf_size > 0) || !(eta_adm > 0)) return fail(FARCUR_EINVAL, "leaf_size and eta_adm must be > 0");
    *out = new farcur_partition{build_partition(mesh->m, leaf_size, eta_adm)};
    return FARCUR_OK;
}

int farcur_partition_sizes(const farcur_partition* p, int* ngroups, int* nrwg, int* nfar, int* nnear) {
    if (!p) return fail(FARCUR_EINVAL, "NULL partition");
    if (ngroups) *ngroups = (int)p->p.group_diam.size();
    if (nrwg) *nrwg = (int)p->p.group_rwg.size();
    if (nfar) *nfar = (int)(p->p.far_pairs.size() / 2);
    if (nnear) *nnear = (int)(p->p.near_pairs.size() / 2);
    return FARCUR_OK;
}

int farcur_partition_tables(const farcur_partition* p, int* group_offsets, int* group_rwg, double* centre,
                            double* diam, int* far_pairs, int* near_pairs) {
    if (!p) return fail(FARCUR_EINVAL, "NULL partition");
    const BlockPartition& P = p->p;
    if (group_offsets) std::copy(P.group_offsets.begin(), P.group_offsets.end(), group_offsets);
    if (group_rwg) std::copy(P.group_rwg.begin(), P.group_rwg.end(), group_rwg);
    if (centre) std::copy(P.group_centre.begin(), P.group_centre.end(), centre);
    if (diam) std::copy(P.group_diam.begin(), P.group_diam.end(), diam);
    if (far_pairs) std::copy(P.far_pairs.begin(), P.far_pairs.end(), far_pairs);
    if (near_pairs) std::copy(P.near_pairs.begin(), P.near_pairs.end(), near_pairs);
    return FARCUR_OK;
}

int farcur_partition_blocks(const farcur_partition* p, int* a, int* b, int* comp, int* self) {
    if (!p) return fail(FARCUR_EINVAL, "NULL partition");
    const BlockPartition& P = p->p;
    const size_t nf = P.far_pairs.size() / 2, nn = P.near_pairs.size() / 2;
    for (size_t i = 0; i < nf + nn; ++i) {
        const bool far = i < nf;
        const int32_t ga = far ? P.far_pairs[2 * i] : P.near_pairs[2 * (i - nf)];
        const int32_t gb = far ? P.far_pairs[2 * i + 1] : P.near_pairs[2 * (i - nf) + 1];
        if (a) a[i] = ga;
        if (b) b[i] = gb;
        if (comp) comp[i] = far ? 1 : 0;
        if (self) self[i] = ga == gb ? 1 : 0;
    }
    return FARCUR_OK;
}

int farcur_partition_save(const farcur_partition* p, const char* path) {
    if (!p || !path) return fail(FARCUR_EINVAL, "NULL argument");
    std::string err;
    if (!save_partition(p->p, path, err)) return fail(FARCUR_EINVAL, "%s", err.c_str());
    return FARCUR_OK;
}

int farcur_partition_load(const char* path, farcur_partition** out) {
    if (!path || !out) return fail(FARCUR_EINVAL, "NULL argument");
    *out = nullptr;
    auto p = std::make_unique<farcur_partition>();
    std::string err;
    if (!load_partition(path, p->p, err)) return fail(FARCUR_EINVAL, "%s", err.c_str());
    *out = p.release();
    return FARCUR_OK;
}

int farcur_partition_params(const farcur_partition* p, double* leaf_size, double* eta_adm, int* nrwg_mesh) {
    if (!p) return fail(FARCUR_EINVAL, "NULL partition");
    if (leaf_size) *leaf_size = p->p.leaf_size;
    if (eta_adm) *eta_adm = p->p.eta_adm;
    if (nrwg_mesh) *nrwg_mesh = p->p.nrwg_mesh;
    return FARCUR_OK;
}

void farcur_partition_destroy(farcur_partition* p) { delete p; }

// ------------------------------------------------------------------ context
int farcur_create(int device, farcur_ctx** out) {
    if (!out) return fail(FARCUR_EINVAL, "out is NULL");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(FARCUR_EHIP, "no HIP device available");
    if (device < 0 || device >= n) return fail(FARCUR_EINVAL, "device %d out of range (have %d)", device, n);
    auto ctx = std::make_unique<farcur_ctx>();
    ctx->device = device;
    HIPTRY(hipSetDevice(device));
    // device address of the allocation-free pinv's static scratch, once per device
    if (const int e = noalloc_init(device))
        return fail(FARCUR_EHIP, "hipGetSymbolAddress of the no-allocation scratch failed (%d)", e);
    HIPTRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    HIPTRY(hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking));
    HIPTRY(hipEventCreate(&ctx->ev0));
    HIPTRY(hipEventCreate(&ctx->ev1));
    HIPTRY(hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming));
    HIPTRY(hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming));
    HIPTRY(hipEventCreateWithFlags(&ctx->ev_pfork, hipEventDisableTiming));
    HIPTRY(hipEventCreateWithFlags(&ctx->ev_pjoin, hipEventDisableTiming));
    *out = ctx.release();
    return FARCUR_OK;
}

void farcur_destroy(farcur_ctx* ctx) {
    if (!ctx) return;
    hipSetDevice(ctx->device);
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    DBuf* bufs[] = {&ctx->rec64, &ctx->rec32, &ctx->cen, &ctx->tri32, &ctx->tcen, &ctx->lrec32, &ctx->rtri, &ctx->d_rows, &ctx->d_cols, &ctx->d_cs,    &ctx->d_rs, &ctx->d_rpos,
                    &ctx->d_C,   &ctx->d_R,   &ctx->d_W,    &ctx->d_V,    &ctx->d_U,     &ctx->d_sigma,
                    &ctx->d_v,   &ctx->d_y0,  &ctx->d_y1,   &ctx->d_plan, &ctx->d_out};
    for (DBuf* b : bufs) b->release();
    if (ctx->stream2) hipStreamSynchronize(ctx->stream2);
    if (ctx->ev0) hipEventDestroy(ctx->ev0);
    if (ctx->ev1) hipEventDestroy(ctx->ev1);
    if (ctx->ev_fork) hipEventDestroy(ctx->ev_fork);
    if (ctx->ev_join) hipEventDestroy(ctx->ev_join);
    if (ctx->ev_pfork) hipEventDestroy(ctx->ev_pfork);
    if (ctx->ev_pjoin) hipEventDestroy(ctx->ev_pjoin);
    if (ctx->stream2) hipStreamDestroy(ctx->stream2);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
}

int farcur_set_em(farcur_ctx* ctx, double k, double eta) {
    if (!ctx) return fail(FARCUR_EINVAL, "NULL context");
    if (!(k > 0) || !(eta > 0) || !std::isfinite(k) || !std::isfinite(eta))
        return fail(FARCUR_EINVAL, "k and eta must be positive finite");
    ctx->k = k;
    ctx->eta = eta;
    return FARCUR_OK;
}

int farcur_set_geometry(farcur_ctx* ctx, const double* vertex, int nv, const double* topol, int nt,
                        const double* edges, int ne, const double* ln) {
    if (!ctx || !vertex || !topol || !edges || !ln) return fail(FARCUR_EINVAL, "NULL argument");
    if (nv <= 0 || nt <= 0 || ne <= 0) return fail(FARCUR_EINVAL, "empty geometry");
    std::vector<double> V(vertex, vertex + 3 * (size_t)nv), L(ln, ln + ne);
    std::vector<int32_t> T(3 * (size_t)nt), E(4 * (size_t)ne);
    for (size_t i = 0; i < T.size(); ++i) T[i] = (int32_t)std::lround(topol[i]) - 1;
    for (size_t i = 0; i < E.size(); ++i) E[i] = (int32_t)std::lround(edges[i]) - 1;
    return set_geometry_impl(ctx, V, T, E, L);
}

int farcur_set_mesh(farcur_ctx* ctx, const farcur_mesh* mesh) {
    if (!ctx || !mesh) return fail(FARCUR_EINVAL, "NULL argument");
    if (!mesh->rwg) return fail(FARCUR_ESTATE, "RWG tables not built");
    return set_geometry_impl(ctx, mesh->m.vertex, mesh->m.topol, mesh->m.edges, mesh->m.ln);
}

int farcur_fill(farcur_ctx* ctx, const int* rows, int nr, const int* cols, int nc, int precision, farcur_c128* Z) {
    if (!ctx || !rows || !cols || !Z) return fail(FARCUR_EINVAL, "NULL argument");
    if (ctx->ne == 0) return fail(FARCUR_ESTATE, "no geometry");
    if (nr <= 0 || nc <= 0) return fail(FARCUR_EINVAL, "empty panel");
    if (precision != FARCUR_F64 && precision != FARCUR_F32) return fail(FARCUR_EINVAL, "bad precision");
    TRY(check_ids(rows, nr, ctx->ne, "rows"));
    TRY(check_ids(cols, nc, ctx->ne, "cols"));
    TRY(ctx->select());
    const int f32 = precision == FARCUR_F32;
    TRY(ctx->d_rows.ensure(nr * 4));
    TRY(ctx->d_cols.ensure(nc * 4));
    TRY(ctx->d_C.ensure((size_t)nr * nc * psize(f32)));
    HIPTRY(hipMemcpyAsync(ctx->d_rows.p, rows, nr * 4, hipMemcpyHostToDevice, ctx->stream));
    HIPTRY(hipMemcpyAsync(ctx->d_cols.p, cols, nc * 4, hipMemcpyHostToDevice, ctx->stream));
    launch_fill_panel(ctx->d_C.p, nr, ctx->d_rows.as<int32_t>(), nr, ctx->d_cols.as<int32_t>(), nc,
                      ctx->fill_params(f32), ctx->stream);
    HIPTRY(hipGetLastError());
    if (f32) {
        std::vector<c64> tmp((size_t)nr * nc);
        HIPTRY(hipMemcpyAsync(tmp.data(), ctx->d_C.p, tmp.size() * sizeof(c64), hipMemcpyDeviceToHost, ctx->stream));
        HIPTRY(hipStreamSynchronize(ctx->stream));
        for (size_t i = 0; i < tmp.size(); ++i) Z[i] = farcur_c128{tmp[i].re, tmp[i].im};
    } else {
        HIPTRY(hipMemcpyAsync(Z, ctx->d_C.p, (size_t)nr * nc * sizeof(c128), hipMemcpyDeviceToHost, ctx->stream));
        HIPTRY(hipStreamSynchronize(ctx->stream));
    }
    return FARCUR_OK;
}

int farcur_pinv(farcur_ctx* ctx, const farcur_c128* W, int r, double atol, double rtol, farcur_c128* U,
                double* sigma, int* kept) {
    if (!ctx || !W || !U) return fail(FARCUR_EINVAL, "NULL argument");
    if (r <= 0 || r > kMaxSamples) return fail(FARCUR_EINVAL, "r = %d outside [1, %d]", r, kMaxSamples);
    if (!(atol >= 0) || !(rtol >= 0)) return fail(FARCUR_EINVAL, "thresholds must be >= 0");
    TRY(ctx->select());
    const size_t rr = (size_t)r * r * sizeof(c128);
    TRY(ctx->d_W.ensure(rr));
    TRY(ctx->d_V.ensure(pinv_v_elems(r) * sizeof(c128)));
    TRY(ctx->d_U.ensure(rr));
    TRY(ctx->d_sigma.ensure(r * sizeof(double)));
    HIPTRY(hipMemcpyAsync(ctx->d_W.p, W, rr, hipMemcpyHostToDevice, ctx->stream));
    BlockJob J{};
    J.r = r;
    J.W = ctx->d_W.as<c128>();
    J.V = ctx->d_V.as<c128>();
    J.U = ctx->d_U.as<c128>();
    J.sigma = ctx->d_sigma.as<double>();
    struct Plan {
        BlockJob job;
        int32_t list[4];
        JobOut out;
    } plan{};
    plan.job = J;
    TRY(ctx->d_plan.ensure(sizeof(Plan)));
    HIPTRY(hipMemcpyAsync(ctx->d_plan.p, &plan, sizeof(Plan), hipMemcpyHostToDevice, ctx->stream));
    Plan* dp = ctx->d_plan.as<Plan>();
    const int cls = pinv_class(r);
    const int threads = cls < kPinvClasses - 1 ? ctx->tune.pinv_threads[cls] : (r > 128 ? 1024 : 256);
    launch_pinv(&dp->job, dp->list, 1, atol, rtol, 0.0, kPinvClassMax[cls], threads, 0, 0, 0, 0, &dp->out,
                ctx->stream);
    HIPTRY(hipGetLastError());
    HIPTRY(hipMemcpyAsync(U, ctx->d_U.p, rr, hipMemcpyDeviceToHost, ctx->stream));
    std::vector<double> s(r);
    HIPTRY(hipMemcpyAsync(s.data(), ctx->d_sigma.p, r * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPTRY(hipMemcpyAsync(&plan, ctx->d_plan.p, sizeof(Plan), hipMemcpyDeviceToHost, ctx->stream));
    HIPTRY(hipStreamSynchronize(ctx->stream));
    std::sort(s.begin(), s.end(), [](double a, double b) { return a > b; });
    if (sigma) std::copy(s.begin(), s.end(), sigma);
    if (kept) *kept = plan.out.kept;
    return FARCUR_OK;
}

int farcur_pinv_qr(farcur_ctx* ctx, const farcur_c128* W, int r, double atol, double rtol, farcur_c128* U,
                   double* diag, int* kept) {
    if (!ctx || !W || !U) return fail(FARCUR_EINVAL, "NULL argument");
    if (r <= 0 || r > kPinvLdsOneBufMaxR)
        return fail(FARCUR_EINVAL, "r = %d outside [1, %d] (the learned core's LDS form)", r, kPinvLdsOneBufMaxR);
    if (!(atol >= 0) || !(rtol >= 1e-8))
        return fail(FARCUR_EINVAL, "thresholds: atol >= 0 and rtol >= 1e-8 (the learned core's LDS form)");
    TRY(ctx->select());
    const size_t rr = (size_t)r * r * sizeof(c128);
    TRY(ctx->d_W.ensure(rr));
    TRY(ctx->d_U.ensure(rr));
    TRY(ctx->d_sigma.ensure(r * sizeof(double)));
    HIPTRY(hipMemcpyAsync(ctx->d_W.p, W, rr, hipMemcpyHostToDevice, ctx->stream));
    // the core is gathered as in a learned round: W = C[rpos, :] with C = W (m = r)
    struct Plan {
        BlockJob job;
        int32_t list[4];
        JobOut out;
        int32_t rpos[kPinvLdsOneBufMaxR];
    } plan{};
    for (int i = 0; i < r; ++i) plan.rpos[i] = i;
    TRY(ctx->d_plan.ensure(sizeof(Plan)));
    Plan* dp = ctx->d_plan.as<Plan>();
    BlockJob& J = plan.job;
    J.m = J.n = J.r = r;
    J.C = ctx->d_W.p;
    J.W = ctx->d_W.as<c128>();
    J.U = ctx->d_U.as<c128>();
    J.sigma = ctx->d_sigma.as<double>();
    J.rpos = dp->rpos;
    HIPTRY(hipMemcpyAsync(ctx->d_plan.p, &plan, sizeof(Plan), hipMemcpyHostToDevice, ctx->stream));
    const int cls = pinv_class(r);
    launch_pinv(&dp->job, dp->list, 1, atol, rtol, 0.0, kPinvClassMax[cls], ctx->tune.pinv_threads[cls], 0, 1, 0, 2,
                &dp->out, ctx->stream);
    HIPTRY(hipGetLastError());
    HIPTRY(hipMemcpyAsync(U, ctx->d_U.p, rr, hipMemcpyDeviceToHost, ctx->stream));
    std::vector<double> d(r);
    HIPTRY(hipMemcpyAsync(d.data(), ctx->d_sigma.p, r * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPTRY(hipMemcpyAsync(&plan, ctx->d_plan.p, sizeof(Plan), hipMemcpyDeviceToHost, ctx->stream));
    HIPTRY(hipStreamSynchronize(ctx->stream));
    if (diag) std::copy(d.begin(), d.end(), diag);
    if (kept) *kept = plan.out.kept;
    return FARCUR_OK;
}

int farcur_pinv_noalloc(farcur_ctx* ctx, const farcur_c128* W, int r, double atol, double rtol, farcur_c128* U,
                        double* sigma, int* kept, void* stream) {
    if (!ctx || !W || !U) return fail(FARCUR_EINVAL, "NULL argument");
    if (r <= 0 || r > kNoAllocMaxR)
        return fail(FARCUR_EINVAL, "r = %d outside [1, %d], the largest core the no-allocation scratch holds", r,
                    kNoAllocMaxR);
    if (!(atol >= 0) || !(rtol >= 0)) return fail(FARCUR_EINVAL, "thresholds must be >= 0");
    unsigned char* base = noalloc_scratch(ctx->device);
    if (!base) return fail(FARCUR_ESTATE, "no-allocation scratch not set up on device %d", ctx->device);
    TRY(ctx->select());
    // the same class, launch shape and threshold semantics as farcur_pinv
    const int cls = pinv_class(r);
    const int threads = cls < kPinvClasses - 1 ? ctx->tune.pinv_threads[cls] : (r > 128 ? 1024 : 256);
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    launch_pinv_noalloc(base, (const c128*)W, r, atol, rtol, kPinvClassMax[cls], threads, (c128*)U, sigma, kept, s);
    HIPTRY(hipGetLastError());
    return FARCUR_OK;
}

int farcur_cur_residual(farcur_ctx* ctx, const farcur_c128* C, int m, int r, const farcur_c128* U,
                        const farcur_c128* R, int n, const farcur_c128* v, const farcur_c128* y_old,
                        farcur_c128* y_new, double* err) {
    if (!ctx || !C || !U || !R || !v || !y_old || !y_new || !err) return fail(FARCUR_EINVAL, "NULL argument");
    if (m <= 0 || n <= 0 || r <= 0 || r > kMaxSamples) return fail(FARCUR_EINVAL, "bad sizes m=%d n=%d r=%d", m, n, r);
    TRY(ctx->select());
    TRY(ctx->d_C.ensure((size_t)m * r * 16));
    TRY(ctx->d_U.ensure((size_t)r * r * 16));
    TRY(ctx->d_R.ensure((size_t)r * n * 16));
    TRY(ctx->d_v.ensure((size_t)n * 16));
    TRY(ctx->d_y0.ensure((size_t)m * 16));
    TRY(ctx->d_y1.ensure((size_t)m * 16));
    TRY(ctx->d_W.ensure((size_t)std::max(r * r, 2 * r) * 16));
    hipStream_t s = ctx->stream;
    HIPTRY(hipMemcpyAsync(ctx->d_C.p, C, (size_t)m * r * 16, hipMemcpyHostToDevice, s));
    HIPTRY(hipMemcpyAsync(ctx->d_U.p, U, (size_t)r * r * 16, hipMemcpyHostToDevice, s));
    HIPTRY(hipMemcpyAsync(ctx->d_R.p, R, (size_t)r * n * 16, hipMemcpyHostToDevice, s));
    HIPTRY(hipMemcpyAsync(ctx->d_v.p, v, (size_t)n * 16, hipMemcpyHostToDevice, s));
    HIPTRY(hipMemcpyAsync(ctx->d_y0.p, y_old, (size_t)m * 16, hipMemcpyHostToDevice, s));
    struct Plan {
        BlockJob job;
        JobOut out;
    } plan{};
    plan.job.m = m;
    plan.job.n = n;
    plan.job.r = r;
    plan.job.C = ctx->d_C.p;
    plan.job.R = ctx->d_R.p;
    plan.job.U = ctx->d_U.as<c128>();
    plan.job.W = ctx->d_W.as<c128>();
    plan.job.v = ctx->d_v.as<c128>();
    plan.job.y_old = ctx->d_y0.as<c128>();
    plan.job.y_new = ctx->d_y1.as<c128>();
    TRY(ctx->d_plan.ensure(sizeof(Plan)));
    HIPTRY(hipMemcpyAsync(ctx->d_plan.p, &plan, sizeof(Plan), hipMemcpyHostToDevice, s));
    Plan* dp = ctx->d_plan.as<Plan>();
    launch_residual(&dp->job, 1, 0, 1024, 0, &dp->out, s);
    HIPTRY(hipGetLastError());
    HIPTRY(hipMemcpyAsync(y_new, ctx->d_y1.p, (size_t)m * 16, hipMemcpyDeviceToHost, s));
    HIPTRY(hipMemcpyAsync(&plan, ctx->d_plan.p, sizeof(Plan), hipMemcpyDeviceToHost, s));
    HIPTRY(hipStreamSynchronize(s));
    *err = plan.out.err;
    return FARCUR_OK;
}

int farcur_sample(farcur_ctx* ctx, uint64_t seed, int block, int round, int purpose, int N, int r, int* out) {
    if (!ctx || !out) return fail(FARCUR_EINVAL, "NULL argument");
    if (N <= 0 || r <= 0 || r > N) return fail(FARCUR_EINVAL, "need 0 < r <= N (r=%d N=%d)", r, N);
    if (purpose != kStreamCols && purpose != kStreamRows) return fail(FARCUR_EINVAL, "purpose must be 1 or 2");
    if (block < 0 || round < 0 || round > 0xffffff) return fail(FARCUR_EINVAL, "bad block/round");
    TRY(ctx->select());
    std::vector<int32_t> ids(N);
    for (int i = 0; i < N; ++i) ids[i] = i;
    TRY(ctx->d_rows.ensure(N * 4));
    TRY(ctx->d_cs.ensure(r * 4));
    HIPTRY(hipMemcpyAsync(ctx->d_rows.p, ids.data(), N * 4, hipMemcpyHostToDevice, ctx->stream));
    BlockJob J{};
    J.b = block;
    J.round = round;
    J.r = r;
    J.sample_mode = 0;
    if (purpose == kStreamCols) {
        J.n = N;
        J.cols = ctx->d_rows.as<int32_t>();
        J.cs = ctx->d_cs.as<int32_t>();
        J.m = 0;
        J.rs = nullptr;
    } else {
        J.m = N;
        J.rows = ctx->d_rows.as<int32_t>();
        J.rs = ctx->d_cs.as<int32_t>();
        J.n = 0;
        J.cs = nullptr;
    }
    TRY(ctx->d_plan.ensure(sizeof(BlockJob)));
    HIPTRY(hipMemcpyAsync(ctx->d_plan.p, &J, sizeof J, hipMemcpyHostToDevice, ctx->stream));
    // k_sample runs both kinds for the job; the kind without candidates draws nothing.
    launch_sample(ctx->d_plan.as<BlockJob>(), 1, seed, ctx->stream);
    HIPTRY(hipGetLastError());
    HIPTRY(hipMemcpyAsync(out, ctx->d_cs.p, r * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPTRY(hipStreamSynchronize(ctx->stream));
    return FARCUR_OK;
}

int farcur_vtest(uint64_t seed, int block, int n, farcur_c128* v) {
    if (!v || n < 0 || block < 0) return fail(FARCUR_EINVAL, "bad arguments");
    for (int j = 0; j < n; ++j) vtest_entry(seed, (uint32_t)block, (uint32_t)j, &v[j].re, &v[j].im);
    return FARCUR_OK;
}

int farcur_pivot_rows(farcur_ctx* ctx, const farcur_c128* C, int m, int r, int* rows) {
    if (!ctx || !C || !rows) return fail(FARCUR_EINVAL, "NULL argument");
    if (m <= 0 || r <= 0) return fail(FARCUR_EINVAL, "bad sizes");
    TRY(ctx->select());
    std::vector<int32_t> ids(m);
    for (int i = 0; i < m; ++i) ids[i] = i;
    TRY(ctx->d_rows.ensure(m * 4));
    TRY(ctx->d_rs.ensure(r * 4));
    TRY(ctx->d_C.ensure((size_t)m * r * 16));
    HIPTRY(hipMemcpyAsync(ctx->d_rows.p, ids.data(), m * 4, hipMemcpyHostToDevice, ctx->stream));
    HIPTRY(hipMemcpyAsync(ctx->d_C.p, C, (size_t)m * r * 16, hipMemcpyHostToDevice, ctx->stream));
    BlockJob J{};
    J.m = m;
    J.r = r;
    J.sample_mode = 2;
    J.rows = ctx->d_rows.as<int32_t>();
    J.rs = ctx->d_rs.as<int32_t>();
    J.C = ctx->d_C.p;
    TRY(ctx->d_plan.ensure(sizeof(BlockJob)));
    HIPTRY(hipMemcpyAsync(ctx->d_plan.p, &J, sizeof J, hipMemcpyHostToDevice, ctx->stream));
    launch_pivot_rows(ctx->d_plan.as<BlockJob>(), 1, 0, ctx->stream);
    HIPTRY(hipGetLastError());
    HIPTRY(hipMemcpyAsync(rows, ctx->d_rs.p, r * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPTRY(hipStreamSynchronize(ctx->stream));
    return FARCUR_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ per-block loop
namespace {

// Device buffers of one block's round.
struct RoundBufs {
    int32_t *cs, *rs;
    int32_t* rpos;  // sampled row positions (W gathered from C); NULL for given samples
    void *C, *R;
    c128 *W, *V, *U;
    double* sigma;
    int32_t *hs, *hpos;  // held-out rows (learned mode)
    void* H;
};

struct BlockLoopIO {
    const int32_t* d_rows;
    const int32_t* d_cols;
    const c128* d_v;
    c128* d_y[2];
    int m, n, b;
};

// Runs the cuda_cur_compress3 loop (SW2:878-973) for one block with a sync per
// round. `alloc` returns device buffers for a round of sample count r.
template <typename Alloc>
int block_loop(farcur_ctx* ctx, const BlockLoopIO& io, const farcur_opts& o, Alloc alloc, DBuf& plan_buf,
               KTimer& kt, farcur_block_info* info, RoundBufs* final_bufs, int* final_y) {
    const int D = std::min(io.m, io.n);
    const int cap = o.max_rank > 0 ? std::min(D, o.max_rank) : D;
    if (cap > kMaxSamples) return fail(FARCUR_EINVAL, "sample count cap %d exceeds %d (set max_rank)", cap, kMaxSamples);
    int r = o.r_start > 0 ? std::min(o.r_start, cap) : std::min(r_start_rule(D), cap);
    const int f32 = o.precision == FARCUR_F32;
    int round = 0, ycur = 0;
    double err = INFINITY;
    JobOut res{};
    RoundBufs rb{};
    double dev_ms = 0;
    int final_r = r;
    bool converged = false;
    while (true) {
        TRY(alloc(r, rb));
        BlockJob J{};
        J.b = io.b;
        J.m = io.m;
        J.n = io.n;
        J.r = r;
        J.round = round;
        J.sample_mode = o.row_mode == FARCUR_ROWS_PIVOT ? 2 : 0;
        J.rows = io.d_rows;
        J.cols = io.d_cols;
        J.cs = rb.cs;
        J.rs = rb.rs;
        J.rpos = rb.rpos;
        J.C = rb.C;
        J.R = rb.R;
        J.W = rb.W;
        J.V = rb.V;
        J.U = rb.U;
        J.sigma = rb.sigma;
        J.v = io.d_v;
        J.y_old = io.d_y[ycur];
        J.y_new = io.d_y[ycur ^ 1];
        std::vector<BlockJob> jobs{J};
        PackedRound pk;
        std::vector<char> bytes;
        pack_round(jobs, o.row_mode == FARCUR_ROWS_PIVOT, pk, bytes);
        TRY(plan_buf.ensure(bytes.size()));
        HIPTRY(hipMemcpyAsync(plan_buf.p, bytes.data(), bytes.size(), hipMemcpyHostToDevice, ctx->stream));
        RoundPlan P = plan_from(pk, plan_buf.p, 1);
        P.f32 = f32;
        P.pivot = o.row_mode == FARCUR_ROWS_PIVOT;
        P.atol = o.pinv_tol;
        P.rtol = o.pinv_rtol;
        P.seed = o.seed;
        const bool big = (size_t)io.m * r > (size_t)256 * 1024 || r > 128;
        P.tune(ctx->tune);
        P.label_eps = o.label_eps;
        if (r > kPinvLdsMaxR) P.pinv_threads[kPinvClasses - 2] = P.pinv_threads[kPinvClasses - 1] = r > 128 ? 1024 : 256;
        if (big) P.resid_threads = 1024;
        HIPTRY(hipEventRecord(ctx->ev0, ctx->stream));
        run_round(ctx, P, kt, ctx->stream);
        HIPTRY(hipEventRecord(ctx->ev1, ctx->stream));
        HIPTRY(hipGetLastError());
        HIPTRY(hipMemcpyAsync(&res, P.out, sizeof res, hipMemcpyDeviceToHost, ctx->stream));
        HIPTRY(hipStreamSynchronize(ctx->stream));
        float ms = 0;
        hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
        dev_ms += ms;
        kt.collect();
        ycur ^= 1;
        final_r = r;
        *final_bufs = rb;
        if (round == 0) {
            // The first round's error compares against ones(M): discarded (SW2:902-906).
            int rn;
            next_r(r, cap, 0, &rn);
            r = rn;
            round = 1;
            continue;
        }
        err = res.err;
        ++round;
        // The reference loop runs while err > tol: a NaN residual also ends it,
        // but is not reported as converged.
        if (!(err > o.tol)) {
            converged = err <= o.tol;
            break;
        }
        int rn;
        if (!next_r(r, cap, round, &rn)) break;
        r = rn;
    }
    info->r = final_r;
    info->rounds = round;
    info->kept = res.kept;
    info->converged = converged ? 1 : 0;
    info->err = err;
    info->heldout_err = -1;
    info->sigma_max = res.sigma_max;
    info->time_ms = dev_ms;
    info->rank_eps = res.rank_eps;
    *final_y = ycur;
    return FARCUR_OK;
}

}  // namespace

extern "C" int farcur_compress_block(farcur_ctx* ctx, const int* rows, int m, const int* cols, int n,
                                     const farcur_opts* opts, farcur_block_info* info) {
    if (!ctx || !rows || !cols || !info) return fail(FARCUR_EINVAL, "NULL argument");
    TRY(validate_opts(opts));
    if (ctx->ne == 0) return fail(FARCUR_ESTATE, "no geometry");
    if (m <= 0 || n <= 0) return fail(FARCUR_EINVAL, "empty block");
    TRY(check_ids(rows, m, ctx->ne, "rows"));
    TRY(check_ids(cols, n, ctx->ne, "cols"));
    {
        FarCheck fc;
        TRY(check_far_block(ctx, rows, m, cols, n, 0.0, fc, "block", opts->block_id));
    }
    TRY(ctx->select());
    ctx->has_result = false;
    const int f32 = opts->precision == FARCUR_F32;
    TRY(ctx->d_rows.ensure(m * 4));
    TRY(ctx->d_cols.ensure(n * 4));
    TRY(ctx->d_v.ensure((size_t)n * 16));
    TRY(ctx->d_y0.ensure((size_t)m * 16));
    TRY(ctx->d_y1.ensure((size_t)m * 16));
    HIPTRY(hipMemcpyAsync(ctx->d_rows.p, rows, m * 4, hipMemcpyHostToDevice, ctx->stream));
    HIPTRY(hipMemcpyAsync(ctx->d_cols.p, cols, n * 4, hipMemcpyHostToDevice, ctx->stream));
    std::vector<c128> v(n);
    for (int j = 0; j < n; ++j) vtest_entry(opts->seed, (uint32_t)opts->block_id, (uint32_t)j, &v[j].re, &v[j].im);
    HIPTRY(hipMemcpyAsync(ctx->d_v.p, v.data(), (size_t)n * 16, hipMemcpyHostToDevice, ctx->stream));
    launch_ones(ctx->d_y0.as<c128>(), m, ctx->stream);
    BlockLoopIO io{ctx->d_rows.as<int32_t>(), ctx->d_cols.as<int32_t>(), ctx->d_v.as<c128>(),
                   {ctx->d_y0.as<c128>(), ctx->d_y1.as<c128>()}, m, n, opts->block_id};
    auto alloc = [&](int r, RoundBufs& rb) -> int {
        TRY(ctx->d_cs.ensure(r * 4));
        TRY(ctx->d_rs.ensure(r * 4));
        TRY(ctx->d_rpos.ensure(r * 4));
        TRY(ctx->d_C.ensure((size_t)m * r * psize(f32)));
        TRY(ctx->d_R.ensure((size_t)r * n * psize(f32)));
        TRY(ctx->d_W.ensure((size_t)r * r * 16));
        TRY(ctx->d_V.ensure(pinv_v_elems(r) * 16));
        TRY(ctx->d_U.ensure((size_t)r * r * 16));
        TRY(ctx->d_sigma.ensure(r * 8));
        rb = RoundBufs{};
        rb.cs = ctx->d_cs.as<int32_t>();
        rb.rs = ctx->d_rs.as<int32_t>();
        rb.rpos = ctx->d_rpos.as<int32_t>();
        rb.C = ctx->d_C.p;
        rb.R = ctx->d_R.p;
        rb.W = ctx->d_W.as<c128>();
        rb.V = ctx->d_V.as<c128>();
        rb.U = ctx->d_U.as<c128>();
        rb.sigma = ctx->d_sigma.as<double>();
        return FARCUR_OK;
    };
    KTimer kt;
    RoundBufs fb{};
    int fy = 0;
    TRY(block_loop(ctx, io, *opts, alloc, ctx->d_plan, kt, info, &fb, &fy));
    ctx->last_m = m;
    ctx->last_n = n;
    ctx->last_r = info->r;
    ctx->last_f32 = f32;
    ctx->last_y = fy;
    ctx->has_result = true;
    return FARCUR_OK;
}

namespace {
int copy_panel(farcur_c128* dst, const void* src, size_t count, int f32, hipStream_t s) {
    if (!dst) return FARCUR_OK;
    if (f32) {
        std::vector<c64> tmp(count);
        HIPTRY(hipMemcpyAsync(tmp.data(), src, count * sizeof(c64), hipMemcpyDeviceToHost, s));
        HIPTRY(hipStreamSynchronize(s));
        for (size_t i = 0; i < count; ++i) dst[i] = farcur_c128{tmp[i].re, tmp[i].im};
    } else {
        HIPTRY(hipMemcpyAsync(dst, src, count * sizeof(c128), hipMemcpyDeviceToHost, s));
        HIPTRY(hipStreamSynchronize(s));
    }
    return FARCUR_OK;
}
}  // namespace

extern "C" int farcur_block_result(farcur_ctx* ctx, farcur_c128* C, farcur_c128* U, farcur_c128* R,
                                   int* row_samples, int* col_samples, farcur_c128* y) {
    if (!ctx) return fail(FARCUR_EINVAL, "NULL context");
    if (!ctx->has_result) return fail(FARCUR_ESTATE, "no compress_block result");
    TRY(ctx->select());
    const int m = ctx->last_m, n = ctx->last_n, r = ctx->last_r, f32 = ctx->last_f32;
    hipStream_t s = ctx->stream;
    TRY(copy_panel(C, ctx->d_C.p, (size_t)m * r, f32, s));
    TRY(copy_panel(R, ctx->d_R.p, (size_t)r * n, f32, s));
    TRY(copy_panel(U, ctx->d_U.p, (size_t)r * r, 0, s));
    TRY(copy_panel(y, ctx->last_y ? ctx->d_y1.p : ctx->d_y0.p, (size_t)m, 0, s));
    if (row_samples) HIPTRY(hipMemcpyAsync(row_samples, ctx->d_rs.p, r * 4, hipMemcpyDeviceToHost, s));
    if (col_samples) HIPTRY(hipMemcpyAsync(col_samples, ctx->d_cs.p, r * 4, hipMemcpyDeviceToHost, s));
    HIPTRY(hipStreamSynchronize(s));
    return FARCUR_OK;
}

// ------------------------------------------------------------------ drop-in
namespace {
std::mutex g_compat_mu;
std::vector<farcur_ctx*> g_compat_ctx;  // one per device, kept for the process lifetime

int compat_ctx(farcur_ctx** out) {
    int dev = 0;
    HIPTRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_compat_mu);
    if ((int)g_compat_ctx.size() <= dev) g_compat_ctx.resize(dev + 1, nullptr);
    if (!g_compat_ctx[dev]) TRY(farcur_create(dev, &g_compat_ctx[dev]));
    *out = g_compat_ctx[dev];
    return FARCUR_OK;
}

int to_ids(const double* src, int n, int ne, std::vector<int32_t>& out, const char* what) {
    out.resize(n);
    for (int i = 0; i < n; ++i) {
        const double x = src[i];
        if (!(x >= 1 && x <= ne) || x != std::floor(x))
            return fail(FARCUR_EINVAL, "%s[%d] = %g is not a 1-based RWG id in [1, %d]", what, i, x, ne);
        out[i] = (int32_t)x - 1;
    }
    return FARCUR_OK;
}

int compat_impl(farcur_c128* Cc, farcur_c128* Uc, farcur_c128* Rc, int r, double* vec_time, double pinv_tol,
                farcur_c128* y_old, farcur_c128* y_new, farcur_c128* vtest, double* col_s, double* row_s,
                int field, double k, double eta, double rnear, double* r1, double* r2, double* vertex, double* topol,
                double* edges, double* ln, int n_r1, int n_r2, int rows_vertex, int cols_vertex, int rows_topol,
                int cols_topol, int rows_edges, int cols_edges, int n_ln, double* err_out) {
    if (!Cc || !Uc || !Rc || !vec_time || !y_old || !y_new || !vtest || !col_s || !row_s || !r1 || !r2 || !vertex ||
        !topol || !edges || !ln)
        return fail(FARCUR_EINVAL, "NULL argument");
    if (field != 1) return fail(FARCUR_EINVAL, "field = %d: only the EFIE (field = 1) is implemented", field);
    if (rows_vertex != 3 || rows_topol != 3 || rows_edges != 4)
        return fail(FARCUR_EINVAL, "table shapes must be vertex 3xNv, topol 3xNt, edges 4xNe");
    if (cols_edges != n_ln) return fail(FARCUR_EINVAL, "edges has %d columns but ln has %d entries", cols_edges, n_ln);
    if (n_r1 <= 0 || n_r2 <= 0 || r <= 0) return fail(FARCUR_EINVAL, "empty block or cur_order");
    if (r > std::min(n_r1, n_r2)) return fail(FARCUR_EINVAL, "cur_order %d exceeds min(n_r1, n_r2)", r);
    if (r > kMaxSamples) return fail(FARCUR_EINVAL, "cur_order %d exceeds %d", r, kMaxSamples);
    farcur_ctx* ctx;
    TRY(compat_ctx(&ctx));
    TRY(ctx->select());
    // Geometry cache keyed by the table contents (the reference re-uploads per call).
    uint64_t h = 0xcbf29ce484222325ull;
    h = fnv1a(h, vertex, sizeof(double) * 3 * (size_t)cols_vertex);
    h = fnv1a(h, topol, sizeof(double) * 3 * (size_t)cols_topol);
    h = fnv1a(h, edges, sizeof(double) * 4 * (size_t)cols_edges);
    h = fnv1a(h, ln, sizeof(double) * (size_t)n_ln);
    if (h != ctx->geom_hash || ctx->ne != n_ln) {
        TRY(farcur_set_geometry(ctx, vertex, cols_vertex, topol, cols_topol, edges, cols_edges, ln));
        ctx->geom_hash = h;
    }
    TRY(farcur_set_em(ctx, k, eta));
    std::vector<int32_t> rows, cols, cs, rs;
    TRY(to_ids(r1, n_r1, ctx->ne, rows, "r1"));
    TRY(to_ids(r2, n_r2, ctx->ne, cols, "r2"));
    TRY(to_ids(col_s, r, ctx->ne, cs, "col_samples"));
    TRY(to_ids(row_s, r, ctx->ne, rs, "row_samples"));
    {
        FarCheck fc;
        TRY(check_far_block(ctx, rows.data(), n_r1, cols.data(), n_r2, rnear, fc, "block", 0));
    }
    hipStream_t s = ctx->stream;
    const int m = n_r1, n = n_r2;
    TRY(ctx->d_rows.ensure(m * 4));
    TRY(ctx->d_cols.ensure(n * 4));
    TRY(ctx->d_cs.ensure(r * 4));
    TRY(ctx->d_rs.ensure(r * 4));
    TRY(ctx->d_C.ensure((size_t)m * r * 16));
    TRY(ctx->d_R.ensure((size_t)r * n * 16));
    TRY(ctx->d_W.ensure((size_t)r * r * 16));
    TRY(ctx->d_V.ensure(pinv_v_elems(r) * 16));
    TRY(ctx->d_U.ensure((size_t)r * r * 16));
    TRY(ctx->d_sigma.ensure(r * 8));
    TRY(ctx->d_v.ensure((size_t)n * 16));
    TRY(ctx->d_y0.ensure((size_t)m * 16));
    TRY(ctx->d_y1.ensure((size_t)m * 16));
    HIPTRY(hipMemcpyAsync(ctx->d_rows.p, rows.data(), m * 4, hipMemcpyHostToDevice, s));
    HIPTRY(hipMemcpyAsync(ctx->d_cols.p, cols.data(), n * 4, hipMemcpyHostToDevice, s));
    HIPTRY(hipMemcpyAsync(ctx->d_cs.p, cs.data(), r * 4, hipMemcpyHostToDevice, s));
    HIPTRY(hipMemcpyAsync(ctx->d_rs.p, rs.data(), r * 4, hipMemcpyHostToDevice, s));
    HIPTRY(hipMemcpyAsync(ctx->d_v.p, vtest, (size_t)n * 16, hipMemcpyHostToDevice, s));
    HIPTRY(hipMemcpyAsync(ctx->d_y0.p, y_old, (size_t)m * 16, hipMemcpyHostToDevice, s));
    BlockJob J{};
    J.m = m;
    J.n = n;
    J.r = r;
    J.sample_mode = 1;
    J.rows = ctx->d_rows.as<int32_t>();
    J.cols = ctx->d_cols.as<int32_t>();
    J.cs = ctx->d_cs.as<int32_t>();
    J.rs = ctx->d_rs.as<int32_t>();
    J.C = ctx->d_C.p;
    J.R = ctx->d_R.p;
    J.W = ctx->d_W.as<c128>();
    J.V = ctx->d_V.as<c128>();
    J.U = ctx->d_U.as<c128>();
    J.sigma = ctx->d_sigma.as<double>();
    J.v = ctx->d_v.as<c128>();
    J.y_old = ctx->d_y0.as<c128>();
    J.y_new = ctx->d_y1.as<c128>();
    std::vector<BlockJob> jobs{J};
    PackedRound pk;
    std::vector<char> bytes;
    pack_round(jobs, false, pk, bytes);
    TRY(ctx->d_plan.ensure(bytes.size()));
    HIPTRY(hipMemcpyAsync(ctx->d_plan.p, bytes.data(), bytes.size(), hipMemcpyHostToDevice, s));
    RoundPlan P = plan_from(pk, ctx->d_plan.p, 1);
    P.atol = pinv_tol;
    P.rtol = 0;
    P.tune(ctx->tune);
    if (r > kPinvLdsMaxR) P.pinv_threads[kPinvClasses - 2] = P.pinv_threads[kPinvClasses - 1] = r > 128 ? 1024 : 256;
    if ((size_t)m * r > (size_t)256 * 1024 || r > 128) P.resid_threads = 1024;
    KTimer kt;
    HIPTRY(hipEventRecord(ctx->ev0, s));
    run_round(ctx, P, kt, ctx->stream);
    HIPTRY(hipEventRecord(ctx->ev1, s));
    HIPTRY(hipGetLastError());
    JobOut res{};
    HIPTRY(hipMemcpyAsync(&res, P.out, sizeof res, hipMemcpyDeviceToHost, s));
    HIPTRY(hipMemcpyAsync(Cc, ctx->d_C.p, (size_t)m * r * 16, hipMemcpyDeviceToHost, s));
    HIPTRY(hipMemcpyAsync(Uc, ctx->d_U.p, (size_t)r * r * 16, hipMemcpyDeviceToHost, s));
    HIPTRY(hipMemcpyAsync(Rc, ctx->d_R.p, (size_t)r * n * 16, hipMemcpyDeviceToHost, s));
    HIPTRY(hipMemcpyAsync(y_new, ctx->d_y1.p, (size_t)m * 16, hipMemcpyDeviceToHost, s));
    HIPTRY(hipStreamSynchronize(s));
    float ms = 0;
    hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    vec_time[0] = ms;
    *err_out = res.err;
    return FARCUR_OK;
}
}  // namespace

extern "C" double cuda_complete_compression(
    farcur_c128* Ccompressed, farcur_c128* Ucompressed, farcur_c128* Rcompressed, int cur_order, double* vec_time_cuda,
    double pinv_tol, farcur_c128* Avtest_old, farcur_c128* Avtest_new, farcur_c128* vtest, double* col_samples,
    double* row_samples, int, int, int, int, int, int, int, int, int, int, int, int, farcur_c128* /*Ztotal*/,
    int field_arg, double k_arg, double eta_arg, double Rinteg_s_arg, double Ranal_s_arg, double Rinteg_f_arg,
    int /*cor_solid*/, int /*flag*/, double* r1_arg, double* r2_arg, double* vertex_arg, double* topol_arg,
    double* /*trian*/, double* edges_arg, double* /*un*/, double* /*ds*/, double* ln_arg, double* /*cent*/,
    int /*N_arg*/, int n_r1_arg, int n_r2_arg, int rows_vertex_arg, int cols_vertex_arg, int rows_topol_arg,
    int cols_topol_arg, int /*rows_trian*/, int /*cols_trian*/, int rows_edges_arg, int cols_edges_arg,
    int /*rows_un*/, int /*cols_un*/, int /*rows_ds*/, int /*cols_ds*/, int n_ln_arg, int /*rows_cent*/,
    int /*cols_cent*/, int, int, int, int, int, int) {
    double err = NAN;
    // the reference integrates triangle pairs closer than these radii with its singular rules
    const double rnear = std::max({0.0, Rinteg_s_arg, Ranal_s_arg, Rinteg_f_arg});
    const int rc = compat_impl(Ccompressed, Ucompressed, Rcompressed, cur_order, vec_time_cuda, pinv_tol, Avtest_old,
                               Avtest_new, vtest, col_samples, row_samples, field_arg, k_arg, eta_arg, rnear, r1_arg,
                               r2_arg,
                               vertex_arg, topol_arg, edges_arg, ln_arg, n_r1_arg, n_r2_arg, rows_vertex_arg,
                               cols_vertex_arg, rows_topol_arg, cols_topol_arg, rows_edges_arg, cols_edges_arg,
                               n_ln_arg, &err);
    if (rc != FARCUR_OK) {
        if (vec_time_cuda) vec_time_cuda[0] = -1.0;
        return NAN;
    }
    return err;
}

// ------------------------------------------------------------------ batches
int model_rank_predict(farcur_model* M, const float* d_X, const int32_t* d_cap, int n, int p, int32_t* d_r,
                       hipStream_t s);

namespace {

// Rank-model features of the block rows x cols (see farcur_batch_features).
void block_features(const farcur_ctx* ctx, const int32_t* rows, int m, const int32_t* cols, int n, float* f) {
    const double lambda = 2.0 * M_PI / ctx->k;
    auto group = [&](const int32_t* ids, int cnt, double c[3], double nrm[3], double* diam, double* hsum) {
        c[0] = c[1] = c[2] = 0;
        nrm[0] = nrm[1] = nrm[2] = 0;
        *hsum = 0;
        for (int i = 0; i < cnt; ++i) {
            const int e = ids[i];
            const int tp = ctx->h_edges[4 * e], vp = ctx->h_edges[4 * e + 2];
            double mid[3] = {0, 0, 0};
            for (int k = 0; k < 3; ++k) {
                const int v = ctx->h_topol[3 * tp + k];
                if (v == vp) continue;
                for (int x = 0; x < 3; ++x) mid[x] += 0.5 * ctx->h_vertex[3 * v + x];
            }
            for (int x = 0; x < 3; ++x) c[x] += mid[x] / cnt;
            *hsum += ctx->h_ln[e];
            for (int hh = 0; hh < 2; ++hh) {
                const int t = ctx->h_edges[4 * e + hh];
                const double* v1 = &ctx->h_vertex[3 * ctx->h_topol[3 * t]];
                const double* v2 = &ctx->h_vertex[3 * ctx->h_topol[3 * t + 1]];
                const double* v3 = &ctx->h_vertex[3 * ctx->h_topol[3 * t + 2]];
                const double a[3] = {v3[0] - v1[0], v3[1] - v1[1], v3[2] - v1[2]};
                const double b[3] = {v2[0] - v1[0], v2[1] - v1[1], v2[2] - v1[2]};
                nrm[0] += a[1] * b[2] - a[2] * b[1];
                nrm[1] += a[2] * b[0] - a[0] * b[2];
                nrm[2] += a[0] * b[1] - a[1] * b[0];
            }
        }
        double r2 = 0;
        for (int i = 0; i < cnt; ++i) {
            for (int hh = 0; hh < 2; ++hh) {
                const int t = ctx->h_edges[4 * ids[i] + hh];
                for (int k = 0; k < 3; ++k) {
                    const double* v = &ctx->h_vertex[3 * ctx->h_topol[3 * t + k]];
                    const double d2 = (v[0] - c[0]) * (v[0] - c[0]) + (v[1] - c[1]) * (v[1] - c[1]) +
                                      (v[2] - c[2]) * (v[2] - c[2]);
                    r2 = std::max(r2, d2);
                }
            }
        }
        *diam = 2.0 * std::sqrt(r2);
        const double nn = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
        for (int x = 0; x < 3; ++x) nrm[x] = nn > 0 ? nrm[x] / nn : 0.0;
    };
    double ca[3], na[3], da, ha, cb[3], nb[3], db, hb;
    group(rows, m, ca, na, &da, &ha);
    group(cols, n, cb, nb, &db, &hb);
    double d[3] = {cb[0] - ca[0], cb[1] - ca[1], cb[2] - ca[2]};
    const double dist = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (int x = 0; x < 3; ++x) d[x] = dist > 0 ? d[x] / dist : 0.0;
    f[0] = (float)(da / lambda);
    f[1] = (float)(db / lambda);
    f[2] = (float)(dist / lambda);
    f[3] = (float)(na[0] * d[0] + na[1] * d[1] + na[2] * d[2]);
    f[4] = (float)(-(nb[0] * d[0] + nb[1] * d[1] + nb[2] * d[2]));
    f[5] = (float)(ctx->k * (ha + hb) / (m + n));
    f[6] = (float)std::log((double)m);
    f[7] = (float)std::log((double)n);
}

struct BState {
    int m = 0, n = 0, cap = 0, r = 0, round = 0;
    int status = 0;  // 0 active, 1 converged, 2 stopped at the cap, 3 NaN residual
    int ycur = 0;
    int kept = 0, rank_eps = 0, sweeps = 0;
    int h = 0;       // held-out rows of the current round (learned mode)
    double err = INFINITY, smax = 0, heldout_err = -1;
    RoundBufs fin{};
    int fin_r = 0;
};

// Distinct triangles touched by a set of RWGs.
int distinct_triangles(const std::vector<int32_t>& tri_of_rwg, const int32_t* ids, int n,
                       std::unordered_set<int32_t>& scratch) {
    scratch.clear();
    for (int i = 0; i < n; ++i) {
        scratch.insert(tri_of_rwg[2 * (size_t)ids[i]]);
        scratch.insert(tri_of_rwg[2 * (size_t)ids[i] + 1]);
    }
    return (int)scratch.size();
}

}  // namespace

// A captured round: the plan upload (from pinned host memory), the round's
// kernels and the JobOut download, replayed as one hipGraph launch. The entry
// owns the device plan the kernels read and the pinned staging buffers the
// memcpy nodes name, so nothing it references is freed or moved while cached.
// The kernels read the plan from device memory at run time; what the graph
// fixes are the values passed by value (counts, section offsets, thresholds,
// launch shapes), which `key` holds byte for byte. A round whose key matches
// replays the graph after writing its own plan into `h_plan`.
struct RoundGraph {
    std::vector<char> key;
    hipGraphExec_t exec = nullptr;
    DBuf d_plan;
    char* h_plan = nullptr;
    JobOut* h_out = nullptr;
    size_t plan_bytes = 0;
    int njobs = 0;
    uint64_t last_use = 0;
    void release() {
        if (exec) hipGraphExecDestroy(exec);
        exec = nullptr;
        d_plan.release();
        if (h_plan) hipHostFree(h_plan);
        if (h_out) hipHostFree(h_out);
        h_plan = nullptr;
        h_out = nullptr;
    }
};

struct farcur_batch {
    farcur_ctx* ctx = nullptr;
    int nb = 0;
    std::vector<int32_t> row_off, col_off, row_ids, col_ids;
    DBuf d_rows, d_cols, d_rperm, d_cperm, d_v, d_ya, d_yb, d_vo, d_plan;
    Arena arena;
    std::vector<BState> st;
    KTimer kt;
    bool profile = false;
    int f32 = 0;
    bool has_result = false;
    // Jobs of every round of the last compress (host copies), for accounting.
    std::vector<std::vector<BlockJob>> rounds;
    double work[KTimer::kClasses] = {};
    // Rank-model features (nb x 8) and the model used by the learned mode.
    std::vector<float> feat;
    DBuf d_feat, d_cap, d_rpred;
    farcur_model* model = nullptr;
    bool learned = false;
    int first_round_misses = 0;
    // Host-side wall time of the last compress (ms): rank prediction round
    // trip, round planning, waiting for the device, advancing the block states.
    double host_ms[4] = {};
    // Two-stream pipeline of the batched rounds: number of chunks (1 = off).
    int pipeline_chunks = 1;
    // pinv r classes split over the context's two streams within a round
    bool pinv_side_stream = true;
    std::vector<char> h_plan;
    std::vector<char> key_buf;  // graph key of the round being planned (kept for its capacity)
    std::vector<PackScratch> pack_sc;  // per chunk, reused across rounds
    std::vector<JobDims> dims_buf;      // the round's job sizes, reused across rounds
    std::vector<RoundBufs> bufs_buf;    // the round's buffers and results, reused across rounds
    std::vector<JobOut> outs_buf;
    // Rounds captured as hipGraphs, keyed by plan content and launch
    // parameters (at most kMaxGraphs, least recently used evicted).
    static constexpr int kMaxGraphs = 8;
    bool use_graphs = true;
    std::vector<RoundGraph> graphs;
    uint64_t graph_clock = 0;
    int graph_captures = 0, graph_replays = 0;
    // Interpolation matrices P_b = C_b U_b for the product (arena memory).
    std::vector<void*> interp;
    bool interp_ready = false;
    // SVD recompression of the last result: outputs (C', R', sigma per block at
    // rc_off, k in rc_k), workspace and job records
    DBuf d_rc, d_rc_ws, d_rc_jobs, d_rc_k;
    std::vector<int32_t> rc_k;
    std::vector<size_t> rc_off;
    bool rc_ready = false;
    // ACA of every block (k_aca): per-block regions of d_aca at aca_off
    DBuf d_aca, d_aca_jobs, d_aca_ks;  // d_aca_ks: k then status, nblocks each
    // learned mode: blocks whose final sample took every row or column (stored exactly)
    std::vector<ExactJob> exact_jobs;
    DBuf d_exact;
    std::vector<size_t> aca_off;
    std::vector<int32_t> aca_kmax, aca_k, aca_status;
    bool aca_ready = false;
};

namespace {

// A round's buffers of one block (and its h held-out rows), carved from one
// arena allocation: sample ids and row positions, C, R, W (then the pinv's
// scratch), V when r exceeds the LDS pinv, U, sigma, held-out ids, positions
// and rows; each piece 256-byte aligned.
int batch_alloc_round(farcur_batch* B, int r, int m, int n, int h, RoundBufs& rb) {
    const size_t ps = psize(B->f32);
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t sz_ids = al((size_t)r * 4), sz_C = al((size_t)m * r * ps), sz_R = al((size_t)r * n * ps);
    const size_t sz_rr = al((size_t)r * r * 16), sz_V = r > kPinvLdsMaxR ? al(pinv_v_elems(r) * 16) : 0;
    const size_t sz_sig = al((size_t)r * 8), sz_h = h > 0 ? al((size_t)h * 4) : 0;
    const size_t sz_H = h > 0 ? al((size_t)h * n * ps) : 0;
    void* p;
    TRY(B->arena.alloc(3 * sz_ids + sz_C + sz_R + 2 * sz_rr + sz_V + sz_sig + 2 * sz_h + sz_H, &p));
    char* q = (char*)p;
    auto take = [&](size_t bytes) {
        char* at = q;
        q += bytes;
        return at;
    };
    rb.cs = (int32_t*)take(sz_ids);
    rb.rs = (int32_t*)take(sz_ids);
    rb.rpos = (int32_t*)take(sz_ids);
    rb.C = take(sz_C);
    rb.R = take(sz_R);
    rb.W = (c128*)take(sz_rr);
    rb.V = sz_V ? (c128*)take(sz_V) : nullptr;
    rb.U = (c128*)take(sz_rr);
    rb.sigma = (double*)take(sz_sig);
    rb.hs = h > 0 ? (int32_t*)take(sz_h) : nullptr;
    rb.hpos = h > 0 ? (int32_t*)take(sz_h) : nullptr;
    rb.H = h > 0 ? (void*)take(sz_H) : nullptr;
    return FARCUR_OK;
}

BlockJob make_job(farcur_batch* B, int b, const RoundBufs& rb, const farcur_opts& o) {
    const BState& s = B->st[b];
    BlockJob J{};
    J.b = b;
    J.m = s.m;
    J.n = s.n;
    J.r = s.r;
    J.round = s.round;
    J.sample_mode = B->learned ? 3 : (o.row_mode == FARCUR_ROWS_PIVOT ? 2 : 0);
    J.h = B->learned ? s.h : 0;
    J.hs = rb.hs;
    J.hpos = rb.hpos;
    J.H = rb.H;
    J.rows = B->d_rows.as<int32_t>() + B->row_off[b];
    J.cols = B->d_cols.as<int32_t>() + B->col_off[b];
    J.cs = rb.cs;
    J.rs = rb.rs;
    J.rpos = rb.rpos;
    J.C = rb.C;
    J.R = rb.R;
    J.W = rb.W;
    J.V = rb.V;
    J.U = rb.U;
    J.sigma = rb.sigma;
    J.v = B->d_v.as<c128>() + B->col_off[b];
    J.row_perm = B->d_rperm.as<int32_t>() + B->row_off[b];
    J.col_perm = B->d_cperm.as<int32_t>() + B->col_off[b];
    c128* ya = B->d_ya.as<c128>() + B->row_off[b];
    c128* yb = B->d_yb.as<c128>() + B->row_off[b];
    J.y_old = s.ycur ? yb : ya;
    J.y_new = s.ycur ? ya : yb;
    return J;
}

// Held-out rows of a learned round: at least `heldout`, one per 32 block rows
// on tall blocks (a fixed count under-samples the error of blocks with
// hundreds of rows), at most 4 x heldout, and no more than the rows left
// over by the r samples.
int heldout_rows(int m, int r, int heldout) {
    const int h = std::min(std::max(heldout, (m + 31) / 32), 4 * heldout);
    return std::max(0, std::min(h, m - r));
}

// Learned mode: the predicted round is accepted when its held-out rows meet
// tol; otherwise the sample count doubles (capped) and the block runs again.
void advance_learned(BState& s, const JobOut& out, const RoundBufs& rb, double accept, double tol, int heldout,
                     double growth, int oversample) {
    s.fin = rb;
    s.fin_r = s.r;
    s.kept = out.kept;
    s.rank_eps = out.rank_eps;
    s.sweeps = out.sweeps;
    s.smax = out.sigma_max;
    // acceptance on the ratio estimate of the whole block's error; the plain
    // held-out rows' error is reported beside it
    s.err = s.h > 0 ? out.err_est : 0.0;
    s.heldout_err = s.h > 0 ? out.err : -1.0;
    s.round += 1;
    if (s.r >= std::min(s.m, s.n)) {
        // the sample took every row or every column: the block is stored
        // exactly once the compress ends (ExactJob), nothing to check
        s.status = 1;
        s.err = 0.0;
        return;
    }
    if (s.err <= accept) {
        s.status = 1;
        return;
    }
    if (!(s.err == s.err)) {
        s.status = 3;
        return;
    }
    if (s.r >= s.cap) {
        // No larger sample exists: the tolerance itself (without the safety
        // margin of the acceptance test) decides.
        s.status = s.err <= tol ? 1 : 2;
        return;
    }
    // grow by the factor, at least by the oversampling
    s.r = std::min(std::max((int)std::ceil(growth * s.r), s.r + std::max(oversample, 1)), s.cap);
    s.h = heldout_rows(s.m, s.r, heldout);
}

// Applies one round's result to a block's state (cuda_cur_compress3 control flow).
void advance(BState& s, const JobOut& out, const RoundBufs& rb, double tol) {
    s.ycur ^= 1;
    s.fin = rb;
    s.fin_r = s.r;
    s.kept = out.kept;
    s.rank_eps = out.rank_eps;
    s.sweeps = out.sweeps;
    s.smax = out.sigma_max;
    if (s.round == 0) {
        int rn;
        next_r(s.r, s.cap, 0, &rn);
        s.r = rn;
        s.round = 1;
        return;
    }
    s.err = out.err;
    s.round += 1;
    if (!(s.err > tol)) {
        s.status = s.err <= tol ? 1 : 3;  // 3: stopped on a NaN residual
        return;
    }
    int rn;
    if (!next_r(s.r, s.cap, s.round, &rn)) {
        s.status = 2;
        return;
    }
    s.r = rn;
}

// Cost estimate of a job for balancing pipeline chunks: fill entries (36 Green
// evaluations each) plus the pinv (~ r^3 x sweeps, in the same units).
double job_cost(int m, int n, int r, int h) {
    const double entries = (double)(m + n + r) * r + (double)h * n;
    return 36.0 * entries + 60.0 * (double)r * r * r;
}

// One round over the active blocks. With pipelining, the jobs are cut into
// chunks of equal estimated cost that alternate between the context's two
// streams, so that one chunk's fill (VALU-bound) runs beside another's pinv
// (latency-bound); each chunk is its own launch sequence with its own plan.
// Profiling keeps one chunk on one stream, so that every kernel class is timed
// alone.
int batch_round(farcur_batch* B, const std::vector<int>& active, const farcur_opts& o, std::vector<RoundBufs>& bufs,
                std::vector<JobOut>& outs) {
    farcur_ctx* ctx = B->ctx;
    const auto t_plan0 = std::chrono::steady_clock::now();
    // Sizes and buffers first; the job records (192 bytes each) are built
    // once, straight into the plan staging the round uploads, after the
    // layout and the graph are known. The size list keeps its capacity
    // across rounds and calls.
    const int nj = (int)active.size();
    bufs.resize(nj);
    std::vector<JobDims>& dims = B->dims_buf;
    dims.resize(nj);
    for (int i = 0; i < nj; ++i) {
        const BState& st = B->st[active[i]];
        const int h = B->learned ? st.h : 0;
        TRY(batch_alloc_round(B, st.r, st.m, st.n, h, bufs[i]));
        dims[i] = JobDims{st.m, st.n, st.r, h, 1, B->learned ? 1 : 0};  // batch jobs always carry row positions
    }
    int nchunks = 1;
    if (!B->profile && B->pipeline_chunks > 1 && nj >= 64 * B->pipeline_chunks) nchunks = B->pipeline_chunks;
    // chunk boundaries at equal cost
    std::vector<int> cut{0};
    if (nchunks > 1) {
        auto cost = [&](int j) { return job_cost(dims[j].m, dims[j].n, dims[j].r, dims[j].h); };
        double total_cost = 0;
        for (int j = 0; j < nj; ++j) total_cost += cost(j);
        double acc = 0;
        int c = 1;
        for (int j = 0; j < nj && c < nchunks; ++j) {
            acc += cost(j);
            if (acc >= total_cost * c / nchunks) {
                cut.push_back(j + 1);
                ++c;
            }
        }
    }
    if (cut.back() != nj) cut.push_back(nj);
    nchunks = (int)cut.size() - 1;
    std::vector<PackedRound> pk(nchunks);
    std::vector<size_t> off(nchunks + 1, 0);
    if ((int)B->pack_sc.size() < nchunks) B->pack_sc.resize(nchunks);
    for (int c = 0; c < nchunks; ++c) {
        std::memset((void*)&pk[c], 0, sizeof(PackedRound));  // padding bytes go into the graph key
        const JobDims* dc = dims.data() + cut[c];
        pack_layout(cut[c + 1] - cut[c], o.row_mode == FARCUR_ROWS_PIVOT, pk[c], B->pack_sc[c],
                    [dc](int j) { return dc[j]; });
        off[c + 1] = off[c] + pk[c].size;  // sizes are multiples of 256
    }
    const size_t plan_bytes = off[nchunks];
    auto write_plan = [&](char* dst) {
        for (int c = 0; c < nchunks; ++c) {
            BlockJob* rec = (BlockJob*)(dst + off[c] + pk[c].off_jobs);
            for (int j = cut[c]; j < cut[c + 1]; ++j) rec[j - cut[c]] = make_job(B, active[j], bufs[j], o);
            pack_write_tail(cut[c + 1] - cut[c], pk[c], B->pack_sc[c], dst + off[c]);
        }
    };
    hipStream_t s0 = ctx->stream, s1 = ctx->stream2;
    // Launch parameters that the kernels take by value.
    struct KeyParams {
        int f32, pivot, nchunks, nj;
        double atol, rtol, label_eps, k, eta;
        uint64_t seed, geom;
        Tuning tune;
        const void* geom_ptr[5];  // device geometry the fill takes by value
        int side, qr_core;
    } kp;
    std::memset((void*)&kp, 0, sizeof kp);
    kp.f32 = B->f32;
    kp.pivot = o.row_mode == FARCUR_ROWS_PIVOT;
    kp.nchunks = nchunks;
    kp.nj = nj;
    kp.atol = o.pinv_tol;
    kp.rtol = o.pinv_rtol;
    kp.label_eps = o.label_eps;
    kp.k = ctx->k;
    kp.eta = ctx->eta;
    kp.seed = o.seed;
    kp.geom = ctx->geom_hash;
    kp.tune = ctx->tune;
    {
        const FillParams fp = ctx->fill_params(B->f32);
        const void* gp[5] = {fp.rec, fp.cen, fp.tri.tri, fp.tri.tcen, fp.tri.lrec};
        std::memcpy(kp.geom_ptr, gp, sizeof gp);
    }
    kp.side = B->pinv_side_stream ? 1 : 0;
    kp.qr_core = o.learned_core == FARCUR_CORE_QR ? 1 : 0;
    // Enqueues the round: plan upload from `hplan` to `dplan`, kernels, JobOut
    // download to `hout`.
    std::vector<RoundPlan> plans(nchunks);
    auto enqueue = [&](const void* hplan, void* dplan, JobOut* hout) -> int {
        HIPTRY(hipMemcpyAsync(dplan, hplan, plan_bytes, hipMemcpyHostToDevice, s0));
        if (nchunks > 1) {
            HIPTRY(hipEventRecord(ctx->ev_fork, s0));
            HIPTRY(hipStreamWaitEvent(s1, ctx->ev_fork, 0));
        }
        for (int c = 0; c < nchunks; ++c) {
            RoundPlan& P = plans[c];
            P = plan_from(pk[c], (char*)dplan + off[c], cut[c + 1] - cut[c]);
            P.f32 = B->f32;
            P.pivot = o.row_mode == FARCUR_ROWS_PIVOT;
            P.atol = o.pinv_tol;
            P.rtol = o.pinv_rtol;
            P.seed = o.seed;
            P.label_eps = o.label_eps;
            P.qr_core = kp.qr_core;
            P.tune(ctx->tune);
            // one chunk: the pinv classes use the second stream (not while
            // profiling, which times every class alone)
            hipStream_t side = nchunks == 1 && !B->profile && B->pinv_side_stream ? s1 : nullptr;
            run_round(ctx, P, B->kt, (c & 1) ? s1 : s0, side);
        }
        HIPTRY(hipGetLastError());
        if (nchunks > 1) {
            HIPTRY(hipEventRecord(ctx->ev_join, s1));
            HIPTRY(hipStreamWaitEvent(s0, ctx->ev_join, 0));
        }
        for (int c = 0; c < nchunks; ++c)
            HIPTRY(hipMemcpyAsync(hout + cut[c], plans[c].out, (size_t)(cut[c + 1] - cut[c]) * sizeof(JobOut),
                                  hipMemcpyDeviceToHost, s0));
        return FARCUR_OK;
    };
    outs.resize(nj);
    const auto t_plan1 = std::chrono::steady_clock::now();
    B->host_ms[1] += std::chrono::duration<double, std::milli>(t_plan1 - t_plan0).count();
    if (!B->use_graphs || B->profile) {
        if (B->h_plan.size() < plan_bytes) B->h_plan.resize(plan_bytes);
        write_plan(B->h_plan.data());
        TRY(B->d_plan.ensure(plan_bytes));
        TRY(enqueue(B->h_plan.data(), B->d_plan.p, outs.data()));
        HIPTRY(hipStreamSynchronize(s0));
    } else {
        // Key: every value the captured launches take by value (the launch
        // parameters, the chunk cuts and each chunk's section layout, counts
        // and class sizes); the plan contents are read from device memory.
        std::vector<char>& key = B->key_buf;
        key.clear();
        auto put_key = [&](const void* p, size_t n) { key.insert(key.end(), (const char*)p, (const char*)p + n); };
        put_key(&kp, sizeof kp);
        put_key(cut.data(), cut.size() * sizeof(int));
        put_key(pk.data(), pk.size() * sizeof(PackedRound));
        RoundGraph* g = nullptr;
        for (RoundGraph& e : B->graphs)
            if (e.exec && e.plan_bytes == plan_bytes && e.njobs == nj && e.key == key) g = &e;
        if (g) {
            // the previous launch of this graph has completed (every round
            // ends with a stream synchronisation), so its staging is free
            write_plan(g->h_plan);
            ++B->graph_replays;
        } else {
            if ((int)B->graphs.size() < farcur_batch::kMaxGraphs) {
                B->graphs.emplace_back();
                g = &B->graphs.back();
            } else {
                g = &B->graphs[0];
                for (RoundGraph& e : B->graphs)
                    if (e.last_use < g->last_use) g = &e;
                HIPTRY(hipStreamSynchronize(s0));
                g->release();
            }
            g->key = key;
            g->plan_bytes = plan_bytes;
            g->njobs = nj;
            TRY(g->d_plan.ensure(plan_bytes));
            HIPTRY(hipHostMalloc((void**)&g->h_plan, plan_bytes, hipHostMallocDefault));
            HIPTRY(hipHostMalloc((void**)&g->h_out, std::max<size_t>(nj, 1) * sizeof(JobOut), hipHostMallocDefault));
            write_plan(g->h_plan);
            hipGraph_t graph = nullptr;
            HIPTRY(hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal));
            const int rc = enqueue(g->h_plan, g->d_plan.p, g->h_out);
            const hipError_t ec = hipStreamEndCapture(s0, &graph);
            if (rc != FARCUR_OK) {
                if (graph) hipGraphDestroy(graph);
                g->release();
                return rc;
            }
            if (ec != hipSuccess) {
                g->release();
                return fail(FARCUR_EHIP, "graph capture of a round failed: %s", hipGetErrorString(ec));
            }
            const hipError_t ei = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
            hipGraphDestroy(graph);
            if (ei != hipSuccess) {
                g->exec = nullptr;
                g->release();
                return fail(FARCUR_EHIP, "graph instantiation failed: %s", hipGetErrorString(ei));
            }
            ++B->graph_captures;
        }
        g->last_use = ++B->graph_clock;
        HIPTRY(hipGraphLaunch(g->exec, s0));
        HIPTRY(hipStreamSynchronize(s0));
        std::memcpy(outs.data(), g->h_out, (size_t)nj * sizeof(JobOut));
    }
    B->host_ms[2] += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_plan1).count();
    B->kt.collect();
    if (B->profile) {
        std::vector<BlockJob> jobs(nj);
        for (int j = 0; j < nj; ++j) jobs[j] = make_job(B, active[j], bufs[j], o);
        B->rounds.push_back(std::move(jobs));
    }
    return FARCUR_OK;
}

// Algorithmic work of the recorded rounds: Green evaluations of the fill
// (9 x distinct test/source triangle pairs of each panel), 96 r^3 flops per
// core pinv, bytes of the residual's panel reads.
int batch_account(farcur_batch* B) {
    farcur_ctx* ctx = B->ctx;
    std::unordered_set<int32_t> scratch;
    std::vector<int> trows(B->nb), tcols(B->nb);
    for (int b = 0; b < B->nb; ++b) {
        trows[b] = distinct_triangles(ctx->tri_of_rwg, &B->row_ids[B->row_off[b]], B->row_off[b + 1] - B->row_off[b],
                                      scratch);
        tcols[b] = distinct_triangles(ctx->tri_of_rwg, &B->col_ids[B->col_off[b]], B->col_off[b + 1] - B->col_off[b],
                                      scratch);
    }
    for (double& w : B->work) w = 0;
    const double ps = (double)psize(B->f32);
    for (auto& jobs : B->rounds) {
        for (const BlockJob& J : jobs) {
            std::vector<int32_t> cs(J.r), rs(J.r);
            HIPTRY(hipMemcpy(cs.data(), J.cs, J.r * 4, hipMemcpyDeviceToHost));
            HIPTRY(hipMemcpy(rs.data(), J.rs, J.r * 4, hipMemcpyDeviceToHost));
            const double tcs = distinct_triangles(ctx->tri_of_rwg, cs.data(), J.r, scratch);
            const double trs = distinct_triangles(ctx->tri_of_rwg, rs.data(), J.r, scratch);
            double ths = 0;
            if (J.h > 0) {
                std::vector<int32_t> hs(J.h);
                HIPTRY(hipMemcpy(hs.data(), J.hs, J.h * 4, hipMemcpyDeviceToHost));
                ths = distinct_triangles(ctx->tri_of_rwg, hs.data(), J.h, scratch);
            }
            // panels C, R, the held-out rows H and, unless gathered from C, W
            B->work[1] += 9.0 * (trows[J.b] * tcs + (J.rpos ? 0.0 : trs * tcs) + trs * tcols[J.b] + ths * tcols[J.b]);
            B->work[2] += 96.0 * (double)J.r * J.r * J.r;
            B->work[4 + pinv_class(J.r)] += 96.0 * (double)J.r * J.r * J.r;
            B->work[3] += ps * ((double)J.m * J.r + (double)J.r * J.n) + 16.0 * J.r * J.r + 16.0 * (J.n + 2.0 * J.m);
        }
    }
    return FARCUR_OK;
}

}  // namespace

extern "C" int farcur_batch_create(farcur_ctx* ctx, int nblocks, const int* row_off, const int* row_ids,
                                   const int* col_off, const int* col_ids, farcur_batch** out) {
    if (!ctx || !row_off || !col_off || !out) return fail(FARCUR_EINVAL, "NULL argument");
    *out = nullptr;
    if (nblocks <= 0) return fail(FARCUR_EINVAL, "nblocks must be > 0");
    if (ctx->ne == 0) return fail(FARCUR_ESTATE, "no geometry");
    if (row_off[0] != 0 || col_off[0] != 0) return fail(FARCUR_EINVAL, "offsets must start at 0");
    for (int b = 0; b < nblocks; ++b)
        if (row_off[b + 1] <= row_off[b] || col_off[b + 1] <= col_off[b])
            return fail(FARCUR_EINVAL, "block %d is empty or offsets decrease", b);
    const int nr = row_off[nblocks], nc = col_off[nblocks];
    if (!row_ids || !col_ids) return fail(FARCUR_EINVAL, "NULL ids");
    TRY(check_ids(row_ids, nr, ctx->ne, "row_ids"));
    TRY(check_ids(col_ids, nc, ctx->ne, "col_ids"));
    {
        FarCheck fc;
        for (int b = 0; b < nblocks; ++b)
            TRY(check_far_block(ctx, row_ids + row_off[b], row_off[b + 1] - row_off[b], col_ids + col_off[b],
                                col_off[b + 1] - col_off[b], 0.0, fc, "block", b));
    }
    TRY(ctx->select());
    auto B = std::make_unique<farcur_batch>();
    B->ctx = ctx;
    B->nb = nblocks;
    B->row_off.assign(row_off, row_off + nblocks + 1);
    B->col_off.assign(col_off, col_off + nblocks + 1);
    B->row_ids.assign(row_ids, row_ids + nr);
    B->col_ids.assign(col_ids, col_ids + nc);
    TRY(B->d_rows.ensure((size_t)nr * 4));
    TRY(B->d_cols.ensure((size_t)nc * 4));
    {
        std::vector<int32_t> rp(nr), cp(nc);
        for (int b = 0; b < nblocks; ++b) {
            tile_order(row_ids + row_off[b], row_off[b + 1] - row_off[b], ctx->h_cen.data(), rp.data() + row_off[b]);
            tile_order(col_ids + col_off[b], col_off[b + 1] - col_off[b], ctx->h_cen.data(), cp.data() + col_off[b]);
        }
        TRY(B->d_rperm.ensure((size_t)nr * 4));
        TRY(B->d_cperm.ensure((size_t)nc * 4));
        HIPTRY(hipMemcpyAsync(B->d_rperm.p, rp.data(), (size_t)nr * 4, hipMemcpyHostToDevice, ctx->stream));
        HIPTRY(hipMemcpyAsync(B->d_cperm.p, cp.data(), (size_t)nc * 4, hipMemcpyHostToDevice, ctx->stream));
        HIPTRY(hipStreamSynchronize(ctx->stream));
    }
    TRY(B->d_v.ensure((size_t)nc * 16));
    TRY(B->d_vo.ensure((size_t)(nblocks + 1) * 4));
    TRY(B->d_ya.ensure((size_t)nr * 16));
    TRY(B->d_yb.ensure((size_t)nr * 16));
    HIPTRY(hipMemcpyAsync(B->d_rows.p, row_ids, (size_t)nr * 4, hipMemcpyHostToDevice, ctx->stream));
    HIPTRY(hipMemcpyAsync(B->d_cols.p, col_ids, (size_t)nc * 4, hipMemcpyHostToDevice, ctx->stream));
    HIPTRY(hipMemcpyAsync(B->d_vo.p, col_off, (size_t)(nblocks + 1) * 4, hipMemcpyHostToDevice, ctx->stream));
    HIPTRY(hipStreamSynchronize(ctx->stream));
    B->st.resize(nblocks);
    B->feat.resize((size_t)nblocks * kMlpFeatures);
    for (int b = 0; b < nblocks; ++b)
        block_features(ctx, row_ids + row_off[b], row_off[b + 1] - row_off[b], col_ids + col_off[b],
                       col_off[b + 1] - col_off[b], &B->feat[(size_t)b * kMlpFeatures]);
    TRY(B->d_feat.ensure(B->feat.size() * 4));
    TRY(B->d_cap.ensure((size_t)nblocks * 4));
    TRY(B->d_rpred.ensure((size_t)nblocks * 4));
    HIPTRY(hipMemcpyAsync(B->d_feat.p, B->feat.data(), B->feat.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    HIPTRY(hipStreamSynchronize(ctx->stream));
    *out = B.release();
    return FARCUR_OK;
}

extern "C" void farcur_batch_destroy(farcur_batch* B) {
    if (!B) return;
    hipSetDevice(B->ctx->device);
    hipStreamSynchronize(B->ctx->stream);
    DBuf* bufs[] = {&B->d_exact, &B->d_rows, &B->d_cols, &B->d_rperm, &B->d_cperm, &B->d_v, &B->d_ya, &B->d_yb,
                    &B->d_vo,   &B->d_plan, &B->d_feat, &B->d_cap, &B->d_rpred, &B->d_rc,  &B->d_rc_ws,
                    &B->d_rc_jobs, &B->d_rc_k, &B->d_aca, &B->d_aca_jobs, &B->d_aca_ks};
    for (DBuf* b : bufs) b->release();
    for (RoundGraph& g : B->graphs) g.release();
    B->arena.release();
    delete B;
}

// Byte offsets of one block's recompression outputs within its region:
// C' (m x r), R' (r x n) in panel precision, sigma (r doubles).
struct RcLayout {
    size_t c, r, sigma, size;
    RcLayout(int m, int n, int rr, size_t ps) {
        auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
        c = 0;
        r = al((size_t)m * rr * ps);
        sigma = r + al((size_t)rr * n * ps);
        size = sigma + al((size_t)rr * 8);
    }
};

// SVD recompression of every block of the last result (svd_compress.m:19-34
// through postcompress_wrapper.m:1-5): C' (m x k), sigma (k), R' (k x n) with
// C U R = C' diag(sigma) R' up to the dropped singular values.
extern "C" int farcur_batch_recompress(farcur_batch* B, double thresh, int* ranks, double* ms) {
    if (!B) return fail(FARCUR_EINVAL, "NULL batch");
    if (!B->has_result) return fail(FARCUR_ESTATE, "no batch result");
    if (!(thresh >= 0) || !std::isfinite(thresh)) return fail(FARCUR_EINVAL, "thresh must be finite and >= 0");
    farcur_ctx* ctx = B->ctx;
    TRY(ctx->select());
    const size_t ps = psize(B->f32);
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    // outputs: per block C' (m x r), R' (r x n) in panel precision, sigma (r), k
    B->rc_off.assign(B->nb + 1, 0);
    for (int b = 0; b < B->nb; ++b) {
        const BState& s = B->st[b];
        if (s.fin_r > kRecompressMaxR)
            return fail(FARCUR_EINVAL, "block %d: r = %d exceeds the recompression bound %d", b, s.fin_r,
                        kRecompressMaxR);
        B->rc_off[b + 1] = B->rc_off[b] + RcLayout(s.m, s.n, s.fin_r, ps).size;
    }
    TRY(B->d_rc.ensure(B->rc_off[B->nb]));
    TRY(B->d_rc_k.ensure((size_t)B->nb * 4));
    // workspace reused by chunks of blocks of at most ~1 GiB together
    const size_t ws_cap = (size_t)1 << 26;  // c128 elements
    size_t ws_max = 0;
    std::vector<RecompressJob> jobs(B->nb);
    std::vector<size_t> wsoff(B->nb);
    std::vector<int> cut{0};
    size_t ws_acc = 0;
    for (int b = 0; b < B->nb; ++b) {
        const BState& s = B->st[b];
        const size_t w = al(recompress_ws_elems(s.m, s.n, s.fin_r) * 16) / 16;
        if (ws_acc > 0 && ws_acc + w > ws_cap) {
            cut.push_back(b);
            ws_acc = 0;
        }
        char* o = (char*)B->d_rc.p + B->rc_off[b];
        RecompressJob& J = jobs[b];
        J.m = s.m;
        J.n = s.n;
        J.r = s.fin_r;
        J.pad = 0;
        J.C = s.fin.C;
        J.U = s.fin.U;
        J.R = s.fin.R;
        wsoff[b] = ws_acc;  // workspace offset, based below
        const RcLayout L(s.m, s.n, s.fin_r, ps);
        J.Cout = o + L.c;
        J.Rout = o + L.r;
        J.sigma = (double*)(o + L.sigma);
        J.k_out = B->d_rc_k.as<int32_t>() + b;
        ws_acc += w;
        ws_max = std::max(ws_max, ws_acc);
    }
    cut.push_back(B->nb);
    TRY(B->d_rc_ws.ensure(std::max<size_t>(ws_max, 1) * 16));
    for (int b = 0; b < B->nb; ++b) jobs[b].ws = B->d_rc_ws.as<c128>() + wsoff[b];
    TRY(B->d_rc_jobs.ensure(jobs.size() * sizeof(RecompressJob)));
    hipStream_t st = ctx->stream;
    HIPTRY(hipMemcpyAsync(B->d_rc_jobs.p, jobs.data(), jobs.size() * sizeof(RecompressJob), hipMemcpyHostToDevice,
                          st));
    HIPTRY(hipEventRecord(ctx->ev0, st));
    for (size_t c = 0; c + 1 < cut.size(); ++c)
        launch_recompress(B->d_rc_jobs.as<RecompressJob>() + cut[c], cut[c + 1] - cut[c], thresh, B->f32, st);
    HIPTRY(hipGetLastError());
    HIPTRY(hipEventRecord(ctx->ev1, st));
    B->rc_k.assign(B->nb, 0);
    HIPTRY(hipMemcpyAsync(B->rc_k.data(), B->d_rc_k.p, (size_t)B->nb * 4, hipMemcpyDeviceToHost, st));
    HIPTRY(hipStreamSynchronize(st));
    float t = 0;
    HIPTRY(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
    if (ms) *ms = t;
    if (ranks) std::copy(B->rc_k.begin(), B->rc_k.end(), ranks);
    B->rc_ready = true;
    return FARCUR_OK;
}

extern "C" int farcur_batch_recompressed_block(farcur_batch* B, int b, farcur_c128* C, double* sigma, farcur_c128* R,
                                               int* k) {
    if (!B) return fail(FARCUR_EINVAL, "NULL batch");
    if (!B->rc_ready) return fail(FARCUR_ESTATE, "no recompression of the current result");
    if (b < 0 || b >= B->nb) return fail(FARCUR_EINVAL, "block %d out of range", b);
    farcur_ctx* ctx = B->ctx;
    TRY(ctx->select());
    const BState& s = B->st[b];
    const int kk = B->rc_k[b];
    if (k) *k = kk;
    const RcLayout L(s.m, s.n, s.fin_r, psize(B->f32));
    const char* o = (const char*)B->d_rc.p + B->rc_off[b];
    hipStream_t st = ctx->stream;
    if (kk > 0) {
        TRY(copy_panel(C, o + L.c, (size_t)s.m * kk, B->f32, st));
        TRY(copy_panel(R, o + L.r, (size_t)kk * s.n, B->f32, st));
        if (sigma) HIPTRY(hipMemcpyAsync(sigma, o + L.sigma, (size_t)kk * 8, hipMemcpyDeviceToHost, st));
    }
    HIPTRY(hipStreamSynchronize(st));
    return FARCUR_OK;
}

// Byte offsets of one block's ACA region: U (m x kmax), Vt (n x kmax), the
// residual row (n), used flags (m + n), pivots I, J (kmax).
struct AcaLayout {
    size_t U, Vt, row, used, I, J, size;
    AcaLayout(int m, int n, int kmax) {
        auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
        U = 0;
        Vt = U + al((size_t)m * kmax * 16);
        row = Vt + al((size_t)n * kmax * 16);
        used = row + al((size_t)n * 16);
        I = used + al((size_t)(m + n) * 4);
        J = I + al((size_t)kmax * 4);
        size = J + al((size_t)kmax * 4);
    }
};

// Partially pivoted ACA of every block of the batch (ACA.m:40-128) with the
// FP64 element: U (m x k), V (k x n) with Z ~ U V.
extern "C" int farcur_batch_aca(farcur_batch* B, double tol, int max_rank, int* ranks, double* ms) {
    if (!B) return fail(FARCUR_EINVAL, "NULL batch");
    if (!(tol >= 0) || !std::isfinite(tol)) return fail(FARCUR_EINVAL, "tol must be finite and >= 0");
    if (max_rank < 0) return fail(FARCUR_EINVAL, "max_rank must be >= 0");
    farcur_ctx* ctx = B->ctx;
    if (ctx->ne == 0) return fail(FARCUR_ESTATE, "no geometry");
    TRY(ctx->select());
    B->aca_ready = false;
    B->aca_off.assign(B->nb + 1, 0);
    B->aca_kmax.assign(B->nb, 0);
    for (int b = 0; b < B->nb; ++b) {
        const int m = B->row_off[b + 1] - B->row_off[b], n = B->col_off[b + 1] - B->col_off[b];
        const int D = std::min(m, n);
        B->aca_kmax[b] = std::max(1, std::min(D, max_rank > 0 ? max_rank : 128));
        B->aca_off[b + 1] = B->aca_off[b] + AcaLayout(m, n, B->aca_kmax[b]).size;
    }
    TRY(B->d_aca.ensure(B->aca_off[B->nb]));
    TRY(B->d_aca_ks.ensure((size_t)B->nb * 8));
    std::vector<AcaJob> jobs(B->nb);
    for (int b = 0; b < B->nb; ++b) {
        const int m = B->row_off[b + 1] - B->row_off[b], n = B->col_off[b + 1] - B->col_off[b];
        const AcaLayout L(m, n, B->aca_kmax[b]);
        char* o = (char*)B->d_aca.p + B->aca_off[b];
        AcaJob& J = jobs[b];
        J.m = m;
        J.n = n;
        J.kmax = B->aca_kmax[b];
        J.pad = 0;
        J.rows = B->d_rows.as<int32_t>() + B->row_off[b];
        J.cols = B->d_cols.as<int32_t>() + B->col_off[b];
        J.U = (c128*)(o + L.U);
        J.Vt = (c128*)(o + L.Vt);
        J.row = (c128*)(o + L.row);
        J.used = (int32_t*)(o + L.used);
        J.I = (int32_t*)(o + L.I);
        J.J = (int32_t*)(o + L.J);
        J.k_out = B->d_aca_ks.as<int32_t>() + b;
        J.status = B->d_aca_ks.as<int32_t>() + B->nb + b;
    }
    TRY(B->d_aca_jobs.ensure(jobs.size() * sizeof(AcaJob)));
    hipStream_t st = ctx->stream;
    HIPTRY(hipMemcpyAsync(B->d_aca_jobs.p, jobs.data(), jobs.size() * sizeof(AcaJob), hipMemcpyHostToDevice, st));
    HIPTRY(hipEventRecord(ctx->ev0, st));
    launch_aca(B->d_aca_jobs.as<AcaJob>(), B->nb, tol, ctx->fill_params(0), st);
    HIPTRY(hipGetLastError());
    HIPTRY(hipEventRecord(ctx->ev1, st));
    std::vector<int32_t> ks((size_t)B->nb * 2);
    HIPTRY(hipMemcpyAsync(ks.data(), B->d_aca_ks.p, ks.size() * 4, hipMemcpyDeviceToHost, st));
    HIPTRY(hipStreamSynchronize(st));
    B->aca_k.assign(ks.begin(), ks.begin() + B->nb);
    B->aca_status.assign(ks.begin() + B->nb, ks.end());
    float t = 0;
    HIPTRY(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
    if (ms) *ms = t;
    if (ranks) std::copy(B->aca_k.begin(), B->aca_k.end(), ranks);
    B->aca_ready = true;
    return FARCUR_OK;
}

extern "C" int farcur_batch_aca_block(farcur_batch* B, int b, farcur_c128* U, farcur_c128* V, int* row_pivots,
                                      int* col_pivots, int* k, int* status) {
    if (!B) return fail(FARCUR_EINVAL, "NULL batch");
    if (!B->aca_ready) return fail(FARCUR_ESTATE, "no ACA result");
    if (b < 0 || b >= B->nb) return fail(FARCUR_EINVAL, "block %d out of range", b);
    farcur_ctx* ctx = B->ctx;
    TRY(ctx->select());
    const int m = B->row_off[b + 1] - B->row_off[b], n = B->col_off[b + 1] - B->col_off[b];
    const int kk = B->aca_k[b];
    if (k) *k = kk;
    if (status) *status = B->aca_status[b];
    const AcaLayout L(m, n, B->aca_kmax[b]);
    const char* o = (const char*)B->d_aca.p + B->aca_off[b];
    hipStream_t st = ctx->stream;
    if (kk > 0) {
        if (U) HIPTRY(hipMemcpyAsync(U, o + L.U, (size_t)m * kk * 16, hipMemcpyDeviceToHost, st));
        std::vector<c128> vt;
        if (V) {
            vt.resize((size_t)n * kk);
            HIPTRY(hipMemcpyAsync(vt.data(), o + L.Vt, vt.size() * 16, hipMemcpyDeviceToHost, st));
        }
        if (row_pivots) HIPTRY(hipMemcpyAsync(row_pivots, o + L.I, (size_t)kk * 4, hipMemcpyDeviceToHost, st));
        if (col_pivots) HIPTRY(hipMemcpyAsync(col_pivots, o + L.J, (size_t)kk * 4, hipMemcpyDeviceToHost, st));
        HIPTRY(hipStreamSynchronize(st));
        if (V)  // V (k x n, ld k) from its rows stored as columns of Vt
            for (int l = 0; l < kk; ++l)
                for (int j = 0; j < n; ++j) V[l + (size_t)j * kk] = farcur_c128{vt[j + (size_t)l * n].re, vt[j + (size_t)l * n].im};
    }
    return FARCUR_OK;
}

extern "C" int farcur_batch_profile(farcur_batch* B, int enable) {
    if (!B) return fail(FARCUR_EINVAL, "NULL batch");
    B->profile = enable != 0;
    return FARCUR_OK;
}

extern "C" int farcur_batch_compress(farcur_batch* B, const farcur_opts* opts, int mode, double* wall_ms) {
    const auto t0 = std::chrono::steady_clock::now();  // the whole call, state set-up included
    if (!B) return fail(FARCUR_EINVAL, "NULL batch");
    TRY(validate_opts(opts));
    if (mode != FARCUR_MODE_PLAIN && mode != FARCUR_MODE_BATCHED && mode != FARCUR_MODE_LEARNED)
        return fail(FARCUR_EINVAL, "bad mode %d", mode);
    if (mode == FARCUR_MODE_LEARNED && opts->row_mode != FARCUR_ROWS_UNIFORM)
        return fail(FARCUR_EINVAL, "learned mode samples rows uniformly (row_mode must be FARCUR_ROWS_UNIFORM)");
    if (mode == FARCUR_MODE_LEARNED && opts->heldout < 1)
        return fail(FARCUR_EINVAL, "learned mode needs heldout >= 1: the held-out rows are its acceptance check");
    farcur_ctx* ctx = B->ctx;
    TRY(ctx->select());
    const farcur_opts o = *opts;
    B->f32 = o.precision == FARCUR_F32;
    B->learned = mode == FARCUR_MODE_LEARNED;
    B->first_round_misses = 0;
    B->has_result = false;
    B->rc_ready = false;
    B->interp_ready = false;
    B->rounds.clear();
    B->kt.clear();
    B->kt.on = B->profile;
    for (double& w : B->work) w = 0;
    for (double& h : B->host_ms) h = 0;
    for (int b = 0; b < B->nb; ++b) {
        BState& s = B->st[b];
        s = BState{};
        s.m = B->row_off[b + 1] - B->row_off[b];
        s.n = B->col_off[b + 1] - B->col_off[b];
        const int D = std::min(s.m, s.n);
        s.cap = o.max_rank > 0 ? std::min(D, o.max_rank) : D;
        if (s.cap > kMaxSamples)
            return fail(FARCUR_EINVAL, "block %d: sample cap %d exceeds %d (set max_rank)", b, s.cap, kMaxSamples);
        s.r = std::min(o.r_start > 0 ? o.r_start : r_start_rule(D), s.cap);
    }
    B->arena.reset();
    hipStream_t st = ctx->stream;
    if (mode != FARCUR_MODE_LEARNED) {  // the successive residual's vectors (learned rounds check held-out rows)
        launch_vtest(B->d_v.as<c128>(), B->d_vo.as<int32_t>(), B->nb, o.seed, st);
        launch_ones(B->d_ya.as<c128>(), (size_t)B->row_off[B->nb], st);
        HIPTRY(hipGetLastError());
    }
    if (mode == FARCUR_MODE_LEARNED) {
        // Sample counts from the rank model (device inference), then rounds of
        // every block whose held-out rows still miss tol.
        // Without a model the first round uses the reference's rule (the
        // method without the network, for comparison).
        const auto tp0 = std::chrono::steady_clock::now();
        std::vector<int32_t> cap(B->nb), rp(B->nb);
        for (int b = 0; b < B->nb; ++b) cap[b] = B->st[b].cap;
        if (B->model) {
            HIPTRY(hipMemcpyAsync(B->d_cap.p, cap.data(), B->nb * 4, hipMemcpyHostToDevice, st));
            TRY(model_rank_predict(B->model, B->d_feat.as<float>(), B->d_cap.as<int32_t>(), B->nb, o.oversample,
                                   B->d_rpred.as<int32_t>(), st));
            HIPTRY(hipMemcpyAsync(rp.data(), B->d_rpred.p, B->nb * 4, hipMemcpyDeviceToHost, st));
            HIPTRY(hipStreamSynchronize(st));
        } else {
            for (int b = 0; b < B->nb; ++b) rp[b] = B->st[b].r;
        }
        B->host_ms[0] += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tp0).count();
        std::vector<int> active(B->nb);
        for (int b = 0; b < B->nb; ++b) {
            BState& s = B->st[b];
            s.r = std::max(1, std::min(rp[b], s.cap));
            s.h = heldout_rows(s.m, s.r, o.heldout);
            active[b] = b;
        }
        std::vector<RoundBufs>& bufs = B->bufs_buf;
        std::vector<JobOut>& outs = B->outs_buf;
        bool first = true;
        while (!active.empty()) {
            TRY(batch_round(B, active, o, bufs, outs));
            const auto ta0 = std::chrono::steady_clock::now();
            std::vector<int> next;
            for (size_t i = 0; i < active.size(); ++i) {
                BState& s = B->st[active[i]];
                advance_learned(s, outs[i], bufs[i], o.tol * o.heldout_safety, o.tol, o.heldout, o.growth,
                                o.oversample);
                if (s.status == 0) next.push_back(active[i]);
            }
            B->host_ms[3] += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ta0).count();
            if (first) B->first_round_misses = (int)next.size();
            first = false;
            active.swap(next);
        }
        // blocks whose final sample took every row or every column: stored exactly
        std::vector<ExactJob>& ex = B->exact_jobs;
        ex.clear();
        for (int b = 0; b < B->nb; ++b) {
            const BState& s = B->st[b];
            if (s.fin_r < std::min(s.m, s.n)) continue;
            ex.push_back(ExactJob{s.fin.C, s.fin.U, s.fin.R, s.m, s.n, s.fin_r, s.fin_r == s.m ? 0 : 1});
        }
        if (!ex.empty()) {
            TRY(B->d_exact.ensure(ex.size() * sizeof(ExactJob)));
            HIPTRY(hipMemcpyAsync(B->d_exact.p, ex.data(), ex.size() * sizeof(ExactJob), hipMemcpyHostToDevice, st));
            launch_exact(B->d_exact.as<ExactJob>(), (int)ex.size(), B->f32, st);
            HIPTRY(hipGetLastError());
        }
    } else if (mode == FARCUR_MODE_BATCHED) {
        std::vector<int> active(B->nb);
        for (int b = 0; b < B->nb; ++b) active[b] = b;
        std::vector<RoundBufs>& bufs = B->bufs_buf;
        std::vector<JobOut>& outs = B->outs_buf;
        while (!active.empty()) {
            TRY(batch_round(B, active, o, bufs, outs));
            std::vector<int> next;
            next.reserve(active.size());
            for (size_t i = 0; i < active.size(); ++i) {
                BState& s = B->st[active[i]];
                advance(s, outs[i], bufs[i], o.tol);
                if (s.status == 0) next.push_back(active[i]);
            }
            active.swap(next);
        }
    } else {
        // One block at a time, each round synchronised: the reference's per-block loop.
        for (int b = 0; b < B->nb; ++b) {
            std::vector<int> one{b};
            std::vector<RoundBufs> bufs;
            std::vector<JobOut> outs;
            while (B->st[b].status == 0) {
                TRY(batch_round(B, one, o, bufs, outs));
                advance(B->st[b], outs[0], bufs[0], o.tol);
            }
        }
    }
    HIPTRY(hipStreamSynchronize(st));
    const auto t1 = std::chrono::steady_clock::now();
    if (wall_ms) *wall_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    if (B->profile) TRY(batch_account(B));
    B->has_result = true;
    return FARCUR_OK;
}

// Accuracy of every block of the last compress against its FP64 dense block,
// on the device (verify.hip). Blocks go in chunks that bound the FP64
// interpolation matrices P = C U to ~256 MiB at a time.
extern "C" int farcur_batch_verify(farcur_batch* B, const farcur_c128* V, int nrhs, int symmetric, double* fro_err,
                                   double* mv_err, farcur_c128* Yref, double* ms) {
    if (!B || !V) return fail(FARCUR_EINVAL, "NULL argument");
    if (!B->has_result) return fail(FARCUR_ESTATE, "no compress result to verify");
    if (nrhs < 1 || nrhs > kVerifyMaxRhs) return fail(FARCUR_EINVAL, "nrhs = %d outside [1, %d]", nrhs, kVerifyMaxRhs);
    farcur_ctx* ctx = B->ctx;
    TRY(ctx->select());
    hipStream_t st = ctx->stream;
    const int N = ctx->ne, nb = B->nb;
    const size_t Mtot = (size_t)B->row_off[nb], Ntot = (size_t)B->col_off[nb];
    DBuf dV, dsums, dyerr, dyref, dys, dyst, dP, djobs, dint, dtiles;
    TRY(dV.ensure((size_t)N * nrhs * sizeof(c128)));
    TRY(dsums.ensure((size_t)2 * nb * sizeof(double)));
    TRY(dyerr.ensure(Mtot * sizeof(c128)));
    TRY(dyref.ensure(Mtot * sizeof(c128)));
    if (Yref) {
        TRY(dys.ensure(Mtot * nrhs * sizeof(c128)));
        if (symmetric) TRY(dyst.ensure(Ntot * nrhs * sizeof(c128)));
    }
    HIPTRY(hipMemcpyAsync(dV.p, V, (size_t)N * nrhs * sizeof(c128), hipMemcpyHostToDevice, st));
    HIPTRY(hipMemsetAsync(dsums.p, 0, (size_t)2 * nb * sizeof(double), st));
    HIPTRY(hipEventRecord(ctx->ev0, st));
    const FillParams fp = ctx->fill_params(0);
    const size_t kChunkP = (size_t)256 << 20;
    std::vector<VerifyInterp> ij;
    std::vector<VerifyJob> vj;
    std::vector<int32_t> tj, tr;
    for (int b0 = 0; b0 < nb;) {
        // chunk [b0, b1): P bytes within kChunkP (at least one block)
        size_t pbytes = 0;
        int b1 = b0;
        while (b1 < nb) {
            const size_t need = (size_t)B->st[b1].m * B->st[b1].fin_r * sizeof(c128);
            if (b1 > b0 && pbytes + need > kChunkP) break;
            pbytes += need;
            ++b1;
        }
        TRY(dP.ensure(pbytes));
        ij.clear();
        vj.clear();
        tj.clear();
        tr.clear();
        size_t poff = 0;
        for (int b = b0; b < b1; ++b) {
            const BState& S = B->st[b];
            c128* P = dP.as<c128>() + poff;
            poff += (size_t)S.m * S.fin_r;
            ij.push_back(VerifyInterp{S.fin.C, S.fin.U, P, S.m, S.fin_r});
            const int ro = B->row_off[b], co = B->col_off[b];
            VerifyJob J{};
            J.rows = B->d_rows.as<int32_t>() + ro;
            J.cols = B->d_cols.as<int32_t>() + co;
            J.m = S.m;
            J.n = S.n;
            J.r = S.fin_r;
            J.ref_only = 0;
            J.R = S.fin.R;
            J.P = P;
            J.sums = dsums.as<double>() + 2 * (size_t)b;
            J.yerr = dyerr.as<c128>() + ro;
            J.yref = dyref.as<c128>() + ro;
            J.ys = Yref ? dys.as<c128>() + (size_t)ro * nrhs : nullptr;
            for (int t = 0; t < S.m; t += 64) {
                tj.push_back((int32_t)vj.size());
                tr.push_back(t);
            }
            vj.push_back(J);
            if (Yref && symmetric) {  // the transposed block Z_b^T = Z(cols, rows), product only
                VerifyJob T{};
                T.rows = J.cols;
                T.cols = J.rows;
                T.m = S.n;
                T.n = S.m;
                T.ref_only = 1;
                T.ys = dyst.as<c128>() + (size_t)co * nrhs;
                for (int t = 0; t < S.n; t += 64) {
                    tj.push_back((int32_t)vj.size());
                    tr.push_back(t);
                }
                vj.push_back(T);
            }
        }
        TRY(dint.ensure(ij.size() * sizeof(VerifyInterp)));
        TRY(djobs.ensure(vj.size() * sizeof(VerifyJob)));
        TRY(dtiles.ensure(tj.size() * 2 * sizeof(int32_t)));
        HIPTRY(hipMemcpyAsync(dint.p, ij.data(), ij.size() * sizeof(VerifyInterp), hipMemcpyHostToDevice, st));
        HIPTRY(hipMemcpyAsync(djobs.p, vj.data(), vj.size() * sizeof(VerifyJob), hipMemcpyHostToDevice, st));
        HIPTRY(hipMemcpyAsync(dtiles.p, tj.data(), tj.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
        HIPTRY(hipMemcpyAsync(dtiles.as<int32_t>() + tj.size(), tr.data(), tr.size() * sizeof(int32_t),
                              hipMemcpyHostToDevice, st));
        for (size_t j0 = 0; j0 < ij.size(); j0 += 65535)
            launch_verify_interp(dint.as<VerifyInterp>() + j0, (int)std::min<size_t>(65535, ij.size() - j0), B->f32, st);
        launch_verify(djobs.as<VerifyJob>(), (int)tj.size(), dtiles.as<int32_t>(), dtiles.as<int32_t>() + tj.size(),
                      B->f32, dV.as<c128>(), N, nrhs, fp, st);
        HIPTRY(hipGetLastError());
        // the job and tile arrays are rewritten by the next chunk
        HIPTRY(hipStreamSynchronize(st));
        b0 = b1;
    }
    HIPTRY(hipEventRecord(ctx->ev1, st));
    std::vector<double> sums(2 * (size_t)nb);
    std::vector<c128> ye(Mtot), yr(Mtot);
    HIPTRY(hipMemcpyAsync(sums.data(), dsums.p, sums.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPTRY(hipMemcpyAsync(ye.data(), dyerr.p, Mtot * sizeof(c128), hipMemcpyDeviceToHost, st));
    HIPTRY(hipMemcpyAsync(yr.data(), dyref.p, Mtot * sizeof(c128), hipMemcpyDeviceToHost, st));
    std::vector<c128> ys, yst;
    if (Yref) {
        ys.resize(Mtot * nrhs);
        HIPTRY(hipMemcpyAsync(ys.data(), dys.p, ys.size() * sizeof(c128), hipMemcpyDeviceToHost, st));
        if (symmetric) {
            yst.resize(Ntot * nrhs);
            HIPTRY(hipMemcpyAsync(yst.data(), dyst.p, yst.size() * sizeof(c128), hipMemcpyDeviceToHost, st));
        }
    }
    HIPTRY(hipStreamSynchronize(st));
    float t = 0;
    hipEventElapsedTime(&t, ctx->ev0, ctx->ev1);
    if (ms) *ms = t;
    for (int b = 0; b < nb; ++b) {
        if (fro_err) fro_err[b] = sums[2 * b + 1] > 0 ? std::sqrt(sums[2 * b] / sums[2 * b + 1]) : 0.0;
        if (mv_err) {
            double ne = 0, nr = 0;
            for (int i = B->row_off[b]; i < B->row_off[b + 1]; ++i) {
                ne += ye[i].re * ye[i].re + ye[i].im * ye[i].im;
                nr += yr[i].re * yr[i].re + yr[i].im * yr[i].im;
            }
            mv_err[b] = nr > 0 ? std::sqrt(ne / nr) : 0.0;
        }
    }
    if (Yref) {
        std::fill(Yref, Yref + (size_t)N * nrhs, farcur_c128{0, 0});
        for (int b = 0; b < nb; ++b) {
            const int ro = B->row_off[b], m = B->st[b].m;
            for (int q = 0; q < nrhs; ++q)
                for (int i = 0; i < m; ++i) {
                    const c128 v = ys[(size_t)ro * nrhs + (size_t)q * m + i];
                    farcur_c128& y = Yref[(size_t)q * N + B->row_ids[ro + i]];
                    y.re += v.re;
                    y.im += v.im;
                }
            if (symmetric) {
                const int co = B->col_off[b], n = B->st[b].n;
                for (int q = 0; q < nrhs; ++q)
                    for (int j = 0; j < n; ++j) {
                        const c128 v = yst[(size_t)co * nrhs + (size_t)q * n + j];
                        farcur_c128& y = Yref[(size_t)q * N + B->col_ids[co + j]];
                        y.re += v.re;
                        y.im += v.im;
                    }
            }
        }
    }
    return FARCUR_OK;
}

extern "C" int farcur_batch_info(farcur_batch* B, farcur_block_info* infos) {
    if (!B || !infos) return fail(FARCUR_EINVAL, "NULL argument");
    if (!B->has_result) return fail(FARCUR_ESTATE, "no batch result");
    for (int b = 0; b < B->nb; ++b) {
        const BState& s = B->st[b];
        farcur_block_info& I = infos[b];
        I.r = s.fin_r;
        I.rounds = s.round;
        I.kept = s.kept;
        I.converged = s.status == 1;
        I.err = s.err;
        I.heldout_err = s.heldout_err;
        I.sigma_max = s.smax;
        I.time_ms = 0;
        I.rank_eps = s.rank_eps;
        I.status = s.status;
        I.sweeps = s.sweeps;
        I.pad = 0;
    }
    return FARCUR_OK;
}

extern "C" int farcur_batch_block(farcur_batch* B, int b, farcur_c128* C, farcur_c128* U, farcur_c128* R,
                                  int* row_samples, int* col_samples) {
    if (!B) return fail(FARCUR_EINVAL, "NULL batch");
    if (!B->has_result) return fail(FARCUR_ESTATE, "no batch result");
    if (b < 0 || b >= B->nb) return fail(FARCUR_EINVAL, "block %d out of range", b);
    farcur_ctx* ctx = B->ctx;
    TRY(ctx->select());
    const BState& s = B->st[b];
    const int r = s.fin_r;
    hipStream_t st = ctx->stream;
    TRY(copy_panel(C, s.fin.C, (size_t)s.m * r, B->f32, st));
    TRY(copy_panel(R, s.fin.R, (size_t)r * s.n, B->f32, st));
    TRY(copy_panel(U, s.fin.U, (size_t)r * r, 0, st));
    if (row_samples) HIPTRY(hipMemcpyAsync(row_samples, s.fin.rs, r * 4, hipMemcpyDeviceToHost, st));
    if (col_samples) HIPTRY(hipMemcpyAsync(col_samples, s.fin.cs, r * 4, hipMemcpyDeviceToHost, st));
    HIPTRY(hipStreamSynchronize(st));
    return FARCUR_OK;
}

extern "C" int farcur_batch_kernel_times(farcur_batch* B, const char** names, double* ms, int* launches,
                                         double* work, int cap) {
    if (!B) return fail(FARCUR_EINVAL, "NULL batch");
    static const char* kNames[KTimer::kClasses] = {"sample",   "fill",     "pinv",     "residual",
                                                   "pinv_r16", "pinv_r20", "pinv_r24", "pinv_r28",
                                                   "pinv_r32", "pinv_r36", "pinv_r40", "pinv_r48",
                                                   "pinv_r64", "pinv_r96", "pinv_global"};
    static const char* kHost[4] = {"host_predict", "host_plan", "host_wait", "host_advance"};
    int n = 0;
    for (int i = 0; i < KTimer::kClasses && n < cap; ++i, ++n) {
        if (names) names[n] = kNames[i];
        if (ms) ms[n] = B->kt.ms[i];
        if (launches) launches[n] = B->kt.launches[i];
        if (work) work[n] = B->work[i];
    }
    for (int i = 0; i < 4 && n < cap; ++i, ++n) {  // host wall time, measured on every compress
        if (names) names[n] = kHost[i];
        if (ms) ms[n] = B->host_ms[i];
        if (launches) launches[n] = 0;
        if (work) work[n] = 0;
    }
    return n;
}

// ------------------------------------------------------------------ autotuner
namespace {

// Times the pinv of nprob random r x r problems with the given launch shape
// (min over reps, ms).
// Times the pinv of nprob random r x r cores. The LDS variant (rmax > 0) runs
// as in a batch round: each core gathered from a panel (here the pristine
// copy W0 through identity row positions) with the FP32 pipeline's relative
// threshold 1e-6, i.e. the one-buffer V-free form.
// qr_core: the gathered cores take the truncated QR pseudo-inverse of the
// learned rounds instead of the SVD one.
int time_pinv(farcur_ctx* ctx, int r, int nprob, int threads, int rmax, double* best_ms, double* sweeps = nullptr,
              bool qr_core = false) {
    hipStream_t s = ctx->stream;
    const size_t rr = (size_t)r * r;
    DBuf W0, W, V, U, sig, plan, ids;
    TRY(W0.ensure(nprob * rr * 16));
    TRY(W.ensure(nprob * rr * 16));
    TRY(V.ensure(nprob * pinv_v_elems(r) * 16));
    TRY(U.ensure(nprob * rr * 16));
    TRY(sig.ensure((size_t)nprob * r * 8));
    std::vector<c128> h(nprob * rr);
    uint64_t st = 0x1234567ull + r;
    for (auto& z : h) {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        z.re = (double)(st >> 11) * (1.0 / 9007199254740992.0) - 0.5;
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        z.im = (double)(st >> 11) * (1.0 / 9007199254740992.0) - 0.5;
    }
    HIPTRY(hipMemcpyAsync(W0.p, h.data(), h.size() * 16, hipMemcpyHostToDevice, s));
    std::vector<int32_t> iden(r);
    for (int i = 0; i < r; ++i) iden[i] = i;
    TRY(ids.ensure((size_t)r * 4));
    HIPTRY(hipMemcpyAsync(ids.p, iden.data(), (size_t)r * 4, hipMemcpyHostToDevice, s));
    const bool gathered = rmax > 0;
    std::vector<BlockJob> jobs(nprob);
    std::vector<int32_t> list(nprob);
    for (int i = 0; i < nprob; ++i) {
        BlockJob& J = jobs[i];
        J = BlockJob{};
        J.r = r;
        J.W = W.as<c128>() + i * rr;
        J.V = V.as<c128>() + i * pinv_v_elems(r);
        J.U = U.as<c128>() + i * rr;
        J.sigma = sig.as<double>() + (size_t)i * r;
        if (gathered) {
            J.m = r;
            J.C = W0.as<c128>() + i * rr;
            J.rpos = ids.as<int32_t>();
        }
        list[i] = i;
    }
    const size_t bj = nprob * sizeof(BlockJob), bl = (size_t)nprob * 4, bo = (size_t)nprob * sizeof(JobOut);
    TRY(plan.ensure(bj + bl + bo + 512));
    char* d = (char*)plan.p;
    HIPTRY(hipMemcpyAsync(d, jobs.data(), bj, hipMemcpyHostToDevice, s));
    HIPTRY(hipMemcpyAsync(d + bj, list.data(), bl, hipMemcpyHostToDevice, s));
    *best_ms = 1e30;
    for (int rep = 0; rep < 3; ++rep) {
        HIPTRY(hipMemcpyAsync(W.p, W0.p, nprob * rr * 16, hipMemcpyDeviceToDevice, s));
        HIPTRY(hipEventRecord(ctx->ev0, s));
        launch_pinv((const BlockJob*)d, (const int32_t*)(d + bj), nprob, 1e-10, gathered ? 1e-6 : 0.0, 0.0, rmax,
                    threads, 0, gathered ? 1 : 0, 0, gathered && qr_core ? 2 : 0, (JobOut*)(d + bj + bl + 256), s);
        HIPTRY(hipEventRecord(ctx->ev1, s));
        HIPTRY(hipGetLastError());
        HIPTRY(hipStreamSynchronize(s));
        float ms = 0;
        HIPTRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        *best_ms = std::min(*best_ms, (double)ms);
    }
    if (sweeps) {
        std::vector<JobOut> o(nprob);
        HIPTRY(hipMemcpy(o.data(), d + bj + bl + 256, nprob * sizeof(JobOut), hipMemcpyDeviceToHost));
        double a = 0;
        for (auto& x : o) a += x.sweeps;
        *sweeps = a / nprob;
    }
    return FARCUR_OK;
}

}  // namespace

extern "C" int farcur_pinv_bench(farcur_ctx* ctx, int r, int nprob, int threads, int use_lds, double* ms,
                                 double* sweeps) {
    if (!ctx || !ms) return fail(FARCUR_EINVAL, "NULL argument");
    if (r <= 0 || r > 1024 || nprob <= 0 || nprob > 65536) return fail(FARCUR_EINVAL, "bad sizes");
    if (threads <= 0 || threads > 1024 || threads % 64) return fail(FARCUR_EINVAL, "threads must be a multiple of 64");
    if (use_lds && r > kPinvLdsOneBufMaxR)
        return fail(FARCUR_EINVAL, "LDS variant needs r <= %d (above %d: one-buffer form)", kPinvLdsOneBufMaxR,
                    kPinvLdsMaxR);
    TRY(ctx->select());
    return time_pinv(ctx, r, nprob, threads, use_lds ? r : 0, ms, sweeps);
}

extern "C" int farcur_autotune(farcur_ctx* ctx, farcur_tuning* out) {
    if (!ctx) return fail(FARCUR_EINVAL, "NULL context");
    TRY(ctx->select());
    Tuning t = ctx->tune;
    farcur_tuning rep{};
    for (int c = 0; c < kPinvClasses; ++c) {
        const bool glb = c == kPinvClasses - 1;
        const int r = glb ? 128 : kPinvClassMax[c];
        const int nprob = glb ? 1024 : 4096;
        // cores up to 32: 64..256 threads; larger LDS classes and the global
        // variant: 128..512 (64 threads would leave one lane pair per column pair)
        const int base = glb || r > 32 ? 128 : 64;
        double best = 1e30;
        for (int k = 0; k < 3; ++k) {
            const int th = base << k;
            double ms;
            // LDS classes as the learned rounds run them (truncated QR core)
            TRY(time_pinv(ctx, r, nprob, th, kPinvClassMax[c], &ms, nullptr, kPinvClassMax[c] > 0));
            if (ms < best) {
                best = ms;
                t.pinv_threads[c] = th;
            }
        }
        rep.pinv_threads[c] = t.pinv_threads[c];
        rep.pinv_ms[c] = best;
    }
    ctx->tune = t;
    rep.fill_grid = t.fill_grid;
    rep.resid_threads = t.resid_threads;
    if (out) *out = rep;
    return FARCUR_OK;
}

extern "C" int farcur_get_tuning(farcur_ctx* ctx, farcur_tuning* out) {
    if (!ctx || !out) return fail(FARCUR_EINVAL, "NULL argument");
    *out = farcur_tuning{};
    for (int c = 0; c < kPinvClasses; ++c) out->pinv_threads[c] = ctx->tune.pinv_threads[c];
    out->fill_grid = ctx->tune.fill_grid;
    out->resid_threads = ctx->tune.resid_threads;
    return FARCUR_OK;
}

extern "C" int farcur_set_tuning(farcur_ctx* ctx, const farcur_tuning* in) {
    if (!ctx || !in) return fail(FARCUR_EINVAL, "NULL argument");
    for (int c = 0; c < kPinvClasses; ++c)
        if (in->pinv_threads[c] <= 0 || in->pinv_threads[c] > 1024 || in->pinv_threads[c] % 64)
            return fail(FARCUR_EINVAL, "pinv_threads[%d] must be a multiple of 64 in [64, 1024]", c);
    if (in->fill_grid <= 0 || in->resid_threads <= 0 || in->resid_threads > 1024 || in->resid_threads % 64)
        return fail(FARCUR_EINVAL, "bad fill_grid / resid_threads");
    for (int c = 0; c < kPinvClasses; ++c) ctx->tune.pinv_threads[c] = in->pinv_threads[c];
    ctx->tune.fill_grid = in->fill_grid;
    ctx->tune.resid_threads = in->resid_threads;
    return FARCUR_OK;
}

// ------------------------------------------------------------------ rank model
struct farcur_model {
    farcur_ctx* ctx = nullptr;
    DBuf params, m, v, grad, loss, X, y, idx, partial, ploss;
    MlpNorm norm{};
    bool has_norm = false;
    int step = 0;
};

int model_rank_predict(farcur_model* M, const float* d_X, const int32_t* d_cap, int n, int p, int32_t* d_r,
                       hipStream_t s) {
    if (!M->has_norm) return fail(FARCUR_ESTATE, "model has no normalisation (train or set it first)");
    launch_rank_predict(M->params.as<float>(), M->norm, d_X, d_cap, n, p, d_r, s);
    HIPTRY(hipGetLastError());
    return FARCUR_OK;
}

namespace {

void norm_from_data(const float* X, const float* y, int n, MlpNorm& nrm) {
    for (int c = 0; c < kMlpFeatures; ++c) {
        double s1 = 0, s2 = 0;
        for (int i = 0; i < n; ++i) {
            const double x = X[(size_t)i * kMlpFeatures + c];
            s1 += x;
            s2 += x * x;
        }
        const double mean = s1 / n, var = std::max(0.0, s2 / n - mean * mean);
        nrm.xm[c] = (float)mean;
        nrm.xs[c] = (float)std::max(1e-6, std::sqrt(var));
    }
    double s1 = 0, s2 = 0;
    for (int i = 0; i < n; ++i) {
        s1 += y[i];
        s2 += (double)y[i] * y[i];
    }
    const double mean = s1 / n, var = std::max(0.0, s2 / n - mean * mean);
    nrm.ym = (float)mean;
    nrm.ys = (float)std::max(1e-6, std::sqrt(var));
}

int model_upload_data(farcur_model* M, const float* X, const float* y, int n) {
    hipStream_t s = M->ctx->stream;
    TRY(M->X.ensure((size_t)n * kMlpFeatures * 4));
    TRY(M->y.ensure((size_t)n * 4));
    HIPTRY(hipMemcpyAsync(M->X.p, X, (size_t)n * kMlpFeatures * 4, hipMemcpyHostToDevice, s));
    if (y) HIPTRY(hipMemcpyAsync(M->y.p, y, (size_t)n * 4, hipMemcpyHostToDevice, s));
    return FARCUR_OK;
}

int check_data(const float* X, const float* y, int n) {
    for (size_t i = 0; i < (size_t)n * kMlpFeatures; ++i)
        if (!std::isfinite(X[i])) return fail(FARCUR_EINVAL, "non-finite feature at %zu", i);
    if (y)
        for (int i = 0; i < n; ++i)
            if (!std::isfinite(y[i])) return fail(FARCUR_EINVAL, "non-finite target at %d", i);
    return FARCUR_OK;
}

}  // namespace

extern "C" int farcur_model_create(farcur_ctx* ctx, uint64_t seed, farcur_model** out) {
    if (!ctx || !out) return fail(FARCUR_EINVAL, "NULL argument");
    TRY(ctx->select());
    auto M = std::make_unique<farcur_model>();
    M->ctx = ctx;
    // Glorot-uniform weights, zero biases.
    std::vector<float> p(kMlpParams, 0.0f);
    SplitMixHost rng{seed};
    auto fill_w = [&](int off, int fan_out, int fan_in) {
        const double a = std::sqrt(6.0 / (fan_in + fan_out));
        for (int i = 0; i < fan_out * fan_in; ++i) p[off + i] = (float)((2.0 * rng.uniform() - 1.0) * a);
    };
    const int H = kMlpHidden, F = kMlpFeatures;
    fill_w(0, H, F);
    fill_w(H * F + H, H, H);
    fill_w(H * F + H + H * H + H, 1, H);
    TRY(M->params.ensure(kMlpParams * 4));
    TRY(M->m.ensure(kMlpParams * 4));
    TRY(M->v.ensure(kMlpParams * 4));
    TRY(M->grad.ensure(kMlpParams * 4));
    HIPTRY(hipMemcpyAsync(M->params.p, p.data(), kMlpParams * 4, hipMemcpyHostToDevice, ctx->stream));
    HIPTRY(hipMemsetAsync(M->m.p, 0, kMlpParams * 4, ctx->stream));
    HIPTRY(hipMemsetAsync(M->v.p, 0, kMlpParams * 4, ctx->stream));
    HIPTRY(hipStreamSynchronize(ctx->stream));
    *out = M.release();
    return FARCUR_OK;
}

extern "C" void farcur_model_destroy(farcur_model* M) {
    if (!M) return;
    hipSetDevice(M->ctx->device);
    hipStreamSynchronize(M->ctx->stream);
    DBuf* bufs[] = {&M->params, &M->m, &M->v, &M->grad, &M->loss, &M->X, &M->y, &M->idx, &M->partial, &M->ploss};
    for (DBuf* b : bufs) b->release();
    delete M;
}

extern "C" int farcur_model_get(farcur_model* M, float* params, float* norm) {
    if (!M) return fail(FARCUR_EINVAL, "NULL model");
    TRY(M->ctx->select());
    if (params) {
        HIPTRY(hipMemcpyAsync(params, M->params.p, kMlpParams * 4, hipMemcpyDeviceToHost, M->ctx->stream));
        HIPTRY(hipStreamSynchronize(M->ctx->stream));
    }
    if (norm) {
        for (int c = 0; c < kMlpFeatures; ++c) {
            norm[c] = M->norm.xm[c];
            norm[kMlpFeatures + c] = M->norm.xs[c];
        }
        norm[2 * kMlpFeatures] = M->norm.ym;
        norm[2 * kMlpFeatures + 1] = M->norm.ys;
    }
    return FARCUR_OK;
}

extern "C" int farcur_model_set(farcur_model* M, const float* params, const float* norm) {
    if (!M) return fail(FARCUR_EINVAL, "NULL model");
    TRY(M->ctx->select());
    if (params) {
        for (int i = 0; i < kMlpParams; ++i)
            if (!std::isfinite(params[i])) return fail(FARCUR_EINVAL, "non-finite parameter %d", i);
        HIPTRY(hipMemcpyAsync(M->params.p, params, kMlpParams * 4, hipMemcpyHostToDevice, M->ctx->stream));
        HIPTRY(hipMemsetAsync(M->m.p, 0, kMlpParams * 4, M->ctx->stream));
        HIPTRY(hipMemsetAsync(M->v.p, 0, kMlpParams * 4, M->ctx->stream));
        HIPTRY(hipStreamSynchronize(M->ctx->stream));
        M->step = 0;
    }
    if (norm) {
        for (int c = 0; c < kMlpFeatures; ++c) {
            if (!(norm[kMlpFeatures + c] > 0)) return fail(FARCUR_EINVAL, "feature scale %d must be > 0", c);
            M->norm.xm[c] = norm[c];
            M->norm.xs[c] = norm[kMlpFeatures + c];
        }
        if (!(norm[2 * kMlpFeatures + 1] > 0)) return fail(FARCUR_EINVAL, "target scale must be > 0");
        M->norm.ym = norm[2 * kMlpFeatures];
        M->norm.ys = norm[2 * kMlpFeatures + 1];
        M->has_norm = true;
    }
    return FARCUR_OK;
}

extern "C" int farcur_model_loss_grad(farcur_model* M, const float* X, const float* y, int n, float* loss,
                                      float* grad) {
    if (!M || !X || !y || n <= 0) return fail(FARCUR_EINVAL, "bad arguments");
    TRY(check_data(X, y, n));
    TRY(M->ctx->select());
    hipStream_t s = M->ctx->stream;
    if (!M->has_norm) {
        norm_from_data(X, y, n, M->norm);
        M->has_norm = true;
    }
    TRY(model_upload_data(M, X, y, n));
    std::vector<int32_t> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    const int ntiles = (n + 63) / 64;
    TRY(M->idx.ensure((size_t)n * 4));
    TRY(M->partial.ensure((size_t)ntiles * kMlpParams * 4));
    TRY(M->ploss.ensure((size_t)ntiles * 4));
    TRY(M->loss.ensure(4));
    HIPTRY(hipMemcpyAsync(M->idx.p, idx.data(), (size_t)n * 4, hipMemcpyHostToDevice, s));
    launch_mlp_train_tiles(M->params.as<float>(), M->norm, M->X.as<float>(), M->y.as<float>(), M->idx.as<int32_t>(),
                           n, M->partial.as<float>(), M->ploss.as<float>(), s);
    launch_mlp_reduce(M->partial.as<float>(), M->ploss.as<float>(), n, M->grad.as<float>(), M->loss.as<float>(), s);
    HIPTRY(hipGetLastError());
    if (grad) HIPTRY(hipMemcpyAsync(grad, M->grad.p, kMlpParams * 4, hipMemcpyDeviceToHost, s));
    float l = 0;
    HIPTRY(hipMemcpyAsync(&l, M->loss.p, 4, hipMemcpyDeviceToHost, s));
    HIPTRY(hipStreamSynchronize(s));
    if (loss) *loss = l;
    return FARCUR_OK;
}

extern "C" int farcur_model_train(farcur_model* M, const float* X, const float* y, int n, int epochs, int batch,
                                  float lr, uint64_t seed, float* epoch_loss) {
    if (!M || !X || !y || n <= 0 || epochs <= 0 || batch <= 0 || !(lr > 0)) return fail(FARCUR_EINVAL, "bad arguments");
    TRY(check_data(X, y, n));
    TRY(M->ctx->select());
    hipStream_t s = M->ctx->stream;
    if (!M->has_norm) {
        norm_from_data(X, y, n, M->norm);
        M->has_norm = true;
    }
    TRY(model_upload_data(M, X, y, n));
    batch = std::min(batch, n);
    const int ntiles = (batch + 63) / 64;
    const int steps_per_epoch = (n + batch - 1) / batch;
    TRY(M->idx.ensure((size_t)n * 4));
    TRY(M->partial.ensure((size_t)ntiles * kMlpParams * 4));
    TRY(M->ploss.ensure((size_t)ntiles * 4));
    TRY(M->loss.ensure((size_t)steps_per_epoch * 4));
    std::vector<int32_t> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    SplitMixHost rng{seed};
    std::vector<float> losses(steps_per_epoch);
    for (int ep = 0; ep < epochs; ++ep) {
        for (int i = n - 1; i > 0; --i) std::swap(idx[i], idx[(int)(rng.next() % (uint64_t)(i + 1))]);
        HIPTRY(hipMemcpyAsync(M->idx.p, idx.data(), (size_t)n * 4, hipMemcpyHostToDevice, s));
        for (int st = 0; st < steps_per_epoch; ++st) {
            const int off = st * batch, nb = std::min(batch, n - off);
            launch_mlp_train_tiles(M->params.as<float>(), M->norm, M->X.as<float>(), M->y.as<float>(),
                                   M->idx.as<int32_t>() + off, nb, M->partial.as<float>(), M->ploss.as<float>(), s);
            launch_mlp_reduce(M->partial.as<float>(), M->ploss.as<float>(), nb, M->grad.as<float>(),
                              M->loss.as<float>() + st, s);
            launch_adam(M->params.as<float>(), M->grad.as<float>(), M->m.as<float>(), M->v.as<float>(), ++M->step, lr,
                        s);
        }
        HIPTRY(hipGetLastError());
        HIPTRY(hipMemcpyAsync(losses.data(), M->loss.p, steps_per_epoch * 4, hipMemcpyDeviceToHost, s));
        HIPTRY(hipStreamSynchronize(s));
        if (epoch_loss) {
            double a = 0;
            for (int st = 0; st < steps_per_epoch; ++st) a += losses[st] * std::min(batch, n - st * batch);
            epoch_loss[ep] = (float)(a / n);
        }
    }
    return FARCUR_OK;
}

extern "C" int farcur_model_predict(farcur_model* M, const float* X, int n, float* y) {
    if (!M || !X || !y || n <= 0) return fail(FARCUR_EINVAL, "bad arguments");
    if (!M->has_norm) return fail(FARCUR_ESTATE, "model has no normalisation");
    TRY(check_data(X, nullptr, n));
    TRY(M->ctx->select());
    hipStream_t s = M->ctx->stream;
    TRY(model_upload_data(M, X, nullptr, n));
    TRY(M->y.ensure((size_t)n * 4));
    launch_mlp_predict(M->params.as<float>(), M->norm, M->X.as<float>(), n, M->y.as<float>(), s);
    HIPTRY(hipGetLastError());
    HIPTRY(hipMemcpyAsync(y, M->y.p, (size_t)n * 4, hipMemcpyDeviceToHost, s));
    HIPTRY(hipStreamSynchronize(s));
    return FARCUR_OK;
}

extern "C" int farcur_batch_features(farcur_batch* B, float* X) {
    if (!B || !X) return fail(FARCUR_EINVAL, "NULL argument");
    std::copy(B->feat.begin(), B->feat.end(), X);
    return FARCUR_OK;
}

extern "C" int farcur_batch_set_model(farcur_batch* B, farcur_model* M) {
    if (!B) return fail(FARCUR_EINVAL, "NULL batch");
    if (M && M->ctx != B->ctx) return fail(FARCUR_EINVAL, "model and batch belong to different contexts");
    B->model = M;
    return FARCUR_OK;
}

extern "C" int farcur_batch_misses(farcur_batch* B) { return B ? B->first_round_misses : -1; }

extern "C" int farcur_batch_set_graphs(farcur_batch* B, int enable) {
    if (!B) return fail(FARCUR_EINVAL, "NULL batch");
    TRY(B->ctx->select());
    HIPTRY(hipStreamSynchronize(B->ctx->stream));
    if (!enable)
        for (RoundGraph& g : B->graphs) g.release();
    if (!enable) B->graphs.clear();
    B->use_graphs = enable != 0;
    return FARCUR_OK;
}

extern "C" int farcur_batch_graph_stats(farcur_batch* B, int* captures, int* replays) {
    if (!B) return fail(FARCUR_EINVAL, "NULL batch");
    if (captures) *captures = B->graph_captures;
    if (replays) *replays = B->graph_replays;
    return FARCUR_OK;
}

extern "C" int farcur_batch_set_pipeline(farcur_batch* B, int chunks) {
    if (!B) return fail(FARCUR_EINVAL, "NULL batch");
    if (chunks < 0 || chunks > 64) return fail(FARCUR_EINVAL, "chunks must be in [0, 64]");
    B->pipeline_chunks = chunks < 1 ? 1 : chunks;
    return FARCUR_OK;
}

// ------------------------------------------------------------------ multi-RHS product
namespace {

// Interpolation matrices P_b = C_b U_b of the last compress (device, panel
// precision), formed on the first product call.
int batch_interp(farcur_batch* B, double* ms) {
    if (B->interp_ready) {
        *ms = 0;
        return FARCUR_OK;
    }
    hipStream_t st = B->ctx->stream;
    std::vector<MrhsInterp> jobs(B->nb);
    B->interp.assign(B->nb, nullptr);
    for (int b = 0; b < B->nb; ++b) {
        const BState& S = B->st[b];
        void* p;
        TRY(B->arena.alloc((size_t)S.m * S.fin_r * psize(B->f32), &p));
        B->interp[b] = p;
        jobs[b] = MrhsInterp{S.fin.C, S.fin.U, p, S.m, S.fin_r};
    }
    DBuf dj;
    TRY(dj.ensure(jobs.size() * sizeof(MrhsInterp)));
    HIPTRY(hipMemcpyAsync(dj.p, jobs.data(), jobs.size() * sizeof(MrhsInterp), hipMemcpyHostToDevice, st));
    hipEvent_t e0, e1;
    HIPTRY(hipEventCreate(&e0));
    HIPTRY(hipEventCreate(&e1));
    HIPTRY(hipEventRecord(e0, st));
    launch_mrhs_interp(dj.as<MrhsInterp>(), B->nb, B->f32, st);
    HIPTRY(hipEventRecord(e1, st));
    HIPTRY(hipGetLastError());
    HIPTRY(hipStreamSynchronize(st));
    float t = 0;
    hipEventElapsedTime(&t, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    *ms = t;
    B->interp_ready = true;
    return FARCUR_OK;
}

// 64-row tiles of a list of problems.
void add_tiles(int prob, int m, std::vector<int32_t>& tp, std::vector<int32_t>& tr) {
    for (int r0 = 0; r0 < m; r0 += 64) {
        tp.push_back(prob);
        tr.push_back(r0);
    }
}

}  // namespace

extern "C" int farcur_batch_matvec(farcur_batch* B, const farcur_c128* V, int s, farcur_c128* Y, int mode,
                                   int symmetric, double* ms) {
    if (!B || !V || !Y) return fail(FARCUR_EINVAL, "NULL argument");
    if (!B->has_result) return fail(FARCUR_ESTATE, "no batch result (compress first)");
    if (s <= 0) return fail(FARCUR_EINVAL, "nrhs must be > 0");
    if (mode != kMrhsValu && mode != kMrhsMfmaF32 && mode != kMrhsMfmaBf16x3) return fail(FARCUR_EINVAL, "bad mode");
    farcur_ctx* ctx = B->ctx;
    TRY(ctx->select());
    hipStream_t st = ctx->stream;
    const int N = ctx->ne;
    double ms_interp = 0;
    TRY(batch_interp(B, &ms_interp));
    // Groups of RWG id lists. A group is an id list in one order (a
    // contribution's rows follow it). Target groups (rows of Y written) must
    // be disjoint or identical, so that every row of Y has one writer; source
    // groups (rows of V read) need only be identified.
    struct Registry {
        bool exclusive;
        std::vector<const int32_t*> ids_dev, ids_host;
        std::vector<int> m;
        std::unordered_map<uint64_t, std::vector<int>> by_hash;
        std::vector<int> owner;
        int of(const int32_t* host_ids, const int32_t* dev_ids, int cnt, int* out) {
            uint64_t h = fnv1a(0xcbf29ce484222325ull, host_ids, (size_t)cnt * 4) ^ (uint64_t)cnt;
            for (int g : by_hash[h]) {
                if (m[g] == cnt && std::memcmp(ids_host[g], host_ids, (size_t)cnt * 4) == 0) {
                    *out = g;
                    return FARCUR_OK;
                }
            }
            const int g = (int)m.size();
            if (exclusive) {
                for (int i = 0; i < cnt; ++i) {
                    if (owner[host_ids[i]] != -1)
                        return fail(FARCUR_EINVAL,
                                    "row/column sets of the blocks overlap without being identical lists (RWG %d)",
                                    host_ids[i]);
                    owner[host_ids[i]] = g;
                }
            }
            by_hash[h].push_back(g);
            m.push_back(cnt);
            ids_dev.push_back(dev_ids);
            ids_host.push_back(host_ids);
            *out = g;
            return FARCUR_OK;
        }
    };
    Registry tgt{true}, src{false};
    tgt.owner.assign(N, -1);
    // Contributions (block, orientation): source group (rows of V read) and
    // target group (rows of Y written).
    struct Contrib {
        int b, o, r, src, tgt;
        const void* a1;  // pass A operand (r x |src|, column-major: R or P^T)
        const void* l;   // pass B operand, read transposed (P^T or R)
        size_t t2_off;
    };
    std::vector<Contrib> cons;
    for (int b = 0; b < B->nb; ++b) {
        const BState& S = B->st[b];
        const int32_t* hrows = &B->row_ids[B->row_off[b]];
        const int32_t* hcols = &B->col_ids[B->col_off[b]];
        const int32_t* drows = B->d_rows.as<int32_t>() + B->row_off[b];
        const int32_t* dcols = B->d_cols.as<int32_t>() + B->col_off[b];
        // o = 0: Y[rows] += P (R V[cols]);  o = 1: Y[cols] += R^T (P^T V[rows])
        int s0, t0;
        TRY(src.of(hcols, dcols, S.n, &s0));
        TRY(tgt.of(hrows, drows, S.m, &t0));
        cons.push_back(Contrib{b, 0, S.fin_r, s0, t0, S.fin.R, B->interp[b], 0});
        if (symmetric) {
            int s1, t1;
            TRY(src.of(hrows, drows, S.m, &s1));
            TRY(tgt.of(hcols, dcols, S.n, &t1));
            cons.push_back(Contrib{b, 1, S.fin_r, s1, t1, B->interp[b], S.fin.R, 0});
        }
    }
    const int nsrc = (int)src.m.size(), ntgt = (int)tgt.m.size();
    // Pass A: one row-segmented problem per source group, its contributions'
    // rows stacked (T of a contribution = r consecutive rows of the problem).
    std::vector<std::vector<int>> by_src(nsrc), by_tgt(ntgt);
    for (int c = 0; c < (int)cons.size(); ++c) {
        by_src[cons[c].src].push_back(c);
        by_tgt[cons[c].tgt].push_back(c);
    }
    std::vector<MrhsProb> pa;
    std::vector<MrhsRowSeg> rsa;
    std::vector<int32_t> ta_p, ta_r;
    size_t t2_total = 0;
    for (int g = 0; g < nsrc; ++g) {
        if (by_src[g].empty()) continue;
        MrhsProb P{};
        P.out_off = t2_total;
        P.b_ids = src.ids_dev[g];
        P.k = src.m[g];
        P.seg_begin = (int)rsa.size();
        int rows = 0;
        for (int c : by_src[g]) {
            Contrib& C = cons[c];
            C.t2_off = t2_total + (size_t)rows * s;
            rsa.push_back(MrhsRowSeg{C.a1, C.r, C.r, rows, 0});
            rows += C.r;
        }
        P.seg_end = (int)rsa.size();
        P.m = rows;
        const int prob = (int)pa.size();
        pa.push_back(P);
        add_tiles(prob, rows, ta_p, ta_r);
        t2_total += (size_t)rows * s;
    }
    // Pass B: one K-segmented problem per target group.
    std::vector<MrhsProb> pb;
    std::vector<MrhsSeg> sb;
    std::vector<int32_t> tb_p, tb_r;
    for (int g = 0; g < ntgt; ++g) {
        if (by_tgt[g].empty()) continue;
        MrhsProb P{};
        P.out_ids = tgt.ids_dev[g];
        P.m = tgt.m[g];
        P.seg_begin = (int)sb.size();
        int kbase = 0;
        for (int c : by_tgt[g]) {
            const Contrib& C = cons[c];
            sb.push_back(MrhsSeg{C.l, C.r, 1, C.r, kbase, nullptr, C.t2_off});
            kbase += C.r;
        }
        P.seg_end = (int)sb.size();
        P.k = kbase;
        const int prob = (int)pb.size();
        pb.push_back(P);
        add_tiles(prob, tgt.m[g], tb_p, tb_r);
    }
    // Device buffers: V and Y row-major complex64, T, descriptors.
    DBuf dV, dY, dT, dpa, drsa, dtap, dtar, dpb, dsb, dtbp, dtbr;
    TRY(dV.ensure((size_t)N * s * 8));
    TRY(dY.ensure((size_t)N * s * 8));
    TRY(dT.ensure(std::max<size_t>(t2_total, 1) * 8));
    auto up = [&](DBuf& d, const void* h, size_t bytes) -> int {
        TRY(d.ensure(std::max<size_t>(bytes, 4)));
        if (bytes) HIPTRY(hipMemcpyAsync(d.p, h, bytes, hipMemcpyHostToDevice, st));
        return FARCUR_OK;
    };
    TRY(up(dpa, pa.data(), pa.size() * sizeof(MrhsProb)));
    TRY(up(drsa, rsa.data(), rsa.size() * sizeof(MrhsRowSeg)));
    TRY(up(dtap, ta_p.data(), ta_p.size() * 4));
    TRY(up(dtar, ta_r.data(), ta_r.size() * 4));
    TRY(up(dpb, pb.data(), pb.size() * sizeof(MrhsProb)));
    TRY(up(dsb, sb.data(), sb.size() * sizeof(MrhsSeg)));
    TRY(up(dtbp, tb_p.data(), tb_p.size() * 4));
    TRY(up(dtbr, tb_r.data(), tb_r.size() * 4));
    std::vector<float2> hv((size_t)N * s);
    for (int j = 0; j < s; ++j)
        for (int i = 0; i < N; ++i)
            hv[(size_t)i * s + j] = make_float2((float)V[(size_t)j * N + i].re, (float)V[(size_t)j * N + i].im);
    HIPTRY(hipMemcpyAsync(dV.p, hv.data(), hv.size() * 8, hipMemcpyHostToDevice, st));
    HIPTRY(hipMemsetAsync(dY.p, 0, (size_t)N * s * 8, st));
    hipEvent_t e0, e1, e2;
    HIPTRY(hipEventCreate(&e0));
    HIPTRY(hipEventCreate(&e1));
    HIPTRY(hipEventCreate(&e2));
    HIPTRY(hipEventRecord(e0, st));
    launch_mrhs_gemm(dpa.as<MrhsProb>(), (int)ta_p.size(), dtap.as<int32_t>(), dtar.as<int32_t>(), nullptr,
                     drsa.as<MrhsRowSeg>(), B->f32, mode, dV.as<float2>(), s, dT.as<float2>(), st);
    HIPTRY(hipEventRecord(e1, st));
    launch_mrhs_gemm(dpb.as<MrhsProb>(), (int)tb_p.size(), dtbp.as<int32_t>(), dtbr.as<int32_t>(), dsb.as<MrhsSeg>(),
                     nullptr, B->f32, mode, dT.as<float2>(), s, dY.as<float2>(), st);
    HIPTRY(hipEventRecord(e2, st));
    HIPTRY(hipGetLastError());
    std::vector<float2> hy((size_t)N * s);
    HIPTRY(hipMemcpyAsync(hy.data(), dY.p, hy.size() * 8, hipMemcpyDeviceToHost, st));
    HIPTRY(hipStreamSynchronize(st));
    float ta = 0, tb = 0;
    hipEventElapsedTime(&ta, e0, e1);
    hipEventElapsedTime(&tb, e1, e2);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipEventDestroy(e2);
    if (ms) {
        ms[0] = ta;
        ms[1] = tb;
        ms[2] = ms_interp;
    }
    for (int j = 0; j < s; ++j)
        for (int i = 0; i < N; ++i) Y[(size_t)j * N + i] = farcur_c128{hy[(size_t)i * s + j].x, hy[(size_t)i * s + j].y};
    return FARCUR_OK;
}

// Single right-hand side of the compressed far field, y = sum_b P_b R_b v[cols_b]
// (+ transposes), one workgroup per (block, orientation): the HBM-bound CUR
// matvec (each panel read once).
extern "C" int farcur_batch_matvec1(farcur_batch* B, const farcur_c128* v, farcur_c128* y, int symmetric,
                                    double* ms) {
    if (!B || !v || !y) return fail(FARCUR_EINVAL, "NULL argument");
    if (!B->has_result) return fail(FARCUR_ESTATE, "no batch result (compress first)");
    farcur_ctx* ctx = B->ctx;
    TRY(ctx->select());
    hipStream_t st = ctx->stream;
    const int N = ctx->ne;
    double ms_interp = 0;
    TRY(batch_interp(B, &ms_interp));
    std::vector<Mv1Job> jobs;
    jobs.reserve((size_t)B->nb * (symmetric ? 2 : 1));
    int rmax = 1;
    for (int b = 0; b < B->nb; ++b) {
        const BState& S = B->st[b];
        if (S.fin_r > 4096) return fail(FARCUR_EINVAL, "block %d: r = %d above 4096", b, S.fin_r);
        rmax = std::max(rmax, S.fin_r);
        const int32_t* rows = B->d_rows.as<int32_t>() + B->row_off[b];
        const int32_t* cols = B->d_cols.as<int32_t>() + B->col_off[b];
        jobs.push_back(Mv1Job{S.fin.R, B->interp[b], cols, rows, S.fin_r, S.n, S.m, 0});
        if (symmetric) jobs.push_back(Mv1Job{B->interp[b], S.fin.R, rows, cols, S.fin_r, S.m, S.n, 0});
    }
    std::vector<c64> xv(N);
    for (int i = 0; i < N; ++i) xv[i] = c64{(float)v[i].re, (float)v[i].im};
    DBuf dj, dx, dy;
    TRY(dj.ensure(jobs.size() * sizeof(Mv1Job)));
    TRY(dx.ensure((size_t)N * sizeof(c64)));
    TRY(dy.ensure((size_t)N * sizeof(c64)));
    HIPTRY(hipMemcpyAsync(dj.p, jobs.data(), jobs.size() * sizeof(Mv1Job), hipMemcpyHostToDevice, st));
    HIPTRY(hipMemcpyAsync(dx.p, xv.data(), (size_t)N * sizeof(c64), hipMemcpyHostToDevice, st));
    HIPTRY(hipMemsetAsync(dy.p, 0, (size_t)N * sizeof(c64), st));
    HIPTRY(hipEventRecord(ctx->ev0, st));
    launch_matvec1(dj.as<Mv1Job>(), (int)jobs.size(), rmax, B->f32, dx.as<float2>(), dy.as<float2>(), st);
    HIPTRY(hipEventRecord(ctx->ev1, st));
    HIPTRY(hipGetLastError());
    HIPTRY(hipMemcpyAsync(xv.data(), dy.p, (size_t)N * sizeof(c64), hipMemcpyDeviceToHost, st));
    HIPTRY(hipStreamSynchronize(st));
    float t = 0;
    hipEventElapsedTime(&t, ctx->ev0, ctx->ev1);
    if (ms) *ms = t;
    for (int i = 0; i < N; ++i) y[i] = farcur_c128{xv[i].re, xv[i].im};
    return FARCUR_OK;
}

// ------------------------------------------------------------------ near field and operator
struct farcur_near {
    farcur_ctx* ctx = nullptr;
    int nb = 0;
    double near_factor = 2.0;
    std::vector<int32_t> row_off, col_off, row_ids, col_ids, transpose;
    DBuf d_rows, d_cols, d_Z, d_mv;
    std::vector<size_t> z_off;  // element offset of block b in d_Z
    int nmv = 0;                // entries of d_mv (block products incl. transposes)
};

namespace {

// Fills the dense blocks (one wave per 64 rows of a block) on the context stream.
int near_fill_blocks(farcur_ctx* ctx, const std::vector<NearJob>& jobs, double near_factor) {
    std::vector<int32_t> tj, tr;
    for (size_t j = 0; j < jobs.size(); ++j)
        for (int t = 0; t < jobs[j].m; t += 64) {
            tj.push_back((int32_t)j);
            tr.push_back(t);
        }
    DBuf dj, dt;
    TRY(dj.ensure(jobs.size() * sizeof(NearJob)));
    TRY(dt.ensure(tj.size() * 2 * sizeof(int32_t)));
    hipStream_t st = ctx->stream;
    HIPTRY(hipMemcpyAsync(dj.p, jobs.data(), jobs.size() * sizeof(NearJob), hipMemcpyHostToDevice, st));
    HIPTRY(hipMemcpyAsync(dt.p, tj.data(), tj.size() * 4, hipMemcpyHostToDevice, st));
    HIPTRY(hipMemcpyAsync(dt.as<int32_t>() + tj.size(), tr.data(), tr.size() * 4, hipMemcpyHostToDevice, st));
    launch_near_fill(dj.as<NearJob>(), (int)tj.size(), dt.as<int32_t>(), dt.as<int32_t>() + tj.size(),
                     ctx->rec64.as<double>(), ctx->k, ctx->eta, near_factor, st);
    HIPTRY(hipGetLastError());
    HIPTRY(hipStreamSynchronize(st));  // dj, dt are freed on return
    return FARCUR_OK;
}

}  // namespace

extern "C" int farcur_fill_near(farcur_ctx* ctx, const int* rows, int nr, const int* cols, int nc, double near_factor,
                                farcur_c128* Z) {
    if (!ctx || !rows || !cols || !Z) return fail(FARCUR_EINVAL, "NULL argument");
    if (ctx->ne == 0) return fail(FARCUR_ESTATE, "no geometry");
    if (nr <= 0 || nc <= 0) return fail(FARCUR_EINVAL, "empty panel");
    if (!(near_factor >= 0)) return fail(FARCUR_EINVAL, "near_factor must be >= 0");
    TRY(check_ids(rows, nr, ctx->ne, "rows"));
    TRY(check_ids(cols, nc, ctx->ne, "cols"));
    TRY(ctx->select());
    DBuf dr, dc, dz;
    TRY(dr.ensure((size_t)nr * 4));
    TRY(dc.ensure((size_t)nc * 4));
    TRY(dz.ensure((size_t)nr * nc * sizeof(c128)));
    HIPTRY(hipMemcpyAsync(dr.p, rows, (size_t)nr * 4, hipMemcpyHostToDevice, ctx->stream));
    HIPTRY(hipMemcpyAsync(dc.p, cols, (size_t)nc * 4, hipMemcpyHostToDevice, ctx->stream));
    std::vector<NearJob> jobs{NearJob{dr.as<int32_t>(), dc.as<int32_t>(), dz.as<c128>(), nr, nc}};
    TRY(near_fill_blocks(ctx, jobs, near_factor));
    HIPTRY(hipMemcpy(Z, dz.p, (size_t)nr * nc * sizeof(c128), hipMemcpyDeviceToHost));
    return FARCUR_OK;
}

extern "C" int farcur_near_create(farcur_ctx* ctx, int nblocks, const int* row_off, const int* row_ids,
                                  const int* col_off, const int* col_ids, const int* transpose, double near_factor,
                                  farcur_near** out) {
    if (!ctx || !row_off || !col_off || !row_ids || !col_ids || !out) return fail(FARCUR_EINVAL, "NULL argument");
    *out = nullptr;
    if (nblocks <= 0) return fail(FARCUR_EINVAL, "nblocks must be > 0");
    if (ctx->ne == 0) return fail(FARCUR_ESTATE, "no geometry");
    if (!(near_factor >= 0)) return fail(FARCUR_EINVAL, "near_factor must be >= 0");
    if (row_off[0] != 0 || col_off[0] != 0) return fail(FARCUR_EINVAL, "offsets must start at 0");
    for (int b = 0; b < nblocks; ++b)
        if (row_off[b + 1] <= row_off[b] || col_off[b + 1] <= col_off[b])
            return fail(FARCUR_EINVAL, "block %d is empty or offsets decrease", b);
    const int nr = row_off[nblocks], nc = col_off[nblocks];
    TRY(check_ids(row_ids, nr, ctx->ne, "row_ids"));
    TRY(check_ids(col_ids, nc, ctx->ne, "col_ids"));
    TRY(ctx->select());
    auto Nr = std::make_unique<farcur_near>();
    Nr->ctx = ctx;
    Nr->nb = nblocks;
    Nr->near_factor = near_factor;
    Nr->row_off.assign(row_off, row_off + nblocks + 1);
    Nr->col_off.assign(col_off, col_off + nblocks + 1);
    Nr->row_ids.assign(row_ids, row_ids + nr);
    Nr->col_ids.assign(col_ids, col_ids + nc);
    Nr->transpose.assign(nblocks, 0);
    size_t total = 0;
    Nr->z_off.resize(nblocks);
    for (int b = 0; b < nblocks; ++b) {
        if (transpose) Nr->transpose[b] = transpose[b] != 0;
        Nr->z_off[b] = total;
        total += (size_t)(row_off[b + 1] - row_off[b]) * (col_off[b + 1] - col_off[b]);
    }
    TRY(Nr->d_rows.ensure((size_t)nr * 4));
    TRY(Nr->d_cols.ensure((size_t)nc * 4));
    TRY(Nr->d_Z.ensure(total * sizeof(c128)));
    HIPTRY(hipMemcpyAsync(Nr->d_rows.p, row_ids, (size_t)nr * 4, hipMemcpyHostToDevice, ctx->stream));
    HIPTRY(hipMemcpyAsync(Nr->d_cols.p, col_ids, (size_t)nc * 4, hipMemcpyHostToDevice, ctx->stream));
    std::vector<NearJob> jobs(nblocks);
    std::vector<NearMv> mv;
    for (int b = 0; b < nblocks; ++b) {
        jobs[b] = NearJob{Nr->d_rows.as<int32_t>() + row_off[b], Nr->d_cols.as<int32_t>() + col_off[b],
                          Nr->d_Z.as<c128>() + Nr->z_off[b], row_off[b + 1] - row_off[b], col_off[b + 1] - col_off[b]};
        mv.push_back(NearMv{jobs[b], 0, 0});
        if (Nr->transpose[b]) mv.push_back(NearMv{jobs[b], 1, 0});
    }
    TRY(near_fill_blocks(ctx, jobs, near_factor));
    TRY(Nr->d_mv.ensure(mv.size() * sizeof(NearMv)));
    HIPTRY(hipMemcpy(Nr->d_mv.p, mv.data(), mv.size() * sizeof(NearMv), hipMemcpyHostToDevice));
    Nr->nmv = (int)mv.size();
    *out = Nr.release();
    return FARCUR_OK;
}

extern "C" int farcur_near_block(farcur_near* Nr, int b, farcur_c128* Z) {
    if (!Nr || !Z) return fail(FARCUR_EINVAL, "NULL argument");
    if (b < 0 || b >= Nr->nb) return fail(FARCUR_EINVAL, "block %d out of range", b);
    TRY(Nr->ctx->select());
    const size_t cnt = (size_t)(Nr->row_off[b + 1] - Nr->row_off[b]) * (Nr->col_off[b + 1] - Nr->col_off[b]);
    HIPTRY(hipMemcpy(Z, Nr->d_Z.as<c128>() + Nr->z_off[b], cnt * sizeof(c128), hipMemcpyDeviceToHost));
    return FARCUR_OK;
}

extern "C" void farcur_near_destroy(farcur_near* Nr) {
    if (!Nr) return;
    hipSetDevice(Nr->ctx->device);
    hipStreamSynchronize(Nr->ctx->stream);
    Nr->d_rows.release();
    Nr->d_cols.release();
    Nr->d_Z.release();
    Nr->d_mv.release();
    delete Nr;
}

struct farcur_operator {
    farcur_ctx* ctx = nullptr;
    farcur_near* near = nullptr;
    farcur_batch* far = nullptr;
    int far_symmetric = 1;
    int N = 0;
    DBuf d_x, d_y, d_x32, d_y32, d_far;
    int nfar = 0, far_rmax = 1;
    std::vector<c128> h_y;
    double ms_near = 0, ms_far = 0;
    int matvecs = 0;
};

extern "C" int farcur_operator_create(farcur_near* near, farcur_batch* far, int far_symmetric, farcur_operator** out) {
    if (!out || (!near && !far)) return fail(FARCUR_EINVAL, "need a near part, a far part or both");
    *out = nullptr;
    farcur_ctx* ctx = near ? near->ctx : far->ctx;
    if (near && far && near->ctx != far->ctx) return fail(FARCUR_EINVAL, "near and far parts on different contexts");
    if (far && !far->has_result) return fail(FARCUR_ESTATE, "far batch has no compress result");
    TRY(ctx->select());
    auto op = std::make_unique<farcur_operator>();
    op->ctx = ctx;
    op->near = near;
    op->far = far;
    op->far_symmetric = far_symmetric ? 1 : 0;
    op->N = ctx->ne;
    TRY(op->d_x.ensure((size_t)op->N * sizeof(c128)));
    TRY(op->d_y.ensure((size_t)op->N * sizeof(c128)));
    if (far) {
        double ms_interp = 0;
        TRY(batch_interp(far, &ms_interp));
        std::vector<Mv1Job> jobs;
        for (int b = 0; b < far->nb; ++b) {
            const BState& S = far->st[b];
            if (S.fin_r > 4096) return fail(FARCUR_EINVAL, "far block %d: r = %d above 4096", b, S.fin_r);
            op->far_rmax = std::max(op->far_rmax, S.fin_r);
            const int32_t* rows = far->d_rows.as<int32_t>() + far->row_off[b];
            const int32_t* cols = far->d_cols.as<int32_t>() + far->col_off[b];
            jobs.push_back(Mv1Job{S.fin.R, far->interp[b], cols, rows, S.fin_r, S.n, S.m, 0});
            if (op->far_symmetric) jobs.push_back(Mv1Job{far->interp[b], S.fin.R, rows, cols, S.fin_r, S.m, S.n, 0});
        }
        TRY(op->d_far.ensure(jobs.size() * sizeof(Mv1Job)));
        HIPTRY(hipMemcpy(op->d_far.p, jobs.data(), jobs.size() * sizeof(Mv1Job), hipMemcpyHostToDevice));
        op->nfar = (int)jobs.size();
        TRY(op->d_x32.ensure((size_t)op->N * sizeof(c64)));
        TRY(op->d_y32.ensure((size_t)op->N * sizeof(c64)));
    }
    op->h_y.resize(op->N);
    *out = op.release();
    return FARCUR_OK;
}

extern "C" void farcur_operator_destroy(farcur_operator* op) {
    if (!op) return;
    hipSetDevice(op->ctx->device);
    hipStreamSynchronize(op->ctx->stream);
    DBuf* bufs[] = {&op->d_x, &op->d_y, &op->d_x32, &op->d_y32, &op->d_far};
    for (DBuf* b : bufs) b->release();
    delete op;
}

namespace {

// y = (near + far) x on the device for one vector (host in, host out).
int operator_apply(farcur_operator* op, const c128* x, c128* y) {
    farcur_ctx* ctx = op->ctx;
    hipStream_t st = ctx->stream;
    const int N = op->N;
    HIPTRY(hipMemcpyAsync(op->d_x.p, x, (size_t)N * sizeof(c128), hipMemcpyHostToDevice, st));
    HIPTRY(hipMemsetAsync(op->d_y.p, 0, (size_t)N * sizeof(c128), st));
    HIPTRY(hipEventRecord(ctx->ev0, st));
    if (op->near) launch_near_matvec(op->near->d_mv.as<NearMv>(), op->near->nmv, op->d_x.as<c128>(), op->d_y.as<c128>(), st);
    HIPTRY(hipEventRecord(ctx->ev1, st));
    if (op->far) {
        launch_to_c64(op->d_x.as<c128>(), op->d_x32.as<c64>(), N, st);
        HIPTRY(hipMemsetAsync(op->d_y32.p, 0, (size_t)N * sizeof(c64), st));
        launch_matvec1(op->d_far.as<Mv1Job>(), op->nfar, op->far_rmax, op->far->f32, op->d_x32.as<float2>(),
                       op->d_y32.as<float2>(), st);
        launch_add_c64(op->d_y.as<c128>(), op->d_y32.as<c64>(), N, st);
    }
    HIPTRY(hipGetLastError());
    HIPTRY(hipMemcpyAsync(y, op->d_y.p, (size_t)N * sizeof(c128), hipMemcpyDeviceToHost, st));
    HIPTRY(hipStreamSynchronize(st));
    float t = 0;
    hipEventElapsedTime(&t, ctx->ev0, ctx->ev1);
    op->ms_near += t;
    op->matvecs += 1;
    return FARCUR_OK;
}

}  // namespace

extern "C" int farcur_operator_matvec(farcur_operator* op, const farcur_c128* x, farcur_c128* y) {
    if (!op || !x || !y) return fail(FARCUR_EINVAL, "NULL argument");
    farcur_ctx* ctx = op->ctx;
    TRY(ctx->select());
    return operator_apply(op, (const c128*)x, (c128*)y);
}

// Restarted GMRES (Saad and Schultz 1986) with modified Gram-Schmidt and
// Givens rotations, zero initial guess, per right-hand side; vectors on the
// host, every product on the device.
extern "C" int farcur_operator_solve(farcur_operator* op, const farcur_c128* b, farcur_c128* x, int nrhs, double tol,
                                     int restart, int maxit, int* iters, double* relres) {
    if (!op || !b || !x) return fail(FARCUR_EINVAL, "NULL argument");
    if (nrhs <= 0 || restart <= 0 || maxit <= 0 || !(tol > 0)) return fail(FARCUR_EINVAL, "bad solver parameters");
    farcur_ctx* ctx = op->ctx;
    TRY(ctx->select());
    typedef std::complex<double> cd;
    const int N = op->N;
    std::vector<cd> Vb((size_t)N * (restart + 1)), w(N), H((size_t)(restart + 1) * restart), g(restart + 1),
        cs(restart), sn(restart), xv(N);
    for (int q = 0; q < nrhs; ++q) {
        const cd* bq = (const cd*)(b + (size_t)q * N);
        std::fill(xv.begin(), xv.end(), cd(0));
        double bn = 0;
        for (int i = 0; i < N; ++i) bn += std::norm(bq[i]);
        bn = std::sqrt(bn);
        int it = 0;
        double res = bn > 0 ? 1.0 : 0.0;
        while (bn > 0 && res > tol && it < maxit) {
            // r = b - A x
            TRY(operator_apply(op, (const c128*)xv.data(), (c128*)w.data()));
            double rn = 0;
            for (int i = 0; i < N; ++i) {
                w[i] = bq[i] - w[i];
                rn += std::norm(w[i]);
            }
            rn = std::sqrt(rn);
            res = rn / bn;
            if (res <= tol) break;
            for (int i = 0; i < N; ++i) Vb[i] = w[i] / rn;
            std::fill(g.begin(), g.end(), cd(0));
            g[0] = rn;
            int j = 0;
            for (; j < restart && it < maxit; ++j, ++it) {
                cd* vj = &Vb[(size_t)j * N];
                TRY(operator_apply(op, (const c128*)vj, (c128*)w.data()));
                for (int i = 0; i <= j; ++i) {  // modified Gram-Schmidt
                    const cd* vi = &Vb[(size_t)i * N];
                    cd h = 0;
                    for (int t = 0; t < N; ++t) h += std::conj(vi[t]) * w[t];
                    H[(size_t)j * (restart + 1) + i] = h;
                    for (int t = 0; t < N; ++t) w[t] -= h * vi[t];
                }
                double wn = 0;
                for (int t = 0; t < N; ++t) wn += std::norm(w[t]);
                wn = std::sqrt(wn);
                H[(size_t)j * (restart + 1) + j + 1] = wn;
                cd* vn = &Vb[(size_t)(j + 1) * N];
                for (int t = 0; t < N; ++t) vn[t] = wn > 0 ? w[t] / wn : cd(0);
                cd* hj = &H[(size_t)j * (restart + 1)];
                for (int i = 0; i < j; ++i) {  // previous rotations
                    const cd t0 = cs[i] * hj[i] + sn[i] * hj[i + 1];
                    hj[i + 1] = -std::conj(sn[i]) * hj[i] + cs[i] * hj[i + 1];
                    hj[i] = t0;
                }
                const double a = std::abs(hj[j]), bb = std::abs(hj[j + 1]);
                const double den = std::sqrt(a * a + bb * bb);
                if (den == 0) {
                    cs[j] = 1;
                    sn[j] = 0;
                } else {
                    cs[j] = a / den;
                    const cd ph = a > 0 ? hj[j] / a : cd(1);
                    sn[j] = ph * std::conj(hj[j + 1]) / den;
                }
                hj[j] = cs[j] * hj[j] + sn[j] * hj[j + 1];
                hj[j + 1] = 0;
                g[j + 1] = -std::conj(sn[j]) * g[j];
                g[j] = cs[j] * g[j];
                res = std::abs(g[j + 1]) / bn;
                if (res <= tol || wn == 0) {
                    ++j;
                    ++it;
                    break;
                }
            }
            // x += V y, H y = g (upper triangular j x j)
            std::vector<cd> yv(j);
            for (int i = j - 1; i >= 0; --i) {
                cd s = g[i];
                for (int l = i + 1; l < j; ++l) s -= H[(size_t)l * (restart + 1) + i] * yv[l];
                yv[i] = s / H[(size_t)i * (restart + 1) + i];
            }
            for (int i = 0; i < j; ++i) {
                const cd* vi = &Vb[(size_t)i * N];
                for (int t = 0; t < N; ++t) xv[t] += yv[i] * vi[t];
            }
        }
        // true residual of the returned solution
        if (bn > 0) {
            TRY(operator_apply(op, (const c128*)xv.data(), (c128*)w.data()));
            double rn = 0;
            for (int i = 0; i < N; ++i) rn += std::norm(bq[i] - w[i]);
            res = std::sqrt(rn) / bn;
        }
        std::copy(xv.begin(), xv.end(), (cd*)(x + (size_t)q * N));
        if (iters) iters[q] = it;
        if (relres) relres[q] = res;
    }
    return FARCUR_OK;
}

// ------------------------------------------------------------------ excitations
// RWG-tested plane waves with the centroid rule of user_compute_Ei.m
// (codes/GPU_implementation/seed/user_compute_Ei.m:40-57): propagation
// k_i = -(sin th cos ph, sin th sin ph, cos th), polarisation
// e_i = cos(rot) (-sin ph, cos ph, 0) + sin(rot) (cos th cos ph, cos th sin ph, -sin th),
// V_m = l_m / 2 [E(c+) . (c+ - v+) + E(c-) . (v- - c-)], E(r) = e_i exp(-j k k_i . r).
extern "C" int farcur_mesh_plane_waves(const farcur_mesh* mesh, double k, const double* theta, const double* phi,
                                       int nrhs, double rot_eh, farcur_c128* V) {
    if (!mesh || !theta || !phi || !V) return fail(FARCUR_EINVAL, "NULL argument");
    if (!mesh->rwg) return fail(FARCUR_ESTATE, "RWG tables not built");
    if (nrhs <= 0 || !(k > 0)) return fail(FARCUR_EINVAL, "bad nrhs or k");
    const Mesh& M = mesh->m;
    const int ne = M.ne();
    for (int s = 0; s < nrhs; ++s) {
        const double th = theta[s], ph = phi[s];
        const double ki[3] = {-std::sin(th) * std::cos(ph), -std::sin(th) * std::sin(ph), -std::cos(th)};
        const double ei[3] = {std::cos(rot_eh) * -std::sin(ph) + std::sin(rot_eh) * std::cos(th) * std::cos(ph),
                              std::cos(rot_eh) * std::cos(ph) + std::sin(rot_eh) * std::cos(th) * std::sin(ph),
                              std::sin(rot_eh) * -std::sin(th)};
        for (int e = 0; e < ne; ++e) {
            double re = 0, im = 0;
            for (int h = 0; h < 2; ++h) {
                const int t = M.edges[4 * e + h];
                const double* c = &M.cent[3 * t];
                const double* v = &M.vertex[3 * M.edges[4 * e + 2 + h]];
                double rho[3];
                for (int x = 0; x < 3; ++x) rho[x] = h == 0 ? c[x] - v[x] : v[x] - c[x];
                const double ph0 = -k * (ki[0] * c[0] + ki[1] * c[1] + ki[2] * c[2]);
                const double dp = ei[0] * rho[0] + ei[1] * rho[1] + ei[2] * rho[2];
                re += dp * std::cos(ph0);
                im += dp * std::sin(ph0);
            }
            V[(size_t)s * ne + e] = farcur_c128{0.5 * M.ln[e] * re, 0.5 * M.ln[e] * im};
        }
    }
    return FARCUR_OK;
}
