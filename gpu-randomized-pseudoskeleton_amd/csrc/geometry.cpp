// Host-side mesh generators, RWG tables (get_edge.m semantics) and the octree
// block partition. None of this is on the GPU: it runs once per geometry.
#include "geometry.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <numeric>
#include <unordered_map>

namespace farcur {

namespace {

inline void sub3(const double* a, const double* b, double* c) {
    c[0] = a[0] - b[0];
    c[1] = a[1] - b[1];
    c[2] = a[2] - b[2];
}
inline void cross3(const double* a, const double* b, double* c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// splitmix64: seeds the geometry jitter (host only).
struct SplitMix {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

struct Key3 {
    uint64_t a, b, c;
    bool operator==(const Key3& o) const { return a == o.a && b == o.b && c == o.c; }
};
struct Key3Hash {
    size_t operator()(const Key3& k) const {
        uint64_t h = k.a * 0x9E3779B97F4A7C15ull;
        h ^= k.b + 0x7F4A7C159E3779B9ull + (h << 6) + (h >> 2);
        h ^= k.c + 0x94D049BB133111EBull + (h << 6) + (h >> 2);
        return (size_t)h;
    }
};
inline Key3 bits3(const double* p) {
    Key3 k;
    std::memcpy(&k.a, &p[0], 8);
    std::memcpy(&k.b, &p[1], 8);
    std::memcpy(&k.c, &p[2], 8);
    return k;
}

}  // namespace

void mesh_append(Mesh& dst, const Mesh& src, double dx, double dy, double dz) {
    const int32_t off = dst.nv();
    for (int i = 0; i < src.nv(); ++i) {
        dst.vertex.push_back(src.vertex[3 * i + 0] + dx);
        dst.vertex.push_back(src.vertex[3 * i + 1] + dy);
        dst.vertex.push_back(src.vertex[3 * i + 2] + dz);
    }
    for (int32_t v : src.topol) dst.topol.push_back(v + off);
    dst.edges.clear();
    dst.trian.clear();
    dst.ln.clear();
    dst.cent.clear();
    dst.un.clear();
    dst.ds.clear();
}

Mesh mesh_plate(double lx, double ly, int nx, int ny, double x0, double y0, double z0) {
    Mesh m;
    // Vertex (ix, iy) has id ix*(ny+1) + iy, as the column-major meshgrid of
    // sq_plate.m numbers them.
    for (int ix = 0; ix <= nx; ++ix) {
        for (int iy = 0; iy <= ny; ++iy) {
            m.vertex.push_back(x0 - lx / 2 + lx * (double)ix / nx);
            m.vertex.push_back(y0 - ly / 2 + ly * (double)iy / ny);
            m.vertex.push_back(z0);
        }
    }
    auto V = [&](int ix, int iy) { return ix * (ny + 1) + iy; };
    for (int ix = 0; ix < nx; ++ix) {
        for (int iy = 0; iy < ny; ++iy) {
            const int t1[3] = {V(ix, iy), V(ix, iy + 1), V(ix + 1, iy)};
            const int t2[3] = {V(ix, iy + 1), V(ix + 1, iy + 1), V(ix + 1, iy)};
            m.topol.insert(m.topol.end(), t1, t1 + 3);
            m.topol.insert(m.topol.end(), t2, t2 + 3);
        }
    }
    return m;
}

Mesh mesh_octasphere(double R, int n_levels) {
    Mesh m;
    const double v0[18] = {0, 0, 1, 1, 0, 0, 0, 1, 0, -1, 0, 0, 0, -1, 0, 0, 0, -1};
    for (int i = 0; i < 18; ++i) m.vertex.push_back(R * v0[i]);
    const int t0[24] = {1, 3, 2, 1, 4, 3, 1, 5, 4, 1, 2, 5, 6, 2, 3, 6, 3, 4, 6, 4, 5, 6, 5, 2};
    for (int i = 0; i < 24; ++i) m.topol.push_back(t0[i] - 1);
    std::unordered_map<Key3, int32_t, Key3Hash> lookup;
    for (int i = 0; i < m.nv(); ++i) lookup.emplace(bits3(&m.vertex[3 * i]), i);
    auto midpoint = [&](int a, int b) -> int32_t {
        double r[3];
        for (int c = 0; c < 3; ++c) r[c] = (m.vertex[3 * a + c] + m.vertex[3 * b + c]) / 2;
        const double nr = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        for (int c = 0; c < 3; ++c) r[c] = r[c] * R / nr;
        auto it = lookup.find(bits3(r));
        if (it != lookup.end()) return it->second;
        const int32_t id = m.nv();
        m.vertex.insert(m.vertex.end(), r, r + 3);
        lookup.emplace(bits3(r), id);
        return id;
    };
    for (int it = 0; it < n_levels; ++it) {
        const int nt = m.nt();
        for (int t = 0; t < nt; ++t) {
            const int v1 = m.topol[3 * t], v2 = m.topol[3 * t + 1], v3 = m.topol[3 * t + 2];
            const int v12 = midpoint(v1, v2);
            const int v23 = midpoint(v2, v3);
            const int v31 = midpoint(v3, v1);
            m.topol[3 * t] = v12;
            m.topol[3 * t + 1] = v23;
            m.topol[3 * t + 2] = v31;
            const int add[9] = {v2, v23, v12, v23, v3, v31, v31, v1, v12};
            m.topol.insert(m.topol.end(), add, add + 9);
        }
    }
    return m;
}

Mesh mesh_icosphere(double ax, double ay, double az, int level, double jitter, uint64_t seed) {
    Mesh m;
    const double p = (1.0 + std::sqrt(5.0)) / 2.0;
    const double base[12][3] = {{-1, p, 0}, {1, p, 0}, {-1, -p, 0}, {1, -p, 0}, {0, -1, p}, {0, 1, p},
                                {0, -1, -p}, {0, 1, -p}, {p, 0, -1}, {p, 0, 1}, {-p, 0, -1}, {-p, 0, 1}};
    for (auto& v : base) {
        const double n = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        m.vertex.push_back(v[0] / n);
        m.vertex.push_back(v[1] / n);
        m.vertex.push_back(v[2] / n);
    }
    const int faces[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                              {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
                              {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    for (auto& f : faces) {
        // Orient so that cross(v3 - v1, v2 - v1) points outward (the get_edge.m
        // normal convention).
        const double* a = &m.vertex[3 * f[0]];
        const double* b = &m.vertex[3 * f[1]];
        const double* c = &m.vertex[3 * f[2]];
        double e1[3], e2[3], n[3], cen[3];
        sub3(c, a, e1);
        sub3(b, a, e2);
        cross3(e1, e2, n);
        for (int k = 0; k < 3; ++k) cen[k] = a[k] + b[k] + c[k];
        if (dot3(n, cen) > 0) {
            m.topol.push_back(f[0]);
            m.topol.push_back(f[1]);
            m.topol.push_back(f[2]);
        } else {
            m.topol.push_back(f[0]);
            m.topol.push_back(f[2]);
            m.topol.push_back(f[1]);
        }
    }
    for (int it = 0; it < level; ++it) {
        std::unordered_map<uint64_t, int32_t> cache;
        auto mid = [&](int a, int b) -> int32_t {
            const uint64_t key = ((uint64_t)std::min(a, b) << 32) | (uint32_t)std::max(a, b);
            auto f = cache.find(key);
            if (f != cache.end()) return f->second;
            double r[3];
            for (int c = 0; c < 3; ++c) r[c] = 0.5 * (m.vertex[3 * a + c] + m.vertex[3 * b + c]);
            const double n = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
            for (int c = 0; c < 3; ++c) r[c] /= n;
            const int32_t id = m.nv();
            m.vertex.insert(m.vertex.end(), r, r + 3);
            cache.emplace(key, id);
            return id;
        };
        std::vector<int32_t> nt;
        nt.reserve(m.topol.size() * 4);
        for (int t = 0; t < m.nt(); ++t) {
            const int a = m.topol[3 * t], b = m.topol[3 * t + 1], c = m.topol[3 * t + 2];
            const int ab = mid(a, b), bc = mid(b, c), ca = mid(c, a);
            const int add[12] = {a, ab, ca, ab, b, bc, ca, bc, c, ab, bc, ca};
            nt.insert(nt.end(), add, add + 12);
        }
        m.topol.swap(nt);
    }
    // Scale to the ellipsoid axes.
    for (int i = 0; i < m.nv(); ++i) {
        m.vertex[3 * i + 0] *= ax;
        m.vertex[3 * i + 1] *= ay;
        m.vertex[3 * i + 2] *= az;
    }
    if (jitter > 0) {
        // Mean edge length of the unjittered mesh (each edge counted from both
        // sides of a closed surface, which leaves the mean unchanged).
        double sum = 0;
        for (int t = 0; t < m.nt(); ++t) {
            for (int k = 0; k < 3; ++k) {
                double d[3];
                sub3(&m.vertex[3 * m.topol[3 * t + k]], &m.vertex[3 * m.topol[3 * t + (k + 1) % 3]], d);
                sum += std::sqrt(dot3(d, d));
            }
        }
        const double h = sum / (3.0 * m.nt());
        SplitMix rng{seed};
        for (int i = 0; i < m.nv(); ++i) {
            double* v = &m.vertex[3 * i];
            const double r = std::sqrt(dot3(v, v));
            const double delta = (2.0 * rng.uniform() - 1.0) * jitter * h;
            const double s = (r + delta) / r;
            v[0] *= s;
            v[1] *= s;
            v[2] *= s;
        }
    }
    return m;
}

bool build_rwg(Mesh& m, std::string& err) {
    const int nt = m.nt();
    const int n3 = 3 * nt;
    // Entry k: position block p = k / nt (edges v1v2, v2v3, v3v1), triangle t = k % nt.
    std::vector<int32_t> ea(n3), eb(n3), vv(n3), ov(n3), tt(n3);
    for (int p = 0; p < 3; ++p) {
        for (int t = 0; t < nt; ++t) {
            const int k = p * nt + t;
            const int v1 = m.topol[3 * t], v2 = m.topol[3 * t + 1], v3 = m.topol[3 * t + 2];
            if (p == 0) { ea[k] = v1; eb[k] = v2; vv[k] = v3; ov[k] = 2; }
            if (p == 1) { ea[k] = v2; eb[k] = v3; vv[k] = v1; ov[k] = 0; }
            if (p == 2) { ea[k] = v3; eb[k] = v1; vv[k] = v2; ov[k] = 1; }
            tt[k] = t;
        }
    }
    std::vector<int32_t> order(n3);
    std::iota(order.begin(), order.end(), 0);
    auto lo = [&](int k) { return std::min(ea[k], eb[k]); };
    auto hi = [&](int k) { return std::max(ea[k], eb[k]); };
    // Stable lexicographic sort of (min, max) pairs, as Matlab's sortrows.
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        if (lo(x) != lo(y)) return lo(x) < lo(y);
        return hi(x) < hi(y);
    });
    std::vector<int32_t> first, second;  // sorted positions: predecessor-equal, successor-equal
    for (int i = 0; i < n3; ++i) {
        const bool eq_prev = i > 0 && lo(order[i]) == lo(order[i - 1]) && hi(order[i]) == hi(order[i - 1]);
        const bool eq_next = i + 1 < n3 && lo(order[i]) == lo(order[i + 1]) && hi(order[i]) == hi(order[i + 1]);
        if (eq_prev && eq_next) {
            err = "non-manifold edge (junction) is not supported";
            return false;
        }
        if (eq_prev) first.push_back(order[i]);
        if (eq_next) second.push_back(order[i]);
    }
    const int ne = (int)first.size();
    m.edges.assign(4 * (size_t)ne, 0);
    m.ln.assign(ne, 0.0);
    m.trian.assign(3 * (size_t)nt, 0);
    for (int e = 0; e < ne; ++e) {
        const int kp = first[e], km = second[e];
        m.edges[4 * e + 0] = tt[kp];
        m.edges[4 * e + 1] = tt[km];
        m.edges[4 * e + 2] = vv[kp];
        m.edges[4 * e + 3] = vv[km];
        double d[3];
        sub3(&m.vertex[3 * ea[kp]], &m.vertex[3 * eb[kp]], d);
        m.ln[e] = std::sqrt(dot3(d, d));
        m.trian[3 * tt[kp] + ov[kp]] += e + 1;
        m.trian[3 * tt[km] + ov[km]] -= e + 1;
    }
    m.cent.assign(3 * (size_t)nt, 0.0);
    m.un.assign(3 * (size_t)nt, 0.0);
    m.ds.assign(nt, 0.0);
    for (int t = 0; t < nt; ++t) {
        const double* v1 = &m.vertex[3 * m.topol[3 * t]];
        const double* v2 = &m.vertex[3 * m.topol[3 * t + 1]];
        const double* v3 = &m.vertex[3 * m.topol[3 * t + 2]];
        double a[3], b[3], c[3];
        for (int k = 0; k < 3; ++k) m.cent[3 * t + k] = (v1[k] + v2[k] + v3[k]) / 3;
        sub3(v3, v1, a);
        sub3(v2, v1, b);
        cross3(a, b, c);
        const double nc = std::sqrt(dot3(c, c));
        for (int k = 0; k < 3; ++k) m.un[3 * t + k] = c[k] / nc;
        m.ds[t] = nc / 2;
    }
    return true;
}

BlockPartition build_partition(const Mesh& m, double leaf_size, double eta_adm) {
    BlockPartition P;
    P.leaf_size = leaf_size;
    P.eta_adm = eta_adm;
    P.nrwg_mesh = m.ne();
    const int ne = m.ne();
    std::vector<double> pt(3 * (size_t)ne);
    double bmin[3] = {1e300, 1e300, 1e300};
    for (int e = 0; e < ne; ++e) {
        // The shared edge: the vertices of T+ other than its free vertex.
        const int tp = m.edges[4 * e], vp = m.edges[4 * e + 2];
        int a = -1, b = -1;
        for (int k = 0; k < 3; ++k) {
            const int v = m.topol[3 * tp + k];
            if (v == vp) continue;
            if (a < 0) a = v; else b = v;
        }
        for (int k = 0; k < 3; ++k) {
            pt[3 * e + k] = 0.5 * (m.vertex[3 * a + k] + m.vertex[3 * b + k]);
            bmin[k] = std::min(bmin[k], pt[3 * e + k]);
        }
    }
    std::map<std::tuple<int64_t, int64_t, int64_t>, std::vector<int32_t>> cells;
    for (int e = 0; e < ne; ++e) {
        const int64_t ix = (int64_t)std::floor((pt[3 * e + 0] - bmin[0]) / leaf_size);
        const int64_t iy = (int64_t)std::floor((pt[3 * e + 1] - bmin[1]) / leaf_size);
        const int64_t iz = (int64_t)std::floor((pt[3 * e + 2] - bmin[2]) / leaf_size);
        cells[std::make_tuple(ix, iy, iz)].push_back(e);
    }
    P.group_offsets.push_back(0);
    for (auto& kv : cells) {
        const auto& ids = kv.second;
        P.group_rwg.insert(P.group_rwg.end(), ids.begin(), ids.end());
        P.group_offsets.push_back((int32_t)P.group_rwg.size());
        double c[3] = {0, 0, 0};
        for (int e : ids)
            for (int k = 0; k < 3; ++k) c[k] += pt[3 * e + k];
        for (int k = 0; k < 3; ++k) c[k] /= (double)ids.size();
        double rmax = 0;
        for (int e : ids) {
            for (int s = 0; s < 2; ++s) {
                const int t = m.edges[4 * e + s];
                for (int k = 0; k < 3; ++k) {
                    double d[3];
                    sub3(&m.vertex[3 * m.topol[3 * t + k]], c, d);
                    rmax = std::max(rmax, std::sqrt(dot3(d, d)));
                }
            }
        }
        P.group_centre.insert(P.group_centre.end(), c, c + 3);
        P.group_diam.push_back(2 * rmax);
    }
    const int ng = (int)P.group_diam.size();
    for (int a = 0; a < ng; ++a) {
        P.near_pairs.push_back(a);
        P.near_pairs.push_back(a);
        for (int b = a + 1; b < ng; ++b) {
            double d[3];
            sub3(&P.group_centre[3 * a], &P.group_centre[3 * b], d);
            const double dist = std::sqrt(dot3(d, d));
            const double dm = std::max(P.group_diam[a], P.group_diam[b]);
            auto& dst = dist >= eta_adm * dm ? P.far_pairs : P.near_pairs;
            dst.push_back(a);
            dst.push_back(b);
        }
    }
    return P;
}

namespace {

constexpr char kBlMagic[8] = {'F', 'A', 'R', 'C', 'U', 'R', 'B', 'L'};
constexpr uint32_t kBlVersion = 1;

struct BlEntry {
    int32_t a, b;
    uint8_t comp, self;
    uint16_t pad;
};
static_assert(sizeof(BlEntry) == 12, "block-list entry layout");

uint64_t fnv1a64(uint64_t h, const void* data, size_t n) {
    const unsigned char* p = (const unsigned char*)data;
    for (size_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= 0x100000001B3ull;
    }
    return h;
}

}  // namespace

bool save_partition(const BlockPartition& P, const std::string& path, std::string& err) {
    std::vector<unsigned char> buf;
    auto put = [&](const void* p, size_t n) {
        const unsigned char* c = (const unsigned char*)p;
        buf.insert(buf.end(), c, c + n);
    };
    const uint32_t ng = (uint32_t)P.group_diam.size(), nrwg = (uint32_t)P.group_rwg.size();
    const uint32_t nfar = (uint32_t)(P.far_pairs.size() / 2), nnear = (uint32_t)(P.near_pairs.size() / 2);
    const uint32_t hdr[6] = {kBlVersion, ng, (uint32_t)P.nrwg_mesh, nrwg, nfar, nnear};
    put(kBlMagic, 8);
    put(hdr, sizeof hdr);
    put(&P.leaf_size, 8);
    put(&P.eta_adm, 8);
    put(P.group_offsets.data(), P.group_offsets.size() * 4);
    put(P.group_rwg.data(), P.group_rwg.size() * 4);
    put(P.group_centre.data(), P.group_centre.size() * 8);
    put(P.group_diam.data(), P.group_diam.size() * 8);
    for (uint32_t i = 0; i < nfar; ++i) {
        const BlEntry e{P.far_pairs[2 * i], P.far_pairs[2 * i + 1], 1, 0, 0};
        put(&e, sizeof e);
    }
    for (uint32_t i = 0; i < nnear; ++i) {
        const int32_t a = P.near_pairs[2 * i], b = P.near_pairs[2 * i + 1];
        const BlEntry e{a, b, 0, (uint8_t)(a == b ? 1 : 0), 0};
        put(&e, sizeof e);
    }
    const uint64_t h = fnv1a64(0xcbf29ce484222325ull, buf.data(), buf.size());
    put(&h, 8);
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) {
        err = "cannot open " + path + " for writing";
        return false;
    }
    const bool ok = std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
    if (std::fclose(f) != 0 || !ok) {
        err = "write to " + path + " failed";
        return false;
    }
    return true;
}

bool load_partition(const std::string& path, BlockPartition& P, std::string& err) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    std::vector<unsigned char> buf;
    unsigned char tmp[1 << 16];
    size_t n;
    while ((n = std::fread(tmp, 1, sizeof tmp, f)) > 0) buf.insert(buf.end(), tmp, tmp + n);
    std::fclose(f);
    size_t at = 0;
    auto get = [&](void* p, size_t cnt) {
        if (at + cnt > buf.size()) return false;
        std::memcpy(p, buf.data() + at, cnt);
        at += cnt;
        return true;
    };
    char magic[8];
    uint32_t hdr[6];
    BlockPartition Q;
    if (!get(magic, 8) || std::memcmp(magic, kBlMagic, 8) != 0) {
        err = path + ": not a block-list file";
        return false;
    }
    if (!get(hdr, sizeof hdr) || hdr[0] != kBlVersion) {
        err = path + ": unsupported block-list version";
        return false;
    }
    const uint32_t ng = hdr[1], nrwg = hdr[3], nfar = hdr[4], nnear = hdr[5];
    Q.nrwg_mesh = (int32_t)hdr[2];
    // sizes must fit the file before anything is allocated
    const size_t need = 16 + (size_t)(ng + 1) * 4 + (size_t)nrwg * 4 + (size_t)ng * 32 +
                        (size_t)(nfar + nnear) * sizeof(BlEntry) + 8;
    if (buf.size() < sizeof magic + sizeof hdr + need) {
        err = path + ": truncated block-list file";
        return false;
    }
    uint64_t h_file;
    std::memcpy(&h_file, buf.data() + buf.size() - 8, 8);
    if (fnv1a64(0xcbf29ce484222325ull, buf.data(), buf.size() - 8) != h_file) {
        err = path + ": checksum mismatch (corrupt block-list file)";
        return false;
    }
    Q.group_offsets.resize(ng + 1);
    Q.group_rwg.resize(nrwg);
    Q.group_centre.resize(3 * (size_t)ng);
    Q.group_diam.resize(ng);
    get(&Q.leaf_size, 8);
    get(&Q.eta_adm, 8);
    get(Q.group_offsets.data(), Q.group_offsets.size() * 4);
    get(Q.group_rwg.data(), Q.group_rwg.size() * 4);
    get(Q.group_centre.data(), Q.group_centre.size() * 8);
    get(Q.group_diam.data(), Q.group_diam.size() * 8);
    if (Q.group_offsets[0] != 0 || Q.group_offsets[ng] != (int32_t)nrwg) {
        err = path + ": group offsets do not cover the RWG list";
        return false;
    }
    for (uint32_t g = 0; g < ng; ++g)
        if (Q.group_offsets[g + 1] < Q.group_offsets[g]) {
            err = path + ": group offsets decrease";
            return false;
        }
    for (int32_t e : Q.group_rwg)
        if (e < 0 || e >= Q.nrwg_mesh) {
            err = path + ": RWG id out of range";
            return false;
        }
    for (uint32_t i = 0; i < nfar + nnear; ++i) {
        BlEntry e;
        get(&e, sizeof e);
        const bool far = i < nfar;
        if (e.a < 0 || e.b < 0 || e.a >= (int32_t)ng || e.b >= (int32_t)ng || e.comp != (far ? 1 : 0) ||
            (far ? e.a >= e.b : e.a > e.b) || e.self != (e.a == e.b ? 1 : 0)) {
            err = path + ": bad block entry";
            return false;
        }
        auto& dst = far ? Q.far_pairs : Q.near_pairs;
        dst.push_back(e.a);
        dst.push_back(e.b);
    }
    P = std::move(Q);
    return true;
}

}  // namespace farcur
