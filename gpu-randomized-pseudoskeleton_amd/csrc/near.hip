// Near-field blocks of the EFIE impedance matrix and the compressed
// operator's near part (dense blocks times a vector).
//
// The reference fills its non-admissible blocks densely (the comp == 0 branch
// of rebuild_matrix_cuda2, codes/GPU_implementation/seed/compres_single_block_spheres.jl:2639-2662)
// with an element that integrates close triangle pairs by its singular rules
// (Rinteg_s / Ranal_s options, compres_single_block_spheres_for_sweep2.jl:1095-1098;
// element binary-only, declared at user_impedance_for_julia_cuda.cu:178-214).
// Here, per triangle pair of an entry:
//   - apart (centroid distance >= near_factor x the longer longest edge): the
//     3 x 3 Gauss rule of the far-field element (entry_f64);
//   - near: singularity extraction, G = 1/(4 pi R) + (e^{-jkR} - 1)/(4 pi R),
//     the 1/R part over the source triangle in closed form (Wilton, Rao,
//     Glisson, Schaubert, Al-Bundak, Butler, IEEE Trans. AP-32(3), 1984), the
//     bounded remainder by the 7-point degree-5 rule (Dunavant 1985), and the
//     test triangle by the same 7-point rule.
// Triangle geometry comes from the FP64 RWG records: with quadrature points
// q_i = V_i / 2 + (V_0 + V_1 + V_2) / 6, V_i = 2 q_i - (q_0 + q_1 + q_2) / 3,
// and the free vertex v = q_0 - s rho_0.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "element.hpp"
#include "near.hpp"

namespace farcur {

namespace {

struct D3 {
    double x, y, z;
};
__device__ __forceinline__ D3 operator-(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ D3 operator+(D3 a, D3 b) { return D3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ D3 operator*(double s, D3 a) { return D3{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 cross(D3 a, D3 b) {
    return D3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// Dunavant's 7-point degree-5 rule: barycentric (a, b, c) and weight (sum 1).
__constant__ double kD7[7][4] = {
    {1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0, 0.225},
    {0.059715871789770, 0.470142064105115, 0.470142064105115, 0.132394152788506},
    {0.470142064105115, 0.059715871789770, 0.470142064105115, 0.132394152788506},
    {0.470142064105115, 0.470142064105115, 0.059715871789770, 0.132394152788506},
    {0.797426985353087, 0.101286507323456, 0.101286507323456, 0.125939180544827},
    {0.101286507323456, 0.797426985353087, 0.101286507323456, 0.125939180544827},
    {0.101286507323456, 0.101286507323456, 0.797426985353087, 0.125939180544827},
};

// One half (triangle) of an RWG from its FP64 record.
struct Half {
    D3 v[3];   // triangle vertices (record order)
    D3 free;   // free vertex
    D3 c;      // centroid
    double area, lmax, s;
};

__device__ Half load_half(const double* __restrict__ rec, int h) {
    Half H;
    D3 q[3];
    for (int i = 0; i < 3; ++i) q[i] = D3{rec[h * 9 + i * 3], rec[h * 9 + i * 3 + 1], rec[h * 9 + i * 3 + 2]};
    H.c = (1.0 / 3.0) * (q[0] + q[1] + q[2]);
    for (int i = 0; i < 3; ++i) H.v[i] = 2.0 * q[i] - H.c;
    H.s = h == 0 ? 1.0 : -1.0;
    const D3 p0{rec[18 + h * 9], rec[18 + h * 9 + 1], rec[18 + h * 9 + 2]};
    H.free = q[0] - H.s * p0;
    const D3 cr = cross(H.v[1] - H.v[0], H.v[2] - H.v[0]);
    H.area = 0.5 * sqrt(dot(cr, cr));
    double m = 0;
    for (int i = 0; i < 3; ++i) {
        const D3 d = H.v[(i + 1) % 3] - H.v[i];
        m = fmax(m, dot(d, d));
    }
    H.lmax = sqrt(m);
    return H;
}

// int_T 1/R dS' (is) and int_T (r' - rho)/R dS' (iv) for the point r, rho its
// projection on the plane of T (the closed forms of Wilton et al. 1984).
__device__ void potentials(const Half& T, D3 r, double& is, D3& iv) {
    D3 n = cross(T.v[1] - T.v[0], T.v[2] - T.v[0]);
    n = (1.0 / sqrt(dot(n, n))) * n;
    const double h = dot(r - T.v[0], n), ah = fabs(h);
    const D3 rho = r - h * n;
    is = 0;
    iv = D3{0, 0, 0};
    for (int i = 0; i < 3; ++i) {
        const D3 a = T.v[i], b = T.v[(i + 1) % 3];
        D3 l = b - a;
        const double len = sqrt(dot(l, l));
        l = (1.0 / len) * l;
        const D3 u = cross(l, n);
        const double lp = dot(b - rho, l), lm = dot(a - rho, l), p0 = dot(a - rho, u);
        const double r02 = p0 * p0 + h * h;
        const D3 rb = r - b, ra = r - a;
        const double rp = sqrt(dot(rb, rb)), rm = sqrt(dot(ra, ra));
        double f = 0;
        if (r02 > 1e-28 * len * len)
            f = (lp + lm >= 0) ? log((rp + lp) / (rm + lm)) : log((rm - lm) / (rp - lp));
        // else: on the edge's line, where f's weights p0 and r02 vanish
        is += p0 * f;
        if (ah > 0) is -= ah * (atan2(p0 * lp, r02 + ah * rp) - atan2(p0 * lm, r02 + ah * rm));
        const double vv = 0.5 * (r02 * f + lp * rp - lm * rm);
        iv = iv + vv * u;
    }
}

// int_Ta int_Tb [(r - va).(r' - vb) - 4/k^2] G dS' dS by singularity extraction.
__device__ void near_pair(const Half& A, const Half& B, double k, double& jr, double& ji) {
    D3 nb = cross(B.v[1] - B.v[0], B.v[2] - B.v[0]);
    nb = (1.0 / sqrt(dot(nb, nb))) * nb;
    const double c4 = 4.0 / (k * k);
    double sr = 0, si = 0;
    for (int p = 0; p < 7; ++p) {
        const D3 r = kD7[p][0] * A.v[0] + kD7[p][1] * A.v[1] + kD7[p][2] * A.v[2];
        double is;
        D3 iv;
        potentials(B, r, is, iv);
        const double h = dot(r - B.v[0], nb);
        const D3 i1s = iv + is * ((r - h * nb) - B.free);  // int (r' - vb) / R
        double i0r = 0, i0i = 0;
        D3 i1r{0, 0, 0}, i1i{0, 0, 0};
        for (int q = 0; q < 7; ++q) {
            const D3 rs = kD7[q][0] * B.v[0] + kD7[q][1] * B.v[1] + kD7[q][2] * B.v[2];
            const D3 d = r - rs;
            const double R = sqrt(dot(d, d));
            double gr, gi;  // (e^{-jkR} - 1) / R, with its limit -jk at R = 0
            if (R > 1e-12 * sqrt(B.area)) {
                double sn, cs;
                sincos(k * R, &sn, &cs);
                gr = (cs - 1.0) / R;
                gi = -sn / R;
            } else {
                gr = 0;
                gi = -k;
            }
            const double w = kD7[q][3] * B.area;
            i0r += w * gr;
            i0i += w * gi;
            const D3 x = rs - B.free;
            i1r = i1r + (w * gr) * x;
            i1i = i1i + (w * gi) * x;
        }
        const D3 xa = r - A.free;
        const double tr = dot(xa, i1s + i1r) - c4 * (is + i0r);
        const double ti = dot(xa, i1i) - c4 * i0i;
        const double wp = kD7[p][3] * A.area;
        sr += wp * tr;
        si += wp * ti;
    }
    jr = sr / (4.0 * M_PI);
    ji = si / (4.0 * M_PI);
}

// Z_mn with the near-pair rule. a, b: FP64 records of m (test) and n (source).
__device__ void entry_near(const double* __restrict__ a, const double* __restrict__ b, double k, double eta,
                           double near_factor, double& zr, double& zi) {
    double sr = 0, si = 0;
    const double lm = a[36], ln = b[36];
    for (int ha = 0; ha < 2; ++ha) {
        const Half A = load_half(a, ha);
        for (int hb = 0; hb < 2; ++hb) {
            const Half B = load_half(b, hb);
            const D3 d = A.c - B.c;
            const double lim = near_factor * fmax(A.lmax, B.lmax);
            const double coef = A.s * B.s * lm * ln / (4.0 * A.area * B.area);
            if (dot(d, d) < lim * lim) {
                double jr, ji;
                near_pair(A, B, k, jr, ji);
                sr += coef * jr;
                si += coef * ji;
            } else {
                // the far element's 3 x 3 rule: rho_ai . rho_bj - s_a s_b 4 / k^2
                const double ss = (ha == hb ? 1.0 : -1.0) * 4.0 / (k * k);
                const double w = lm * ln / 36.0 / (4.0 * M_PI);
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) {
                        const D3 qa{a[ha * 9 + i * 3], a[ha * 9 + i * 3 + 1], a[ha * 9 + i * 3 + 2]};
                        const D3 qb{b[hb * 9 + j * 3], b[hb * 9 + j * 3 + 1], b[hb * 9 + j * 3 + 2]};
                        const D3 pa{a[18 + ha * 9 + i * 3], a[18 + ha * 9 + i * 3 + 1], a[18 + ha * 9 + i * 3 + 2]};
                        const D3 pb{b[18 + hb * 9 + j * 3], b[18 + hb * 9 + j * 3 + 1], b[18 + hb * 9 + j * 3 + 2]};
                        const D3 dq = qa - qb;
                        const double R = sqrt(dot(dq, dq));
                        double sn, cs;
                        sincos(k * R, &sn, &cs);
                        const double t = (dot(pa, pb) - ss) * w / R;
                        sr += t * cs;
                        si -= t * sn;
                    }
            }
        }
    }
    // times -j k eta
    zr = k * eta * si;
    zi = -k * eta * sr;
}

// One wave per 64 rows of a block; each lane walks its row's columns.
__global__ __launch_bounds__(64) void k_near_fill(const NearJob* __restrict__ jobs, const int32_t* __restrict__ tile_job,
                                                  const int32_t* __restrict__ tile_row0, const double* __restrict__ rec,
                                                  double k, double eta, double near_factor) {
    const NearJob J = jobs[tile_job[blockIdx.x]];
    const int i = tile_row0[blockIdx.x] + (int)threadIdx.x;
    if (i >= J.m) return;
    const double* a = rec + (size_t)J.rows[i] * kRecStride;
    for (int j = 0; j < J.n; ++j) {
        double zr, zi;
        entry_near(a, rec + (size_t)J.cols[j] * kRecStride, k, eta, near_factor, zr, zi);
        J.Z[(size_t)j * J.m + i] = c128{zr, zi};
    }
}

// y[rows] += Z x[cols] (orientation 0) or y[cols] += Z^T x[rows] (1), one
// workgroup per (block, orientation), one output row per thread, FP64 atomics.
__global__ __launch_bounds__(256) void k_near_matvec(const NearMv* __restrict__ jobs, const c128* __restrict__ x,
                                                     double* __restrict__ y) {
    const NearMv J = jobs[blockIdx.x];
    const NearJob& B = J.blk;
    if (J.o == 0) {
        for (int i = threadIdx.x; i < B.m; i += blockDim.x) {
            double ar = 0, ai = 0;
            for (int j = 0; j < B.n; ++j) {
                const c128 z = B.Z[(size_t)j * B.m + i], v = x[B.cols[j]];
                ar += z.re * v.re - z.im * v.im;
                ai += z.re * v.im + z.im * v.re;
            }
            atomicAdd(&y[2 * (size_t)B.rows[i]], ar);
            atomicAdd(&y[2 * (size_t)B.rows[i] + 1], ai);
        }
    } else {
        for (int j = threadIdx.x; j < B.n; j += blockDim.x) {
            double ar = 0, ai = 0;
            const c128* zc = B.Z + (size_t)j * B.m;
            for (int i = 0; i < B.m; ++i) {
                const c128 z = zc[i], v = x[B.rows[i]];
                ar += z.re * v.re - z.im * v.im;
                ai += z.re * v.im + z.im * v.re;
            }
            atomicAdd(&y[2 * (size_t)B.cols[j]], ar);
            atomicAdd(&y[2 * (size_t)B.cols[j] + 1], ai);
        }
    }
}

__global__ void k_to_c64(const c128* __restrict__ x, c64* __restrict__ y, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        y[i] = c64{(float)x[i].re, (float)x[i].im};
}

__global__ void k_add_c64(c128* __restrict__ y, const c64* __restrict__ a, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        y[i].re += a[i].re;
        y[i].im += a[i].im;
    }
}

}  // namespace

void launch_near_fill(const NearJob* jobs, int ntiles, const int32_t* tile_job, const int32_t* tile_row0,
                      const double* rec, double k, double eta, double near_factor, hipStream_t s) {
    if (ntiles <= 0) return;
    hipLaunchKernelGGL(k_near_fill, dim3(ntiles), dim3(64), 0, s, jobs, tile_job, tile_row0, rec, k, eta, near_factor);
}

void launch_near_matvec(const NearMv* jobs, int njobs, const c128* x, c128* y, hipStream_t s) {
    if (njobs <= 0) return;
    hipLaunchKernelGGL(k_near_matvec, dim3(njobs), dim3(256), 0, s, jobs, x, (double*)y);
}

void launch_to_c64(const c128* x, c64* y, int n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_to_c64, dim3(std::min((n + 255) / 256, 4096)), dim3(256), 0, s, x, y, n);
}

void launch_add_c64(c128* y, const c64* a, int n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_add_c64, dim3(std::min((n + 255) / 256, 4096)), dim3(256), 0, s, y, a, n);
}

}  // namespace farcur
