// Workgroup-level pseudo-inverse of the r x r complex128 core (intersection)
// matrix W by one-sided (Hestenes) Jacobi SVD.
//
// Replaces cusolverDnZgesvd + kernel_Sinv + cublasZdgmm + cublasZgemm at
// codes/GPU_implementation/seed/user_impedance_for_julia_cuda.cu:1873-1924
// (threshold kernel CU:1435-1444): U_core = V diag(s >= tol ? 1/s : 0) U_svd^H
// with an ABSOLUTE threshold on the singular values (pinv_tol = 1e-10 in the
// reference, compres_single_block_spheres_for_sweep2.jl:884 and :1001-1009).
//
// Method: columns of A (= W on entry) are rotated in pairs until mutually
// orthogonal; the same rotations accumulate into V (starting from I), so that
// A_final = W V = U_svd diag(s). Then W^+ = sum_j [s_j >= tol] v_j a_j^H / s_j^2.
// Pairs follow the round-robin (circle) schedule: each step rotates n/2
// disjoint column pairs in parallel, a team of T lanes per pair.
//
// The loop is latency-bound (one barrier per step, a dependent chain of
// reductions and rotation parameters per step), so the step is kept short:
// the team's inner products are reduced with DPP row moves (T <= 16 lanes
// stay inside one 16-lane DPP row), the rotation parameters use the hardware
// reciprocal / reciprocal square root refined by two Newton steps each, and
// with kE > 0 a lane keeps its kE elements of the pair's two columns in
// registers between the inner products and the rotation.
#pragma once
#include "common.hpp"

namespace farcur {

constexpr int kJacobiMaxSweeps = 40;

// Column pair (p, q) of pair index pi at step `step` of the circle schedule
// over n (even) players.
__device__ __forceinline__ void circle_pair(int n, int step, int pi, int& p, int& q) {
    if (pi == 0) {
        p = step;
        q = n - 1;
    } else {
        p = (step + pi) % (n - 1);
        q = (step - pi + (n - 1)) % (n - 1);
    }
}

// Row-local DPP move of a double (both halves); every source lane is inside
// the row, so no lane reads out of range.
template <int kCtrl>
__device__ __forceinline__ double dpp_move(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), kCtrl, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), kCtrl, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// Sum over the kT consecutive lanes of a team (kT a power of two <= 64). Both
// partners of every exchange add the same two operands, so all lanes of the
// team end with the same bits.
template <int kT>
__device__ __forceinline__ double team_sum_t(double v) {
    if constexpr (kT >= 2) v += dpp_move<0xB1>(v);   // quad_perm [1,0,3,2]: lane ^ 1
    if constexpr (kT >= 4) v += dpp_move<0x4E>(v);   // quad_perm [2,3,0,1]: lane ^ 2
    if constexpr (kT >= 8) v += dpp_move<0x141>(v);  // row_half_mirror: lane i <-> 7 - i in 8
    if constexpr (kT >= 16) v += dpp_move<0x140>(v); // row_mirror: lane i <-> 15 - i in 16
    if constexpr (kT >= 32) v += __shfl_xor(v, 16, 32);
    if constexpr (kT >= 64) v += __shfl_xor(v, 32, 64);
    return v;
}

// 1 / sqrt(x) and 1 / x for x > 0 (finite): hardware estimate, then two
// Newton-Raphson steps (y <- y + y (1 - x y^2) / 2, y <- y + y (1 - x y)).
__device__ __forceinline__ double rsqrt_nr(double x) {
    double y = __builtin_amdgcn_rsq(x);
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const double e = fma(-x * y, y, 1.0);
        y = fma(0.5 * y, e, y);
    }
    return y;
}
__device__ __forceinline__ double rcp_nr(double x) {
    double y = __builtin_amdgcn_rcp(x);
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const double e = fma(-x, y, 1.0);
        y = fma(y, e, y);
    }
    return y;
}

// Rotation of columns (a_p, a_q) with |a_p|^2 = al, |a_q|^2 = be and
// a_p^H a_q = g (|g|^2 = g2 > 0) that makes them orthogonal:
//   a_p' = c a_p - s conj(e) a_q,  a_q' = s a_p + c conj(e) a_q,  e = g / |g|,
// with t = tan(theta) the smaller root of t^2 + 2 zeta t - 1 = 0,
// zeta = (be - al) / (2 |g|).
struct Rot {
    double c, s, er, ei;
};
__device__ __forceinline__ Rot rotation(double al, double be, double gr, double gi, double g2) {
    const double ig = rsqrt_nr(g2);
    const double zeta = 0.5 * (be - al) * ig;
    const double z2 = fma(zeta, zeta, 1.0);
    const double t = (zeta >= 0 ? 1.0 : -1.0) * rcp_nr(fabs(zeta) + z2 * rsqrt_nr(z2));
    const double c = rsqrt_nr(fma(t, t, 1.0));
    return Rot{c, c * t, gr * ig, gi * ig};
}

__device__ __forceinline__ void rotate_pair(const Rot& R, c128& a, c128& b) {
    const double btr = b.re * R.er + b.im * R.ei, bti = b.im * R.er - b.re * R.ei;
    const c128 x = a;
    a = c128{R.c * x.re - R.s * btr, R.c * x.im - R.s * bti};
    b = c128{R.s * x.re + R.c * btr, R.s * x.im + R.c * bti};
}

// All threads of the workgroup must call. A, V: r x r col-major scratch (LDS or
// global). Teams of kT lanes; with kE > 0 every r <= kE * kT (elements cached
// in registers), with kE = 0 any r (elements streamed). On exit U (ld = ldu)
// holds W^+ with singular values below tau = max(atol, rtol * sigma_max)
// dropped, sigma[j] the singular value carried by column j (unsorted),
// *kept_out the number of sigma >= tau and *smax_out sigma_max, *rank_eps_out
// the number of sigma >= label_eps * sigma_max, *sweeps_out the sweeps run.
// s_flag: one int of LDS; s_w: r doubles of LDS or global.
template <int kT, int kE>
__device__ __forceinline__ void jacobi_pinv_t(c128* A, c128* V, int r, double atol, double rtol,
                                              c128* __restrict__ U, int ldu, double* __restrict__ sigma,
                                              int* s_flag, double* s_w, int* kept_out, double* smax_out,
                                              double label_eps, int* rank_eps_out, int* sweeps_out) {
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int idx = tid; idx < r * r; idx += nt) {
        const int i = idx % r, j = idx / r;
        V[idx] = c128{i == j ? 1.0 : 0.0, 0.0};
    }
    const int n = r + (r & 1);
    const int npairs = n / 2;
    const int nteams = nt / kT;
    const int team = tid / kT, tl = tid - team * kT;
    const double tol = fmax(1e-14, 8.8e-16 * r);
    const double tol2 = tol * tol;
    __syncthreads();
    // s_flag holds the largest squared relative coupling |a_p^H a_q|^2 /
    // (|a_p|^2 |a_q|^2) rotated in the sweep (float bits; 0 = no rotation).
    // Jacobi converges quadratically, so once every coupling of a sweep is
    // below 1e-7 the next sweep would only see couplings near the rounding
    // level and is skipped.
    int sweeps = 0;
    for (int sweep = 0; sweep < kJacobiMaxSweeps && n > 1; ++sweep) {
        ++sweeps;
        if (tid == 0) *s_flag = 0;
        __syncthreads();
        for (int step = 0; step < n - 1; ++step) {
            for (int pi = team; pi < npairs; pi += nteams) {
                int p, q;
                circle_pair(n, step, pi, p, q);
                if (p >= r || q >= r) continue;
                c128* ap = A + (size_t)p * r;
                c128* aq = A + (size_t)q * r;
                c128* vp = V + (size_t)p * r;
                c128* vq = V + (size_t)q * r;
                double al = 0, be = 0, gr = 0, gi = 0;
                if constexpr (kE > 0) {
                    c128 a[kE], b[kE];
#pragma unroll
                    for (int e = 0; e < kE; ++e) {
                        const int i = tl + e * kT;
                        a[e] = b[e] = c128{0.0, 0.0};
                        if (i < r) {
                            a[e] = ap[i];
                            b[e] = aq[i];
                        }
                        al += a[e].re * a[e].re + a[e].im * a[e].im;
                        be += b[e].re * b[e].re + b[e].im * b[e].im;
                        gr += a[e].re * b[e].re + a[e].im * b[e].im;
                        gi += a[e].re * b[e].im - a[e].im * b[e].re;
                    }
                    al = team_sum_t<kT>(al);
                    be = team_sum_t<kT>(be);
                    gr = team_sum_t<kT>(gr);
                    gi = team_sum_t<kT>(gi);
                    const double g2 = gr * gr + gi * gi;
                    if (!(g2 > tol2 * (al * be))) continue;
                    const Rot R = rotation(al, be, gr, gi, g2);
#pragma unroll
                    for (int e = 0; e < kE; ++e) {
                        const int i = tl + e * kT;
                        if (i < r) {
                            rotate_pair(R, a[e], b[e]);
                            ap[i] = a[e];
                            aq[i] = b[e];
                            c128 x = vp[i], y = vq[i];
                            rotate_pair(R, x, y);
                            vp[i] = x;
                            vq[i] = y;
                        }
                    }
                    if (tl == 0) atomicMax(s_flag, __float_as_int(fmaxf((float)(g2 * rcp_nr(al * be)), 1e-30f)));
                } else {
                    for (int i = tl; i < r; i += kT) {
                        const c128 a = ap[i], b = aq[i];
                        al += a.re * a.re + a.im * a.im;
                        be += b.re * b.re + b.im * b.im;
                        gr += a.re * b.re + a.im * b.im;
                        gi += a.re * b.im - a.im * b.re;
                    }
                    al = team_sum_t<kT>(al);
                    be = team_sum_t<kT>(be);
                    gr = team_sum_t<kT>(gr);
                    gi = team_sum_t<kT>(gi);
                    const double g2 = gr * gr + gi * gi;
                    if (!(g2 > tol2 * (al * be))) continue;
                    const Rot R = rotation(al, be, gr, gi, g2);
                    for (int i = tl; i < r; i += kT) {
                        c128 a = ap[i], b = aq[i];
                        rotate_pair(R, a, b);
                        ap[i] = a;
                        aq[i] = b;
                        c128 x = vp[i], y = vq[i];
                        rotate_pair(R, x, y);
                        vp[i] = x;
                        vq[i] = y;
                    }
                    if (tl == 0) atomicMax(s_flag, __float_as_int(fmaxf((float)(g2 * rcp_nr(al * be)), 1e-30f)));
                }
            }
            __syncthreads();
        }
        const int maxc = *s_flag;
        __syncthreads();
        if (maxc == 0 || __int_as_float(maxc) < 1e-14f) break;
    }
    for (int j = team; j < r; j += nteams) {
        const c128* aj = A + (size_t)j * r;
        double nn = 0;
        for (int i = tl; i < r; i += kT) nn += aj[i].re * aj[i].re + aj[i].im * aj[i].im;
        nn = team_sum_t<kT>(nn);
        if (tl == 0) {
            sigma[j] = sqrt(nn);
            s_w[j] = nn;
        }
    }
    __syncthreads();
    double smax = 0;
    for (int j = 0; j < r; ++j) smax = fmax(smax, sigma[j]);
    const double tau = fmax(atol, rtol * smax);
    int kept = 0, reps = 0;
    for (int j = 0; j < r; ++j) {
        kept += (sigma[j] >= tau && s_w[j] > 0);
        reps += (sigma[j] >= label_eps * smax && smax > 0);
    }
    __syncthreads();
    for (int j = tid; j < r; j += nt) {
        const double nn = s_w[j];
        s_w[j] = (sigma[j] >= tau && nn > 0) ? 1.0 / nn : 0.0;
    }
    __syncthreads();
    // U[i, k] = sum_j V[i, j] w_j conj(A[k, j])
    for (int idx = tid; idx < r * r; idx += nt) {
        const int i = idx % r, k = idx / r;
        double ur = 0, ui = 0;
        for (int j = 0; j < r; ++j) {
            const double w = s_w[j];
            if (w == 0.0) continue;
            const c128 v = V[(size_t)j * r + i];
            const c128 a = A[(size_t)j * r + k];
            // v * conj(a) * w
            ur += w * (v.re * a.re + v.im * a.im);
            ui += w * (v.im * a.re - v.re * a.im);
        }
        U[(size_t)k * ldu + i] = c128{ur, ui};
    }
    *kept_out = kept;
    *smax_out = smax;
    *rank_eps_out = reps;
    *sweeps_out = sweeps;
    __syncthreads();
}

// Team width for n/2 pairs on nt threads: the largest power of two <= 64 with
// T * pairs <= nt (at least 1).
__host__ __device__ inline int jacobi_team(int r, int nt) {
    const int npairs = (r + (r & 1)) / 2;
    int T = 1;
    while (T * 2 <= kWave && T * 2 * npairs <= nt) T *= 2;
    return T;
}

}  // namespace farcur
