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
// disjoint column pairs in parallel, a team of T lanes per pair, with the three
// inner products reduced across the team by lane shuffles.
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

// All threads of the workgroup must call. A, V: r x r col-major scratch (LDS or
// global). On exit U (ld = ldu) holds W^+ with singular values below
// tau = max(atol, rtol * sigma_max) dropped, sigma[j] the singular value carried
// by column j (unsorted), *kept_out the number of sigma >= tau and *smax_out
// sigma_max, *rank_eps_out the number of sigma >= label_eps * sigma_max.
// s_flag: one int of LDS; s_w: r doubles of LDS or global.
__device__ __forceinline__ void jacobi_pinv(c128* A, c128* V, int r, double atol, double rtol,
                                            c128* __restrict__ U, int ldu, double* __restrict__ sigma, int* s_flag,
                                            double* s_w, int* kept_out, double* smax_out, double label_eps,
                                            int* rank_eps_out, int* sweeps_out) {
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int idx = tid; idx < r * r; idx += nt) {
        const int i = idx % r, j = idx / r;
        V[idx] = c128{i == j ? 1.0 : 0.0, 0.0};
    }
    const int n = r + (r & 1);
    const int npairs = n / 2;
    int T = 1;
    while (T * 2 <= kWave && T * 2 * npairs <= nt) T *= 2;
    const int nteams = nt / T;
    const int team = tid / T, tl = tid - team * T;
    const double tol = fmax(1e-14, 8.8e-16 * r);
    __syncthreads();
    // s_flag holds the largest relative coupling |a_p^H a_q| / (|a_p| |a_q|)
    // rotated in the sweep (float bits; 0 = no rotation). Jacobi converges
    // quadratically, so once every coupling of a sweep is below 1e-7 the next
    // sweep would only see couplings near the rounding level and is skipped.
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
                double al = 0, be = 0, gr = 0, gi = 0;
                for (int i = tl; i < r; i += T) {
                    const c128 a = ap[i], b = aq[i];
                    al += a.re * a.re + a.im * a.im;
                    be += b.re * b.re + b.im * b.im;
                    gr += a.re * b.re + a.im * b.im;
                    gi += a.re * b.im - a.im * b.re;
                }
                al = team_sum(al, T);
                be = team_sum(be, T);
                gr = team_sum(gr, T);
                gi = team_sum(gi, T);
                const double g = sqrt(gr * gr + gi * gi);
                const double nrm = sqrt(al * be);
                if (!(g > tol * nrm)) continue;
                const double zeta = (be - al) / (2.0 * g);
                const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / sqrt(1.0 + t * t);
                const double s = c * t;
                const double er = gr / g, ei = gi / g;
                c128* vp = V + (size_t)p * r;
                c128* vq = V + (size_t)q * r;
                for (int i = tl; i < r; i += T) {
                    const c128 a = ap[i], b = aq[i];
                    const double btr = b.re * er + b.im * ei, bti = b.im * er - b.re * ei;
                    ap[i] = c128{c * a.re - s * btr, c * a.im - s * bti};
                    aq[i] = c128{s * a.re + c * btr, s * a.im + c * bti};
                    const c128 x = vp[i], y = vq[i];
                    const double ytr = y.re * er + y.im * ei, yti = y.im * er - y.re * ei;
                    vp[i] = c128{c * x.re - s * ytr, c * x.im - s * yti};
                    vq[i] = c128{s * x.re + c * ytr, s * x.im + c * yti};
                }
                if (tl == 0) atomicMax(s_flag, __float_as_int(fmaxf((float)(g / nrm), 1e-30f)));
            }
            __syncthreads();
        }
        const int maxc = *s_flag;
        __syncthreads();
        if (maxc == 0 || __int_as_float(maxc) < 1e-7f) break;
    }
    for (int j = team; j < r; j += nteams) {
        const c128* aj = A + (size_t)j * r;
        double nn = 0;
        for (int i = tl; i < r; i += T) nn += aj[i].re * aj[i].re + aj[i].im * aj[i].im;
        nn = team_sum(nn, T);
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

}  // namespace farcur
