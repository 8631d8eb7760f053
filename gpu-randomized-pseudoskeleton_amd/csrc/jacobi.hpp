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
// over n (even) players: p = (step + pi) mod (n - 1), q = (step - pi) mod
// (n - 1) for pi > 0, (step, n - 1) for pi = 0. With step < n - 1 and
// pi < n / 2 both sums lie in [0, 2 (n - 1)), so one conditional subtraction
// reduces them.
__device__ __forceinline__ void circle_pair(int n, int step, int pi, int& p, int& q) {
    const int m = n - 1;
    if (pi == 0) {
        p = step;
        q = m;
    } else {
        p = step + pi;
        p -= p >= m ? m : 0;
        q = step - pi + m;
        q -= q >= m ? m : 0;
    }
}

// Row-local DPP move of a double (both halves); every source lane is inside
// the row, so no lane reads out of range and no fill value is needed.
template <int kCtrl>
__device__ __forceinline__ double dpp_move(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), kCtrl, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), kCtrl, 0xF, 0xF, true);
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

// 1 / sqrt(x) and 1 / x for x > 0 (finite): hardware estimate, then kSteps
// Newton-Raphson steps (y <- y + y (1 - x y^2) / 2, y <- y + y (1 - x y));
// each step doubles the correct bits, two give full double precision.
template <int kSteps = 2>
__device__ __forceinline__ double rsqrt_nr(double x) {
    double y = __builtin_amdgcn_rsq(x);
#pragma unroll
    for (int i = 0; i < kSteps; ++i) {
        const double e = fma(-x * y, y, 1.0);
        y = fma(0.5 * y, e, y);
    }
    return y;
}
template <int kSteps = 2>
__device__ __forceinline__ double rcp_nr(double x) {
    double y = __builtin_amdgcn_rcp(x);
#pragma unroll
    for (int i = 0; i < kSteps; ++i) {
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
// c and s come from t as c = 1 / sqrt(1 + t^2), s = c t, so the rotation is
// orthogonal whatever t's rounding: t needs only enough accuracy to cancel
// most of the coupling (the next sweep removes the rest), and its two
// estimates take one Newton step each; e (|e| = 1) and c take two.
__device__ __forceinline__ Rot rotation(double al, double be, double gr, double gi, double g2) {
    const double ig = rsqrt_nr(g2);
    const double zeta = 0.5 * (be - al) * ig;
    const double z2 = fma(zeta, zeta, 1.0);
    const double t = (zeta >= 0 ? 1.0 : -1.0) * rcp_nr<1>(fabs(zeta) + z2 * rsqrt_nr<1>(z2));
    const double c = rsqrt_nr(fma(t, t, 1.0));
    return Rot{c, c * t, gr * ig, gi * ig};
}

__device__ __forceinline__ void rotate_pair(const Rot& R, c128& a, c128& b) {
    const double btr = b.re * R.er + b.im * R.ei, bti = b.im * R.er - b.re * R.ei;
    const c128 x = a;
    a = c128{R.c * x.re - R.s * btr, R.c * x.im - R.s * bti};
    b = c128{R.s * x.re + R.c * btr, R.s * x.im + R.c * bti};
}

// Sweeps of the circle schedule over the columns of A (r x r, col-major with
// leading dimension ld, LDS or global), rotations accumulated into V (same
// layout) unless V is NULL. All threads of the workgroup must call; teams of kT lanes; with
// kE > 0 every r <= kE * kT (elements cached in registers), with kE = 0 any r
// (elements streamed). s_flag: one int of LDS. Returns the number of sweeps.
// kWeak: pairs whose columns both have squared norms below weak2 are not
// rotated nor counted for convergence. Columns at rounding level (the null
// space of a rank-deficient A) otherwise keep random couplings that no
// rotation reduces, and the sweeps run to the limit. With weak2 = tau^2 / r
// the columns kept by a threshold tau are unchanged: a rotation leaves the
// smaller of its two norms no larger than either input, so the skipped
// columns hold together less than tau^2.
template <int kT, int kE, bool kWeak = false>
__device__ __forceinline__ int jacobi_sweeps(c128* A, c128* V, int r, int ld, int* s_flag, double weak2 = 0.0) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const int n = r + (r & 1);
    const int npairs = n / 2;
    const int nteams = nt / kT;
    const int team = tid / kT, tl = tid - team * kT;
    const double tol = fmax(1e-14, 8.8e-16 * r);
    const double tol2 = tol * tol;
    __syncthreads();
    // s_flag holds the largest squared relative coupling |a_p^H a_q|^2 /
    // (|a_p|^2 |a_q|^2) rotated in the sweep (float bits; 0 = no rotation).
    // Jacobi converges quadratically: after a sweep whose couplings were all
    // below c the remaining ones are of order c^2. The loop stops after a
    // sweep with every coupling below 1e-5, leaving the columns orthogonal to
    // about 1e-10 relative: the weights 1 / |a_j|^2 and the products below then
    // differ from those of exactly orthogonal columns at that level only.
    int sweeps = 0;
    for (int sweep = 0; sweep < kJacobiMaxSweeps && n > 1; ++sweep) {
        ++sweeps;
        if (tid == 0) *s_flag = 0;
        __syncthreads();
        float cmax = 0.0f;  // this thread's largest rotated coupling of the sweep
        for (int step = 0; step < n - 1; ++step) {
            for (int pi = team; pi < npairs; pi += nteams) {
                int p, q;
                circle_pair(n, step, pi, p, q);
                if (p >= r || q >= r) continue;
                c128* ap = A + (size_t)p * ld;
                c128* aq = A + (size_t)q * ld;
                c128* vp = V ? V + (size_t)p * ld : nullptr;
                c128* vq = V ? V + (size_t)q * ld : nullptr;
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
                    if constexpr (kWeak)
                        if (al < weak2 && be < weak2) continue;
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
                            if (vp) {
                                c128 x = vp[i], y = vq[i];
                                rotate_pair(R, x, y);
                                vp[i] = x;
                                vq[i] = y;
                            }
                        }
                    }
                    cmax = fmaxf(cmax, fmaxf((float)(g2 * __builtin_amdgcn_rcp(al * be)), 1e-30f));
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
                    if constexpr (kWeak)
                        if (al < weak2 && be < weak2) continue;
                    const double g2 = gr * gr + gi * gi;
                    if (!(g2 > tol2 * (al * be))) continue;
                    const Rot R = rotation(al, be, gr, gi, g2);
                    for (int i = tl; i < r; i += kT) {
                        c128 a = ap[i], b = aq[i];
                        rotate_pair(R, a, b);
                        ap[i] = a;
                        aq[i] = b;
                        if (vp) {
                            c128 x = vp[i], y = vq[i];
                            rotate_pair(R, x, y);
                            vp[i] = x;
                            vq[i] = y;
                        }
                    }
                    cmax = fmaxf(cmax, fmaxf((float)(g2 * __builtin_amdgcn_rcp(al * be)), 1e-30f));
                }
            }
            __syncthreads();
        }
        // non-negative floats order like their bit patterns
        if (cmax > 0.0f) atomicMax(s_flag, __float_as_int(cmax));
        __syncthreads();
        const int maxc = *s_flag;
        __syncthreads();
        if (maxc == 0 || __int_as_float(maxc) < 1e-10f) break;
    }
    return sweeps;
}

// Singular values sigma[j] = |a_j| of the converged columns, the pinv weights
// s_w[j] = 1 / sigma_j^2 (1 / sigma_j^4 with quartic) for sigma_j >= tau =
// max(atol, rtol * sigma_max) (else 0), *kept the number of weights set,
// *smax sigma_max and *reps the number of sigma_j >= label_eps * sigma_max.
// All threads must call.
template <int kT>
__device__ __forceinline__ void jacobi_weights(const c128* A, int r, int ld, double atol, double rtol, double label_eps,
                                               double* __restrict__ sigma, double* s_w, int* kept_out,
                                               double* smax_out, int* reps_out, bool quartic = false) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const int nteams = nt / kT, team = tid / kT, tl = tid - team * kT;
    for (int j = team; j < r; j += nteams) {
        const c128* aj = A + (size_t)j * ld;
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
        s_w[j] = (sigma[j] >= tau && nn > 0) ? (quartic ? 1.0 / (nn * nn) : 1.0 / nn) : 0.0;
    }
    __syncthreads();
    *kept_out = kept;
    *smax_out = smax;
    *reps_out = reps;
}

// U[k, map(i)] = sum_j L[i, j] w_j conj(M[k, j]) (L, M r x r col-major with
// leading dimensions ld, ldm; map the identity when rowmap is NULL). All
// threads must call.
__device__ __forceinline__ void pinv_form(const c128* L, const c128* M, const double* s_w, int r, int ld, int ldm,
                                          const int* rowmap, c128* __restrict__ U, int ldu) {
    for (int idx = threadIdx.x; idx < r * r; idx += blockDim.x) {
        const int i = idx % r, k = idx / r;
        double ur = 0, ui = 0;
        for (int j = 0; j < r; ++j) {
            const double w = s_w[j];
            if (w == 0.0) continue;
            const c128 v = L[(size_t)j * ld + i];
            const c128 a = M[(size_t)j * ldm + k];
            // v * conj(a) * w
            ur += w * (v.re * a.re + v.im * a.im);
            ui += w * (v.im * a.re - v.re * a.im);
        }
        U[(size_t)k * ldu + (rowmap ? rowmap[i] : i)] = c128{ur, ui};
    }
}

// One-sided Jacobi pinv of W = A (r x r, col-major, LDS or global; destroyed),
// V: r x r scratch. On exit U (ld = ldu) holds W^+ = sum_j [s_j >= tau]
// v_j a_j^H / s_j^2 with A_final = W V, sigma[j] the singular value carried by
// column j (unsorted), plus the counts of jacobi_weights and the sweeps.
// s_flag: one int of LDS; s_w: r doubles of LDS or global.
template <int kT, int kE>
__device__ __forceinline__ void jacobi_pinv_t(c128* A, c128* V, int r, double atol, double rtol,
                                              c128* __restrict__ U, int ldu, double* __restrict__ sigma,
                                              int* s_flag, double* s_w, int* kept_out, double* smax_out,
                                              double label_eps, int* rank_eps_out, int* sweeps_out) {
    for (int idx = threadIdx.x; idx < r * r; idx += blockDim.x) {
        const int i = idx % r, j = idx / r;
        V[idx] = c128{i == j ? 1.0 : 0.0, 0.0};
    }
    *sweeps_out = jacobi_sweeps<kT, kE>(A, V, r, r, s_flag);
    jacobi_weights<kT>(A, r, r, atol, rtol, label_eps, sigma, s_w, kept_out, smax_out, rank_eps_out);
    pinv_form(V, A, s_w, r, r, r, nullptr, U, ldu);
    __syncthreads();
}

// ---------------------------------------------------------------- QR-preconditioned
// Householder QR with column pivoting, A P = Q R (A r x r col-major in LDS,
// leading dimension ld).
// Step k takes the trailing column of largest norm (first on ties) as logical
// column k and reflects it onto alpha_k e_k with
//   H_k = I - tau_k v_k v_k^H,  v_k = x - alpha_k e_k,  alpha_k = -e^{i arg x_k} |x|,
//   tau_k = 2 / |v_k|^2 = 1 / (|x|^2 + |x| |x_k|)   (real; H_k Hermitian and unitary).
// The map from logical to physical columns and the trailing norms (indexed
// by logical position, so the pivot search reads them directly) are double
// buffered, so each step needs one barrier. On exit:
//   R[i][j] = A[phys[j] * ld + i] for i < j, R[j][j] = s_alpha[j],
//   v_k (rows k..r-1) in A[phys[k] * ld + i] for i >= k, and in
//   Wg[k * r + i] (Wg global, r x r; not stored when Wg is NULL),
//   tau_k in s_tau[k],
//   phys = s_phys (the final map).
// Returns the buffer (0 or 1) of s_phys / s_norm that holds the final map.
template <int kT>
__device__ __forceinline__ int qrcp_lds(c128* A, int r, int ld, int* s_phys, double* s_norm, double* s_tau,
                                        c128* s_alpha, c128* __restrict__ Wg) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const int nteams = nt / kT, team = tid / kT, tl = tid - team * kT;
    __syncthreads();  // A was just written by every thread of the workgroup
    for (int j = tid; j < r; j += nt) s_phys[j] = j;
    for (int j = team; j < r; j += nteams) {
        const c128* aj = A + (size_t)j * ld;
        double nn = 0;
        for (int i = tl; i < r; i += kT) nn += aj[i].re * aj[i].re + aj[i].im * aj[i].im;
        nn = team_sum_t<kT>(nn);
        if (tl == 0) s_norm[j] = nn;
    }
    __syncthreads();
    int cur = 0;
    c128 v0_prev{0.0, 0.0};  // v_{k-1}[k-1]: written to its column's row k-1 in step k,
                             // which does not touch the columns of earlier pivots
    for (int k = 0; k < r; ++k) {
        const int* phys = s_phys + cur * r;
        if (tid == 0 && k > 0) A[(size_t)phys[k - 1] * ld + k - 1] = v0_prev;
        int* phys_next = s_phys + (cur ^ 1) * r;
        const double* norm = s_norm + cur * r;
        double* norm_next = s_norm + (cur ^ 1) * r;
        // Pivot: the largest trailing norm, first logical column on ties, by a
        // wave-wide argmax that every wave computes alike. (value, column) is
        // totally ordered (a non-finite or negative norm counts as -1), so the
        // result does not depend on the order of the combinations. Lane l
        // takes candidates k + l and, past 64 candidates, k + l + 64 (r <= 128);
        // then DPP exchanges within rows of 16 (the rounds that reach past the
        // occupied lanes are skipped), lane swaps across rows only for more
        // than 16 occupied lanes, and lane 0's result is broadcast.
        double best;
        int p;
        {
            const int j = k + (tid & (kWave - 1));
            double v = j < r ? norm[j] : -1.0;
            if (!(v >= 0.0)) v = -1.0;
            int pj = j;
            auto take = [&](double ov, int oj) {
                if (ov > v || (ov == v && oj < pj)) {
                    v = ov;
                    pj = oj;
                }
            };
            if (r - k > kWave) {
                const int j2 = j + kWave;
                double v2 = j2 < r ? norm[j2] : -1.0;
                if (!(v2 >= 0.0)) v2 = -1.0;
                take(v2, j2);
            }
            const int nc = r - k;  // candidates (uniform); occupied lanes min(nc, 64)
            if (nc > 1) take(dpp_move<0xB1>(v), __builtin_amdgcn_mov_dpp(pj, 0xB1, 0xF, 0xF, true));    // lane ^ 1
            if (nc > 2) take(dpp_move<0x4E>(v), __builtin_amdgcn_mov_dpp(pj, 0x4E, 0xF, 0xF, true));    // lane ^ 2
            if (nc > 4) take(dpp_move<0x141>(v), __builtin_amdgcn_mov_dpp(pj, 0x141, 0xF, 0xF, true));  // 7 - i in 8
            if (nc > 8) take(dpp_move<0x140>(v), __builtin_amdgcn_mov_dpp(pj, 0x140, 0xF, 0xF, true));  // 15 - i in 16
            if (nc > 16) take(__shfl_xor(v, 16, kWave), __shfl_xor(pj, 16, kWave));
            if (nc > 32) take(__shfl_xor(v, 32, kWave), __shfl_xor(pj, 32, kWave));
            best = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 0),
                                    __builtin_amdgcn_readlane(__double2loint(v), 0));
            p = __builtin_amdgcn_readlane(pj, 0);
        }
        const int ck = phys[p], co = phys[k];
        const c128 x0 = A[(size_t)ck * ld + k];
        const double a2 = x0.re * x0.re + x0.im * x0.im;
        const double xn = best > 0 ? best * rsqrt_nr(best) : 0.0;
        const double ia = a2 > 0 ? rsqrt_nr(a2) : 0.0;  // 1 / |x_k|
        const double ax0 = a2 * ia;
        c128 alpha;
        if (ax0 > 0) {
            const double f = -xn * ia;
            alpha = c128{f * x0.re, f * x0.im};
        } else {
            alpha = c128{-xn, 0.0};
        }
        const c128 v0{x0.re - alpha.re, x0.im - alpha.im};
        v0_prev = v0;
        const double den = best + xn * ax0;
        const double tau = den > 0 ? rcp_nr(den) : 0.0;
        for (int j = tid; j < r; j += nt) phys_next[j] = j == k ? ck : (j == p ? co : phys[j]);
        if (tid == 0) {
            s_alpha[k] = alpha;
            s_tau[k] = tau;
            if (Wg) Wg[(size_t)k * r + k] = v0;
        }
        // a lane's first row i0 = k + tl is row k itself for tl = 0, whose
        // element of v is v0 (the panel still holds x_k there): peeled, so
        // the loops over the lane's further rows read v plainly
        const c128* v = A + (size_t)ck * ld;
        const int i0 = k + tl;
        const c128 vf = tl == 0 ? v0 : (i0 < r ? v[i0] : c128{0.0, 0.0});
        for (int j = k + 1 + team; j < r; j += nteams) {
            int c = co;
            if (j != p) c = phys[j];
            c128* a = A + (size_t)c * ld;
            // w = v^H a
            double wr = 0, wi = 0;
            if (i0 < r) {
                const c128 ai = a[i0];
                wr += vf.re * ai.re + vf.im * ai.im;
                wi += vf.re * ai.im - vf.im * ai.re;
            }
            for (int i = i0 + kT; i < r; i += kT) {
                const c128 vi = v[i];
                const c128 ai = a[i];
                wr += vi.re * ai.re + vi.im * ai.im;
                wi += vi.re * ai.im - vi.im * ai.re;
            }
            wr = team_sum_t<kT>(wr) * tau;
            wi = team_sum_t<kT>(wi) * tau;
            double nn = 0;
            if (i0 < r) {
                c128 ai = a[i0];
                ai.re -= vf.re * wr - vf.im * wi;
                ai.im -= vf.re * wi + vf.im * wr;
                a[i0] = ai;
                if (tl > 0) nn += ai.re * ai.re + ai.im * ai.im;  // row k is R's, not trailing
            }
            for (int i = i0 + kT; i < r; i += kT) {
                const c128 vi = v[i];
                c128 ai = a[i];
                ai.re -= vi.re * wr - vi.im * wi;
                ai.im -= vi.re * wi + vi.im * wr;
                a[i] = ai;
                nn += ai.re * ai.re + ai.im * ai.im;
            }
            nn = team_sum_t<kT>(nn);
            if (tl == 0) norm_next[j] = nn;
        }
        __syncthreads();
        cur ^= 1;
    }
    const int* phys = s_phys + cur * r;
    if (tid == 0 && r > 0) A[(size_t)phys[r - 1] * ld + r - 1] = v0_prev;
    // reflector bodies (rows below k) to global
    if (Wg) {
        for (int idx = tid; idx < r * r; idx += nt) {
            const int i = idx % r, k = idx / r;
            if (i > k) Wg[idx] = A[(size_t)phys[k] * ld + i];
        }
    }
    return cur;
}

// Z <- Q Z = H_0 H_1 ... H_{r-1} Z for the reflectors of qrcp_lds (v_k in Wg,
// tau_k in s_tau). Columns of Z are independent: a team applies every
// reflector to its columns with team reductions only.
template <int kT>
__device__ __forceinline__ void apply_q_left(c128* Z, int r, int ld, const c128* __restrict__ Wg,
                                             const double* s_tau) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const int nteams = nt / kT, team = tid / kT, tl = tid - team * kT;
    for (int j = team; j < r; j += nteams) {
        c128* z = Z + (size_t)j * ld;
        for (int k = r - 1; k >= 0; --k) {
            const c128* v = Wg + (size_t)k * r;
            const double tau = s_tau[k];
            double sr = 0, si = 0;
            for (int i = k + tl; i < r; i += kT) {  // s = v^H z
                const c128 vi = v[i], zi = z[i];
                sr += vi.re * zi.re + vi.im * zi.im;
                si += vi.re * zi.im - vi.im * zi.re;
            }
            sr = team_sum_t<kT>(sr) * tau;
            si = team_sum_t<kT>(si) * tau;
            for (int i = k + tl; i < r; i += kT) {
                const c128 vi = v[i];
                c128 zi = z[i];
                zi.re -= vi.re * sr - vi.im * si;
                zi.im -= vi.re * si + vi.im * sr;
                z[i] = zi;
            }
        }
    }
}

// Q_k = H_0 H_1 ... H_{k-1} [I_k; 0] in place of the reflectors of qrcp_lds
// (v_j in A's physical column phys[j], rows j..r-1; tau_j in s_tau): on exit
// column phys[j] holds Q e_j, all r rows. Step i applies H_i to the columns
// j > i, which hold H_{i+1} ... H_{k-1} e_j (zero above row i + 1); column
// i + 1 is set to H_{i+1} e_{i+1} in the same step by its own team, after
// step i + 1 has read v_{i+1} for the last time. One barrier per step; a team
// per column (the team width must leave a team for each of the k - 1 columns
// to run the steps without a serial loop, but any width is correct).
template <int kT>
__device__ __forceinline__ void form_q_inplace(c128* A, int r, int ld, int k, const int* phys, const double* s_tau) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const int nteams = nt / kT, team = tid / kT, tl = tid - team * kT;
    // z = H_j e_j in column phys[j]: rows < j zero, row j 1 - tau |v0|^2,
    // rows > j -tau conj(v0) v_j[l] (v0 = v_j[j]); each lane rewrites its own rows
    auto init_col = [&](int j) {
        c128* z = A + (size_t)phys[j] * ld;
        const double tj = s_tau[j];
        const c128 v0 = z[j];
        __builtin_amdgcn_wave_barrier();  // every lane of the team has v0 before row j is rewritten
        const double fr = -tj * v0.re, fi = tj * v0.im;  // -tau conj(v0)
        for (int l = tl; l < r; l += kT) {
            c128 x{0.0, 0.0};
            if (l == j) {
                x = c128{1.0 - tj * (v0.re * v0.re + v0.im * v0.im), 0.0};
            } else if (l > j) {
                const c128 vl = z[l];
                x = c128{fr * vl.re - fi * vl.im, fr * vl.im + fi * vl.re};
            }
            z[l] = x;
        }
    };
    for (int i = k - 2; i >= 0; --i) {
        const c128* v = A + (size_t)phys[i] * ld;
        const double tau = s_tau[i];
        for (int j = i + 1 + team; j < k; j += nteams) {
            if (j == i + 1) init_col(j);
            c128* z = A + (size_t)phys[j] * ld;
            double sr = 0, si = 0;
            for (int l = i + tl; l < r; l += kT) {  // s = v^H z
                const c128 vl = v[l], zl = z[l];
                sr += vl.re * zl.re + vl.im * zl.im;
                si += vl.re * zl.im - vl.im * zl.re;
            }
            sr = team_sum_t<kT>(sr) * tau;
            si = team_sum_t<kT>(si) * tau;
            for (int l = i + tl; l < r; l += kT) {
                const c128 vl = v[l];
                c128 zl = z[l];
                zl.re -= vl.re * sr - vl.im * si;
                zl.im -= vl.re * si + vl.im * sr;
                z[l] = zl;
            }
        }
        __syncthreads();
    }
    if (k > 0 && team == 0) init_col(0);
}

// Truncated QR pseudo-inverse of W (in A, LDS; destroyed): W P = Q R with
// column pivoting, k = the leading |R_jj| >= tau = max(atol, rtol |R_00|),
// and U = P [R_11^{-1} Q_k^H; 0]. For k = r this is W^{-1}; with truncation
// C U R interpolates the pivot columns (an interpolative decomposition)
// instead of projecting on W's leading singular space. Uses U (ldu) as
// scratch for R_11; Q_k is formed in place of the reflectors. sigma[j] =
// |R_jj| (logical order), smax = |R_00|, rank_eps counts |R_jj| >= label_eps |R_00|.
template <int kT>
__device__ __forceinline__ void pinv_qr_trunc(c128* A, int r, int ld, double atol, double rtol,
                                              c128* __restrict__ U, int ldu, double* __restrict__ sigma,
                                              int* s_phys, double* s_norm, double* s_tau, c128* s_alpha,
                                              int* kept_out, double* smax_out, double label_eps, int* rank_eps_out) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const int cur = qrcp_lds<kT>(A, r, ld, s_phys, s_norm, s_tau, s_alpha, nullptr);
    const int* phys = s_phys + cur * r;
    __syncthreads();
    const double smax = sqrt(s_alpha[0].re * s_alpha[0].re + s_alpha[0].im * s_alpha[0].im);
    const double tau = fmax(atol, rtol * smax);
    int k = 0, reps = 0;
    for (int j = 0; j < r; ++j) {
        const double a = sqrt(s_alpha[j].re * s_alpha[j].re + s_alpha[j].im * s_alpha[j].im);
        if (k == j && a >= tau && a > 0) k = j + 1;
        reps += (a >= label_eps * smax && smax > 0);
    }
    for (int j = tid; j < r; j += nt)
        sigma[j] = sqrt(s_alpha[j].re * s_alpha[j].re + s_alpha[j].im * s_alpha[j].im);
    // R_11's strict upper triangle (logical order) to the U scratch
    for (int idx = tid; idx < r * r; idx += nt) {
        const int i = idx % r, j = idx / r;
        if (i < j && j < k) U[(size_t)j * ldu + i] = A[(size_t)phys[j] * ld + i];
    }
    __syncthreads();
    // 1 / R_jj in place of R_jj
    for (int j = tid; j < k; j += nt) {
        const c128 a = s_alpha[j];
        const double ia = rcp_nr(a.re * a.re + a.im * a.im);
        s_alpha[j] = c128{a.re * ia, -a.im * ia};
    }
    __syncthreads();  // A's R is consumed; Q_k replaces the reflectors
    // A[phys[j] * ld + :] = Q e_j, with teams narrow enough to give every
    // column its own
    if constexpr (kT >= 8) {
        if (blockDim.x / kT < k && blockDim.x / (kT / 2) >= k)
            form_q_inplace<kT / 2>(A, r, ld, k, phys, s_tau);
        else if (blockDim.x / kT < k)
            form_q_inplace<kT / 4>(A, r, ld, k, phys, s_tau);
        else
            form_q_inplace<kT>(A, r, ld, k, phys, s_tau);
    } else if constexpr (kT >= 4) {
        if (blockDim.x / kT < k)
            form_q_inplace<kT / 2>(A, r, ld, k, phys, s_tau);
        else
            form_q_inplace<kT>(A, r, ld, k, phys, s_tau);
    } else {
        form_q_inplace<kT>(A, r, ld, k, phys, s_tau);
    }
    __syncthreads();
    // row c of Q_k: R_11 u = conj(Q[c, 0:k])^T by back substitution, in place;
    // four lanes per row share each step's dot product (quad sums), the
    // first of them writes u_l
    for (int c0 = tid >> 2; c0 < r; c0 += nt >> 2) {  // a quad shares c0: its DPP lanes are all active
        const int c = c0, sub = tid & 3;
        const bool live = c < r;
        for (int l = k - 1; l >= 0; --l) {
            double xr = 0, xi = 0;
            if (live) {
                for (int s2 = l + 1 + sub; s2 < k; s2 += 4) {
                    const c128 rv = U[(size_t)s2 * ldu + l], u = A[(size_t)phys[s2] * ld + c];
                    xr += rv.re * u.re - rv.im * u.im;
                    xi += rv.re * u.im + rv.im * u.re;
                }
            }
            xr = team_sum_t<4>(xr);
            xi = team_sum_t<4>(xi);
            if (live && sub == 0) {
                const c128 q = A[(size_t)phys[l] * ld + c];
                const double yr = q.re - xr, yi = -q.im - xi;
                const c128 ia = s_alpha[l];
                A[(size_t)phys[l] * ld + c] = c128{yr * ia.re - yi * ia.im, yr * ia.im + yi * ia.re};
            }
        }
    }
    __syncthreads();
    for (int idx = tid; idx < r * r; idx += nt) {
        const int l = idx % r, c = idx / r;
        U[(size_t)c * ldu + phys[l]] = l < k ? A[(size_t)phys[l] * ld + c] : c128{0.0, 0.0};
    }
    // every thread computed k, smax and reps alike (the outputs are per thread)
    *kept_out = k;
    *smax_out = smax;
    *rank_eps_out = reps;
    __syncthreads();
}

// Pinv of W (in A, LDS; destroyed) preconditioned by QR with column pivoting:
// W P = Q R, then one-sided Jacobi on X = R^H (X V = X_f, orthogonal columns
// x_j = sigma_j u_j). X = R^H of a column-pivoted R has nearly orthogonal
// columns, so the sweeps drop to about 5 from about 13 for the graded core
// matrices of far-field blocks. Two ways to finish:
//  - with V (accumulated rotations): R = V Sigma U^H gives
//      W^+ = P X_f diag(w) (Q V)^H,  w_j = [sigma_j >= tau] / sigma_j^2.
//    The weak components keep the relative accuracy of the Jacobi; needed
//    when tau may sit many orders below sigma_max (absolute thresholds).
//  - without V (vfree): W^H W = P X X^H P^T = P X_f X_f^H P^T, so
//      W^+ = (W^H W)^+ W^H = P X_f diag(w) Y^H,  Y = W P X_f,
//      w_j = [sigma_j >= tau] / sigma_j^4.
//    No rotations are accumulated and Q is not needed, but Y = W P x_j,
//    of norm sigma_j^2, forms from entries of size sigma_max sigma_j: the
//    weak components carry a relative error ~ eps sigma_max / sigma_j. The
//    caller uses this form only when tau >= 1e-8 sigma_max (rtol >= 1e-8).
// A, B: r x r LDS buffers with leading dimension ld; Wg: the global W panel
// (receives the reflectors, leading dimension r; untouched when vfree);
// wload(k, c): element (k, c) of the original W, used when vfree. B = NULL (vfree only, W not read from Wg): one LDS buffer; X = R^H
// is staged through U (global) and Y goes to Wg, so the launch needs half
// the LDS and fits twice the workgroups per CU. Outputs as jacobi_pinv_t.
template <int kT, int kE, typename WLoad>
__device__ __forceinline__ void pinv_qr_lds(c128* A, c128* B, int r, int ld, double atol, double rtol, bool vfree,
                                            const WLoad& wload, c128* __restrict__ U, int ldu,
                                            double* __restrict__ sigma, c128* __restrict__ Wg, int* s_flag,
                                            double* s_w, int* s_phys, double* s_norm, double* s_tau,
                                            c128* s_alpha, int* kept_out, double* smax_out, double label_eps,
                                            int* rank_eps_out, int* sweeps_out) {
    const int cur = qrcp_lds<kT>(A, r, ld, s_phys, s_norm, s_tau, s_alpha, vfree ? nullptr : Wg);
    const int* phys = s_phys + cur * r;
    __syncthreads();
    // X = R^H into B (X[i][j] = conj(R[j][i]) for i >= j); with one buffer
    // through U into A
    const bool one_buf = B == nullptr;
    c128* Xs = one_buf ? U : B;
    const int ldx = one_buf ? ldu : ld;
    for (int idx = threadIdx.x; idx < r * r; idx += blockDim.x) {
        const int i = idx % r, j = idx / r;
        c128 x{0.0, 0.0};
        if (i == j) {
            x = c128{s_alpha[j].re, -s_alpha[j].im};
        } else if (i > j) {
            const c128 rji = A[(size_t)phys[i] * ld + j];
            x = c128{rji.re, -rji.im};
        }
        Xs[(size_t)j * ldx + i] = x;
    }
    __syncthreads();
    if (one_buf) {
        for (int idx = threadIdx.x; idx < r * r; idx += blockDim.x) {
            const int i = idx % r, j = idx / r;
            A[(size_t)j * ld + i] = U[(size_t)j * ldu + i];
        }
        B = A;  // X lives in A from here on
        __syncthreads();
    }
    if (vfree) {
        *sweeps_out = jacobi_sweeps<kT, kE>(B, nullptr, r, ld, s_flag);
        jacobi_weights<kT>(B, r, ld, atol, rtol, label_eps, sigma, s_w, kept_out, smax_out, rank_eps_out, true);
        // Y[k][j] = sum_l W[k][phys[l]] X_f[l][j] into A (free since X was
        // formed) or, with one buffer, into Wg (ld r; W is not read from it)
        c128* Y = one_buf ? Wg : A;
        const int ldy = one_buf ? r : ld;
        for (int idx = threadIdx.x; idx < r * r; idx += blockDim.x) {
            const int k = idx % r, j = idx / r;
            const c128* xj = B + (size_t)j * ld;
            double yr = 0, yi = 0;
            if (s_w[j] != 0.0) {
#pragma unroll 4
                for (int l = 0; l < r; ++l) {
                    const c128 w = wload(k, phys[l]), x = xj[l];
                    yr += w.re * x.re - w.im * x.im;
                    yi += w.re * x.im + w.im * x.re;
                }
            }
            Y[(size_t)j * ldy + k] = c128{yr, yi};
        }
        __syncthreads();
        pinv_form(B, Y, s_w, r, ld, ldy, phys, U, ldu);
        __syncthreads();
        return;
    } else {
        for (int idx = threadIdx.x; idx < r * r; idx += blockDim.x) {
            const int i = idx % r, j = idx / r;
            A[(size_t)j * ld + i] = c128{i == j ? 1.0 : 0.0, 0.0};
        }
        *sweeps_out = jacobi_sweeps<kT, kE>(B, A, r, ld, s_flag);
        jacobi_weights<kT>(B, r, ld, atol, rtol, label_eps, sigma, s_w, kept_out, smax_out, rank_eps_out);
        apply_q_left<kT>(A, r, ld, Wg, s_tau);
    }
    __syncthreads();
    pinv_form(B, A, s_w, r, ld, ld, phys, U, ldu);
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
