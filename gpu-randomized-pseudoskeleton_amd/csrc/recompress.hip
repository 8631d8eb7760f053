// SVD recompression of compressed blocks C U R to their relative-threshold
// rank (svd_compress.m:19-34, called through postcompress_wrapper.m:1-5 and
// compres_single_block_spheres_for_lambda_sweep.jl:568):
//   C = Q_c R_c, R^H = Q_r R_r (Householder QR),
//   S = R_c U R_r^H = u diag(s) v^H (one-sided Jacobi on S with the rotations
//   accumulated), kept: s_j / s_1 > thresh,
//   C' = Q_c u_k, U' = diag(s_k), R' = (Q_r v_k)^H.
// One workgroup per block, FP64 throughout; the tall factors live in an HBM
// workspace, the r x r matrices too (r reaches 64, three r x r complex
// matrices exceed the LDS). Not on the compression path: it runs once per
// compress, when asked for.
#include "jacobi.hpp"
#include "kernels.hpp"

namespace farcur {

namespace {

constexpr int kRcThreads = 256;
constexpr int kRcWaves = kRcThreads / kWave;
constexpr int kRcMaxR = kRecompressMaxR;  // LDS arrays below are sized for this bound

__device__ __forceinline__ double wave_sum(double v) { return team_sum(v, kWave); }

__device__ __forceinline__ double abs2(c128 a) { return a.re * a.re + a.im * a.im; }

// Householder QR of A (m x r, ld lda, m >= r) in place: reflector j is
// H_j = I - scale_j v_j v_j^H with v_j = A[j:m, j] after the call, R's
// strictly upper part is A[i, c] for i < c and R_jj = beta[j].
__device__ void householder_qr(c128* A, int m, int r, int lda, c128* s_beta, double* s_scale, double* s_red) {
    const int tid = threadIdx.x, lane = tid & (kWave - 1), w = tid / kWave;
    for (int j = 0; j < r; ++j) {
        c128* x = A + (size_t)j * lda;
        double nn = 0;
        for (int i = j + tid; i < m; i += blockDim.x) nn += abs2(x[i]);
        nn = block_sum(nn, s_red);  // synchronises
        const c128 a = x[j];
        const double xn = sqrt(nn), aa = sqrt(abs2(a));
        c128 beta{0.0, 0.0}, v0 = a;
        double scale = 0.0;
        if (xn > 0) {
            const c128 ph = aa > 0 ? c128{a.re / aa, a.im / aa} : c128{1.0, 0.0};
            beta = c128{-ph.re * xn, -ph.im * xn};
            v0 = c128{a.re + ph.re * xn, a.im + ph.im * xn};
            scale = 1.0 / (xn * (xn + aa));  // 2 / (v^H v)
        }
        __syncthreads();
        if (tid == 0) {
            x[j] = v0;
            s_beta[j] = beta;
            s_scale[j] = scale;
        }
        __syncthreads();
        // trailing columns, one wave per column
        for (int c = j + 1 + w; c < r; c += kRcWaves) {
            c128* y = A + (size_t)c * lda;
            double dr = 0, di = 0;
            for (int i = j + lane; i < m; i += kWave) {
                const c128 v = x[i], t = y[i];
                dr += v.re * t.re + v.im * t.im;  // conj(v) t
                di += v.re * t.im - v.im * t.re;
            }
            dr = wave_sum(dr) * scale;
            di = wave_sum(di) * scale;
            for (int i = j + lane; i < m; i += kWave) {
                const c128 v = x[i];
                y[i].re -= v.re * dr - v.im * di;
                y[i].im -= v.re * di + v.im * dr;
            }
        }
        __syncthreads();
    }
}

// M (mm x k, ld ldm) = H_0 H_1 ... H_{r-1} M for the reflectors of
// householder_qr stored in A (ld lda).
__device__ void apply_q(const c128* A, int mm, int r, int lda, const double* s_scale, c128* M, int k, int ldm) {
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    for (int j = r - 1; j >= 0; --j) {
        const c128* x = A + (size_t)j * lda;
        const double scale = s_scale[j];
        if (scale != 0.0) {
            for (int c = w; c < k; c += kRcWaves) {
                c128* y = M + (size_t)c * ldm;
                double dr = 0, di = 0;
                for (int i = j + lane; i < mm; i += kWave) {
                    const c128 v = x[i], t = y[i];
                    dr += v.re * t.re + v.im * t.im;
                    di += v.re * t.im - v.im * t.re;
                }
                dr = wave_sum(dr) * scale;
                di = wave_sum(di) * scale;
                for (int i = j + lane; i < mm; i += kWave) {
                    const c128 v = x[i];
                    y[i].re -= v.re * dr - v.im * di;
                    y[i].im -= v.re * di + v.im * dr;
                }
            }
        }
        __syncthreads();
    }
}

// kMaxR: the LDS bound of the launch; jobs with r above it are left to the
// launch with the larger bound (kLo: the smaller bound below which jobs belong
// to the other launch). kE: rows per lane cached by the Jacobi (0: streamed).
template <typename PC, int kMaxR, int kLo, int kE>
__global__ __launch_bounds__(kRcThreads) void k_recompress(const RecompressJob* __restrict__ jobs, double thresh) {
    __shared__ c128 s_beta_c[kMaxR], s_beta_r[kMaxR];
    __shared__ double s_scale_c[kMaxR], s_scale_r[kMaxR], s_sig[kMaxR];
    __shared__ int s_ord[kMaxR];
    __shared__ double s_red[kRcWaves];
    __shared__ int s_flag;
    const RecompressJob J = jobs[blockIdx.x];
    const int m = J.m, n = J.n, r = J.r, tid = threadIdx.x, nt = blockDim.x;
    if (r > kMaxR || r <= kLo) return;  // the other launch's job
    if (r <= 0) {
        if (tid == 0) *J.k_out = 0;
        return;
    }
    const PC* C = (const PC*)J.C;
    const PC* R = (const PC*)J.R;
    // workspace: Q_c (m x r), Q_r (n x r), S, V (r x r), M (max(m, n) x r)
    c128* Qc = J.ws;
    c128* Qr = Qc + (size_t)m * r;
    c128* S = Qr + (size_t)n * r;
    c128* V = S + (size_t)r * r;
    c128* M = V + (size_t)r * r;
    for (size_t i = tid; i < (size_t)m * r; i += nt) Qc[i] = to_c128(C[i]);
    for (size_t idx = tid; idx < (size_t)n * r; idx += nt) {
        const int i = (int)(idx % n), c = (int)(idx / n);  // R^H[i][c] = conj(R[c][i])
        const c128 v = to_c128(R[c + (size_t)i * r]);
        Qr[idx] = c128{v.re, -v.im};
    }
    __syncthreads();
    householder_qr(Qc, m, r, m, s_beta_c, s_scale_c, s_red);
    householder_qr(Qr, n, r, n, s_beta_r, s_scale_r, s_red);
    // M (r x r, ld r) = R_c U, then S = M R_r^H
    for (int idx = tid; idx < r * r; idx += nt) {
        const int i = idx % r, c = idx / r;
        double tr = 0, ti = 0;
        for (int l = i; l < r; ++l) {
            const c128 rc = l == i ? s_beta_c[i] : Qc[i + (size_t)l * m];
            const c128 u = J.U[l + (size_t)c * r];
            tr += rc.re * u.re - rc.im * u.im;
            ti += rc.re * u.im + rc.im * u.re;
        }
        M[idx] = c128{tr, ti};
    }
    __syncthreads();
    for (int idx = tid; idx < r * r; idx += nt) {
        const int i = idx % r, c = idx / r;
        double sr = 0, si = 0;
        for (int l = c; l < r; ++l) {  // R_r^H[l][c] = conj(R_r[c][l]), nonzero for l >= c
            const c128 rr = l == c ? s_beta_r[c] : Qr[c + (size_t)l * n];
            const c128 t = M[i + (size_t)l * r];
            sr += t.re * rr.re + t.im * rr.im;
            si += t.im * rr.re - t.re * rr.im;
        }
        S[idx] = c128{sr, si};
        V[idx] = c128{i == c ? 1.0 : 0.0, 0.0};
    }
    __syncthreads();
    // S V = X with orthogonal columns: S = X V^H, s_j = |x_j|. S has the rank
    // of C U R: pairs of columns below (tau / sqrt r)^2, tau = thresh x the
    // largest column norm (<= s_1), are not rotated; they are dropped by the
    // threshold whatever rotations among them do, and at rounding level they
    // would keep the sweeps going to the limit. thresh = 0 keeps every nonzero
    // value, so there every pair is rotated. The sweeps stop with couplings ~1e-10, enough for a pinv but not
    // for singular values of nearly equal pairs; a second call adds the sweep
    // that takes them to rounding (its first sweep meets the stop rule).
    double cmax2 = 0;
    for (int j = tid; j < r; j += nt) {
        double nn = 0;
        for (int i = 0; i < r; ++i) nn += abs2(S[i + (size_t)j * r]);
        cmax2 = fmax(cmax2, nn);
    }
    for (int o = kWave / 2; o > 0; o >>= 1) cmax2 = fmax(cmax2, __shfl_xor(cmax2, o));
    if ((tid & (kWave - 1)) == 0) s_red[tid / kWave] = cmax2;
    __syncthreads();
    for (int w = 0; w < kRcWaves; ++w) cmax2 = fmax(cmax2, s_red[w]);
    const double weak2 = thresh * thresh * cmax2 / r;
    jacobi_sweeps<16, kE, true>(S, V, r, r, &s_flag, weak2);
    jacobi_sweeps<16, kE, true>(S, V, r, r, &s_flag, weak2);
    __syncthreads();
    for (int j = tid; j < r; j += nt) {
        double nn = 0;
        for (int i = 0; i < r; ++i) nn += abs2(S[i + (size_t)j * r]);
        s_sig[j] = sqrt(nn);
    }
    __syncthreads();
    // descending order (ties by index) and the kept count
    double smax = 0;
    for (int j = 0; j < r; ++j) smax = fmax(smax, s_sig[j]);
    for (int j = tid; j < r; j += nt) {
        int rank = 0;
        for (int i = 0; i < r; ++i) rank += s_sig[i] > s_sig[j] || (s_sig[i] == s_sig[j] && i < j);
        s_ord[rank] = j;
    }
    __syncthreads();
    int k = 0;
    if (smax > 0)
        for (int j = 0; j < r; ++j) k += s_sig[j] / smax > thresh;
    if (tid == 0) *J.k_out = k;
    for (int c = tid; c < k; c += nt) J.sigma[c] = s_sig[s_ord[c]];
    if (k == 0) return;
    // C' = Q_c [X_k diag(1/s_k); 0]
    for (size_t idx = tid; idx < (size_t)m * k; idx += nt) {
        const int i = (int)(idx % m), c = (int)(idx / m);
        const int j = s_ord[c];
        const double inv = 1.0 / s_sig[j];
        const c128 x = i < r ? S[i + (size_t)j * r] : c128{0.0, 0.0};
        M[idx] = c128{x.re * inv, x.im * inv};
    }
    __syncthreads();
    apply_q(Qc, m, r, m, s_scale_c, M, k, m);
    PC* Co = (PC*)J.Cout;
    for (size_t idx = tid; idx < (size_t)m * k; idx += nt) store_c(Co + idx, M[idx].re, M[idx].im);
    __syncthreads();
    // R' = (Q_r [V_k; 0])^H, k x n with ld k
    for (size_t idx = tid; idx < (size_t)n * k; idx += nt) {
        const int i = (int)(idx % n), c = (int)(idx / n);
        M[idx] = i < r ? V[i + (size_t)s_ord[c] * r] : c128{0.0, 0.0};
    }
    __syncthreads();
    apply_q(Qr, n, r, n, s_scale_r, M, k, n);
    PC* Ro = (PC*)J.Rout;
    for (size_t idx = tid; idx < (size_t)n * k; idx += nt) {
        const int i = (int)(idx % n), c = (int)(idx / n);
        store_c(Ro + c + (size_t)i * k, M[idx].re, -M[idx].im);
    }
}

}  // namespace

size_t recompress_ws_elems(int m, int n, int r) {
    return (size_t)(m + n) * r + 2 * (size_t)r * r + (size_t)std::max(m, n) * r;
}

void launch_recompress(const RecompressJob* jobs, int njobs, double thresh, int f32, hipStream_t s) {
    if (njobs <= 0) return;
    // r <= 64 (the usual far-field core): small LDS arrays, Jacobi rows in
    // registers; larger r: the second launch (its workgroups for small-r jobs
    // return at once)
    if (f32) {
        hipLaunchKernelGGL((k_recompress<c64, 64, -1, 4>), dim3(njobs), dim3(kRcThreads), 0, s, jobs, thresh);
        hipLaunchKernelGGL((k_recompress<c64, kRcMaxR, 64, 0>), dim3(njobs), dim3(kRcThreads), 0, s, jobs, thresh);
    } else {
        hipLaunchKernelGGL((k_recompress<c128, 64, -1, 4>), dim3(njobs), dim3(kRcThreads), 0, s, jobs, thresh);
        hipLaunchKernelGGL((k_recompress<c128, kRcMaxR, 64, 0>), dim3(njobs), dim3(kRcThreads), 0, s, jobs, thresh);
    }
}

}  // namespace farcur
