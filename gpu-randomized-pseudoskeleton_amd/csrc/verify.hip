// Accuracy of every compressed block against its FP64 dense block, on the
// device. For block b with rows (test functions) and cols, panels C, U, R:
//   ||Z_b - C U R||_F^2 and ||Z_b||_F^2 (the reference's error metric
//   norm(C*U*R - Z) / norm(Z), compres_single_block_spheres_for_lambda_sweep.jl:515-544),
//   (Z_b - C U R) v and Z_b v for v = V[cols, 0] (matvec error), and
//   Z_b V[cols, :] for up to kVerifyMaxRhs vectors (the FP64 dense product).
// Z_b's entries come from the FP64 element (entry_f64), the one the FP64
// panels use; C U R from P = C U formed in FP64 (k_verify_interp) times R.
// Nothing is stored per entry: one wave owns 64 rows of a block and walks its
// columns, each lane accumulating its row.
#include <hip/hip_runtime.h>

#include "element.hpp"
#include "verify.hpp"

namespace farcur {

namespace {

// P[i * r + k] = sum_l C[i + l m] U[l + k r], one thread per entry.
template <typename PC>
__global__ __launch_bounds__(256) void k_verify_interp(const VerifyInterp* __restrict__ jobs, int njobs) {
    for (int b = blockIdx.y; b < njobs; b += gridDim.y) {
        const VerifyInterp J = jobs[b];
        const PC* C = (const PC*)J.C;
        const int m = J.m, r = J.r;
        for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < m * r; idx += gridDim.x * blockDim.x) {
            const int i = idx / r, k = idx - i * r;
            const c128* u = J.U + (size_t)k * r;
            double pr = 0, pi = 0;
            for (int l = 0; l < r; ++l) {
                const c128 c = to_c128(C[(size_t)l * m + i]);
                pr += c.re * u[l].re - c.im * u[l].im;
                pi += c.re * u[l].im + c.im * u[l].re;
            }
            J.P[idx] = c128{pr, pi};
        }
    }
}

template <typename PC, int S>
__global__ __launch_bounds__(64) void k_verify(const VerifyJob* __restrict__ jobs, const int32_t* __restrict__ tile_job,
                                               const int32_t* __restrict__ tile_row0, const c128* __restrict__ V, int N,
                                               int s, const double* __restrict__ rec, double k, double c4k2,
                                               double scale) {
    const VerifyJob J = jobs[tile_job[blockIdx.x]];
    const int i = tile_row0[blockIdx.x] + (int)threadIdx.x;
    const bool live = i < J.m;
    const PC* R = (const PC*)J.R;
    const int r = J.r;
    RwgRec<double> a;
    load_rec(rec, J.rows[live ? i : 0], a);
    const c128* prow = J.P + (size_t)(live ? i : 0) * r;
    double num = 0, den = 0;
    double er = 0, ei = 0;  // (Z - CUR) v, v = V[:, 0] (Z v is column 0 of the product)
    double yr[S], yi[S];
#pragma unroll
    for (int q = 0; q < S; ++q) yr[q] = yi[q] = 0;
    for (int j = 0; j < J.n; ++j) {
        const int cj = J.cols[j];
        double zr, zi;
        entry_f64(a, rec + (size_t)cj * kRecStride, k, c4k2, scale, zr, zi);
#pragma unroll
        for (int q = 0; q < S; ++q)
            if (q < s) {
                const c128 v = V[(size_t)q * N + cj];
                yr[q] += zr * v.re - zi * v.im;
                yi[q] += zr * v.im + zi * v.re;
            }
        if (!J.ref_only) {
            double cr = 0, ci = 0;
            const PC* rc = R + (size_t)j * r;
            for (int l = 0; l < r; ++l) {
                const c128 p = prow[l], x = to_c128(rc[l]);
                cr += p.re * x.re - p.im * x.im;
                ci += p.re * x.im + p.im * x.re;
            }
            const double dr = zr - cr, di = zi - ci;
            num += dr * dr + di * di;
            den += zr * zr + zi * zi;
            const c128 v = V[cj];
            er += dr * v.re - di * v.im;
            ei += dr * v.im + di * v.re;
        }
    }
    if (!live) num = den = 0;
    if (live) {
        if (J.ys)
            for (int q = 0; q < s && q < S; ++q) J.ys[(size_t)q * J.m + i] = c128{yr[q], yi[q]};
        if (!J.ref_only && J.yerr) {
            J.yerr[i] = c128{er, ei};
            J.yref[i] = c128{yr[0], yi[0]};
        }
    }
    if (!J.ref_only) {
        num = team_sum(num, kWave);
        den = team_sum(den, kWave);
        if (threadIdx.x == 0) {
            atomicAdd(&J.sums[0], num);
            atomicAdd(&J.sums[1], den);
        }
    }
}

template <typename PC>
void launch_verify_t(const VerifyJob* jobs, int ntiles, const int32_t* tile_job, const int32_t* tile_row0,
                     const c128* V, int N, int s, const FillParams& fp, hipStream_t st) {
    const FillConsts c = fill_consts_f64(fp.k, fp.eta);
    const double* rec = (const double*)fp.rec;
    if (s <= 1)
        hipLaunchKernelGGL((k_verify<PC, 1>), dim3(ntiles), dim3(64), 0, st, jobs, tile_job, tile_row0, V, N, s, rec,
                           c.k, c.c4k2, c.scale);
    else if (s <= 8)
        hipLaunchKernelGGL((k_verify<PC, 8>), dim3(ntiles), dim3(64), 0, st, jobs, tile_job, tile_row0, V, N, s, rec,
                           c.k, c.c4k2, c.scale);
    else
        hipLaunchKernelGGL((k_verify<PC, kVerifyMaxRhs>), dim3(ntiles), dim3(64), 0, st, jobs, tile_job, tile_row0,
                           V, N, s, rec, c.k, c.c4k2, c.scale);
}

}  // namespace

void launch_verify_interp(const VerifyInterp* jobs, int njobs, int f32, hipStream_t s) {
    if (njobs <= 0) return;
    const dim3 grid(16, njobs < 65535 ? njobs : 65535);
    if (f32)
        hipLaunchKernelGGL(k_verify_interp<c64>, grid, dim3(256), 0, s, jobs, njobs);
    else
        hipLaunchKernelGGL(k_verify_interp<c128>, grid, dim3(256), 0, s, jobs, njobs);
}

void launch_verify(const VerifyJob* jobs, int ntiles, const int32_t* tile_job, const int32_t* tile_row0, int f32,
                   const c128* V, int N, int s, const FillParams& fp, hipStream_t st) {
    if (ntiles <= 0) return;
    if (f32)
        launch_verify_t<c64>(jobs, ntiles, tile_job, tile_row0, V, N, s, fp, st);
    else
        launch_verify_t<c128>(jobs, ntiles, tile_job, tile_row0, V, N, s, fp, st);
}

}  // namespace farcur
