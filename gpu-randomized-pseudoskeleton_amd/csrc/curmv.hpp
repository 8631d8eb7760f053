// Workgroup-level CUR products: y = C (U (R v)) and the successive residual
// err = ||y_old - y|| / ||y||.
//
// Replaces the three cublasZgemm (n = 1) calls, the two cublasDznrm2 calls and
// kernel_substract at codes/GPU_implementation/seed/user_impedance_for_julia_cuda.cu:1928-1963
// (kernel_substract at CU:239-254). The denominator is the NEW vector, as in
// the reference (CU:1966).
#pragma once
#include "common.hpp"

namespace farcur {

// y[i] = sum_j A[i + j*lda] x[j] for i < M, j < N, computed by the whole
// workgroup. x and y may live in LDS or global memory; s_part is LDS of
// blockDim.x complex values. Threads are split into rows x column-groups so
// that a column of A is read contiguously.
template <typename AC>
__device__ __forceinline__ void wg_gemv(const AC* __restrict__ A, int lda, int M, int N, const c128* x, c128* y,
                                        c128* s_part) {
    const int tid = threadIdx.x, nt = blockDim.x;
    int rt = kWave;
    while (rt < M && rt < nt) rt *= 2;
    if (rt > nt) rt = nt;
    const int G = nt / rt;
    const int tr = tid % rt, g = tid / rt;
    for (int i0 = 0; i0 < M; i0 += rt) {
        const int i = i0 + tr;
        double ar = 0, ai = 0;
        if (i < M && g < G) {
            for (int j = g; j < N; j += G) {
                const c128 a = to_c128(A[(size_t)j * lda + i]);
                const c128 b = x[j];
                ar += a.re * b.re - a.im * b.im;
                ai += a.re * b.im + a.im * b.re;
            }
        }
        __syncthreads();
        s_part[tid] = c128{ar, ai};
        __syncthreads();
        if (g == 0 && i < M) {
            double sr = 0, si = 0;
            for (int gg = 0; gg < G; ++gg) {
                sr += s_part[gg * rt + tr].re;
                si += s_part[gg * rt + tr].im;
            }
            y[i] = c128{sr, si};
        }
    }
    __syncthreads();
}

// Full residual step for one block. s_t1, s_t2: r complex values (LDS or
// global), s_part: blockDim.x complex (LDS), s_red: blockDim.x/64 doubles (LDS).
// Returns err (identical in all threads).
template <typename PC>
__device__ __forceinline__ double cur_residual(const PC* __restrict__ C, int m, int r, const c128* __restrict__ U,
                                               const PC* __restrict__ R, int n, const c128* v, const c128* y_old,
                                               c128* y_new, c128* s_t1, c128* s_t2, c128* s_part, double* s_red) {
    wg_gemv<PC>(R, r, r, n, v, s_t1, s_part);
    wg_gemv<c128>(U, r, r, r, s_t1, s_t2, s_part);
    wg_gemv<PC>(C, m, m, r, s_t2, y_new, s_part);
    double nn = 0, dd = 0;
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
        const c128 a = y_new[i], b = y_old[i];
        nn += a.re * a.re + a.im * a.im;
        const double dr = b.re - a.re, di = b.im - a.im;
        dd += dr * dr + di * di;
    }
    nn = block_sum(nn, s_red);
    dd = block_sum(dd, s_red);
    return sqrt(dd) / sqrt(nn);
}

}  // namespace farcur
