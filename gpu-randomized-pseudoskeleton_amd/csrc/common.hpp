// Shared device/host definitions for the far-field CUR kernels (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace farcur {

struct c128 {
    double re, im;
};
struct c64 {
    float re, im;
};

template <typename C>
struct ComplexTraits;
template <>
struct ComplexTraits<c128> {
    using real = double;
};
template <>
struct ComplexTraits<c64> {
    using real = float;
};

__device__ __forceinline__ c128 to_c128(c128 a) { return a; }
__device__ __forceinline__ c128 to_c128(c64 a) { return c128{(double)a.re, (double)a.im}; }
__device__ __forceinline__ void store_c(c128* p, double re, double im) { *p = c128{re, im}; }
__device__ __forceinline__ void store_c(c64* p, double re, double im) { *p = c64{(float)re, (float)im}; }
__device__ __forceinline__ void store_c(c64* p, float re, float im) { *p = c64{re, im}; }
__device__ __forceinline__ void store_c(c128* p, float re, float im) { *p = c128{(double)re, (double)im}; }

constexpr int kWave = 64;

// Sum over `width` consecutive lanes (width a power of two <= 64).
__device__ __forceinline__ double team_sum(double v, int width) {
    for (int o = width >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, width);
    return v;
}

// Workgroup-wide sum of a double; `red` is LDS scratch of nthreads/64 doubles.
__device__ __forceinline__ double block_sum(double v, double* red) {
    v = team_sum(v, kWave);
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double s = 0;
    const int nw = blockDim.x / kWave;
    for (int i = 0; i < nw; ++i) s += red[i];
    return s;
}

}  // namespace farcur
