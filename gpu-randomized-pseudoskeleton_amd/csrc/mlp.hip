// Rank predictor: MLP 8 -> 64 -> 64 -> 1 with GELU, trained and evaluated on
// the GPU by the kernels below (forward, backward, gradient reduction, Adam,
// inference). FP32 throughout; a training tile is 64 samples per workgroup
// with the activations in LDS.
//
// Parameter layout (kMlpParams = 4801 floats):
//   W1 [64][8] at 0, b1 [64] at 512, W2 [64][64] at 576 (W2[j][k] = weight
//   from hidden-1 unit k to hidden-2 unit j), b2 [64] at 4672, w3 [64] at
//   4736, b3 at 4800.
#include <hip/hip_runtime.h>

#include <math.h>

#include "mlp.hpp"

namespace farcur {

namespace {

constexpr int kF = kMlpFeatures, kH = kMlpHidden;
constexpr int oW1 = 0, oB1 = kH * kF, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH, oB3 = oW3 + kH;
static_assert(oB3 + 1 == kMlpParams, "parameter layout");
constexpr int kTile = 64;  // samples per workgroup
constexpr int kLd = kH + 1;

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

struct TileLds {
    float x[kTile][kF];
    float t[kTile];
    float dy[kTile];
    float p[kMlpParams];
    float z1[kTile][kLd];
    float h1[kTile][kLd];
    float z2[kTile][kLd];
    float h2[kTile][kLd];
};

// Loss and partial gradient of one tile of the batch. Samples
// idx[tile*64 + s] for s < 64 (those >= nbatch are masked).
__global__ __launch_bounds__(256) void k_mlp_train_tile(const float* __restrict__ params, MlpNorm nrm,
                                                        const float* __restrict__ X, const float* __restrict__ y,
                                                        const int32_t* __restrict__ idx, int nbatch,
                                                        float* __restrict__ partial, float* __restrict__ ploss) {
    extern __shared__ __align__(16) char smem[];
    TileLds& L = *reinterpret_cast<TileLds*>(smem);
    const int tid = threadIdx.x, nt = blockDim.x;
    const int base = blockIdx.x * kTile;
    for (int i = tid; i < kMlpParams; i += nt) L.p[i] = params[i];
    for (int i = tid; i < kTile * kF; i += nt) {
        const int s = i / kF, c = i % kF;
        const int g = base + s;
        L.x[s][c] = g < nbatch ? (X[(size_t)idx[g] * kF + c] - nrm.xm[c]) / nrm.xs[c] : 0.0f;
    }
    for (int s = tid; s < kTile; s += nt) {
        const int g = base + s;
        L.t[s] = g < nbatch ? (y[idx[g]] - nrm.ym) / nrm.ys : 0.0f;
    }
    __syncthreads();
    // hidden 1
    for (int i = tid; i < kTile * kH; i += nt) {
        const int s = i / kH, j = i % kH;
        float z = L.p[oB1 + j];
#pragma unroll
        for (int c = 0; c < kF; ++c) z += L.p[oW1 + j * kF + c] * L.x[s][c];
        L.z1[s][j] = z;
        L.h1[s][j] = gelu(z);
    }
    __syncthreads();
    // hidden 2
    for (int i = tid; i < kTile * kH; i += nt) {
        const int s = i / kH, j = i % kH;
        float z = L.p[oB2 + j];
        for (int k = 0; k < kH; ++k) z += L.p[oW2 + j * kH + k] * L.h1[s][k];
        L.z2[s][j] = z;
        L.h2[s][j] = gelu(z);
    }
    __syncthreads();
    // output, loss, dL/dy
    if (tid < kTile) {
        const int s = tid;
        float o = L.p[oB3];
        for (int k = 0; k < kH; ++k) o += L.p[oW3 + k] * L.h2[s][k];
        const bool valid = base + s < nbatch;
        const float d = o - L.t[s];
        L.dy[s] = valid ? 2.0f * d / (float)nbatch : 0.0f;
        L.t[s] = valid ? d * d : 0.0f;  // per-sample squared error
    }
    __syncthreads();
    float* g = partial + (size_t)blockIdx.x * kMlpParams;
    if (tid == 0) {
        float ls = 0, gb3 = 0;
        for (int s = 0; s < kTile; ++s) {
            ls += L.t[s];
            gb3 += L.dy[s];
        }
        ploss[blockIdx.x] = ls;
        g[oB3] = gb3;
    }
    for (int k = tid; k < kH; k += nt) {
        float a = 0;
        for (int s = 0; s < kTile; ++s) a += L.dy[s] * L.h2[s][k];
        g[oW3 + k] = a;
    }
    __syncthreads();
    // dZ2 (into z2)
    for (int i = tid; i < kTile * kH; i += nt) {
        const int s = i / kH, j = i % kH;
        L.z2[s][j] = L.dy[s] * L.p[oW3 + j] * gelu_d(L.z2[s][j]);
    }
    __syncthreads();
    // dW2, db2
    for (int i = tid; i < kH * kH; i += nt) {
        const int j = i / kH, k = i % kH;
        float a = 0;
        for (int s = 0; s < kTile; ++s) a += L.z2[s][j] * L.h1[s][k];
        g[oW2 + j * kH + k] = a;
    }
    for (int j = tid; j < kH; j += nt) {
        float a = 0;
        for (int s = 0; s < kTile; ++s) a += L.z2[s][j];
        g[oB2 + j] = a;
    }
    // dH1 (into h2), then dZ1 (into z1)
    for (int i = tid; i < kTile * kH; i += nt) {
        const int s = i / kH, k = i % kH;
        float a = 0;
        for (int j = 0; j < kH; ++j) a += L.z2[s][j] * L.p[oW2 + j * kH + k];
        L.h2[s][k] = a;
    }
    __syncthreads();
    for (int i = tid; i < kTile * kH; i += nt) {
        const int s = i / kH, k = i % kH;
        L.z1[s][k] = L.h2[s][k] * gelu_d(L.z1[s][k]);
    }
    __syncthreads();
    // dW1, db1
    for (int i = tid; i < kH * kF; i += nt) {
        const int k = i / kF, c = i % kF;
        float a = 0;
        for (int s = 0; s < kTile; ++s) a += L.z1[s][k] * L.x[s][c];
        g[oW1 + k * kF + c] = a;
    }
    for (int k = tid; k < kH; k += nt) {
        float a = 0;
        for (int s = 0; s < kTile; ++s) a += L.z1[s][k];
        g[oB1 + k] = a;
    }
}

// grad[p] = sum over tiles (fixed order); loss = sum of tile losses / nbatch.
__global__ void k_mlp_reduce(const float* __restrict__ partial, const float* __restrict__ ploss, int ntiles,
                             int nbatch, float* __restrict__ grad, float* __restrict__ loss) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < kMlpParams) {
        float a = 0;
        for (int t = 0; t < ntiles; ++t) a += partial[(size_t)t * kMlpParams + p];
        grad[p] = a;
    }
    if (p == 0) {
        float l = 0;
        for (int t = 0; t < ntiles; ++t) l += ploss[t];
        *loss = l / (float)nbatch;
    }
}

// Adam (Kingma & Ba, 2015) with bias correction.
__global__ void k_adam(float* __restrict__ params, const float* __restrict__ grad, float* __restrict__ m,
                       float* __restrict__ v, int step, float lr, float b1, float b2, float eps) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    const float gr = grad[p];
    const float mm = b1 * m[p] + (1.0f - b1) * gr;
    const float vv = b2 * v[p] + (1.0f - b2) * gr * gr;
    m[p] = mm;
    v[p] = vv;
    const float mh = mm / (1.0f - powf(b1, (float)step));
    const float vh = vv / (1.0f - powf(b2, (float)step));
    params[p] -= lr * mh / (sqrtf(vh) + eps);
}

__device__ float mlp_forward(const float* __restrict__ P, const MlpNorm& nrm, const float* __restrict__ x) {
    float xn[kF];
#pragma unroll
    for (int c = 0; c < kF; ++c) xn[c] = (x[c] - nrm.xm[c]) / nrm.xs[c];
    float h1[kH];
    for (int j = 0; j < kH; ++j) {
        float z = P[oB1 + j];
#pragma unroll
        for (int c = 0; c < kF; ++c) z += P[oW1 + j * kF + c] * xn[c];
        h1[j] = gelu(z);
    }
    float o = P[oB3];
    for (int j = 0; j < kH; ++j) {
        float z = P[oB2 + j];
        for (int k = 0; k < kH; ++k) z += P[oW2 + j * kH + k] * h1[k];
        o += P[oW3 + j] * gelu(z);
    }
    return o * nrm.ys + nrm.ym;
}

// Inference: y[i] = model(X[i]) (log rank). Parameters staged in LDS.
__global__ __launch_bounds__(256) void k_mlp_predict(const float* __restrict__ params, MlpNorm nrm,
                                                     const float* __restrict__ X, int n, float* __restrict__ y) {
    __shared__ float P[kMlpParams];
    for (int i = threadIdx.x; i < kMlpParams; i += blockDim.x) P[i] = params[i];
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = mlp_forward(P, nrm, X + (size_t)i * kF);
}

// Sample counts from the model: r = min(cap, max(1, ceil(exp(pred))) + p).
__global__ __launch_bounds__(256) void k_rank_predict(const float* __restrict__ params, MlpNorm nrm,
                                                      const float* __restrict__ X, const int32_t* __restrict__ cap,
                                                      int n, int p, int32_t* __restrict__ r) {
    __shared__ float P[kMlpParams];
    for (int i = threadIdx.x; i < kMlpParams; i += blockDim.x) P[i] = params[i];
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float lr = mlp_forward(P, nrm, X + (size_t)i * kF);
    const float k = ceilf(expf(fminf(lr, 20.0f)));
    int rr = (k < 1.0f ? 1 : (int)k) + p;
    if (rr > cap[i]) rr = cap[i];
    r[i] = rr;
}

}  // namespace

size_t mlp_tile_lds_bytes() { return sizeof(TileLds); }

void launch_mlp_train_tiles(const float* params, const MlpNorm& nrm, const float* X, const float* y,
                            const int32_t* idx, int nbatch, float* partial, float* ploss, hipStream_t s) {
    const int ntiles = (nbatch + kTile - 1) / kTile;
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute((const void*)k_mlp_train_tile, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)sizeof(TileLds));
        attr = true;
    }
    hipLaunchKernelGGL(k_mlp_train_tile, dim3(ntiles), dim3(256), sizeof(TileLds), s, params, nrm, X, y, idx, nbatch,
                       partial, ploss);
}

void launch_mlp_reduce(const float* partial, const float* ploss, int nbatch, float* grad, float* loss,
                       hipStream_t s) {
    const int ntiles = (nbatch + kTile - 1) / kTile;
    hipLaunchKernelGGL(k_mlp_reduce, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, partial, ploss, ntiles, nbatch,
                       grad, loss);
}

void launch_adam(float* params, const float* grad, float* m, float* v, int step, float lr, hipStream_t s) {
    hipLaunchKernelGGL(k_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, params, grad, m, v, step, lr, 0.9f,
                       0.999f, 1e-8f);
}

void launch_mlp_predict(const float* params, const MlpNorm& nrm, const float* X, int n, float* y, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_mlp_predict, dim3((n + 255) / 256), dim3(256), 0, s, params, nrm, X, n, y);
}

void launch_rank_predict(const float* params, const MlpNorm& nrm, const float* X, const int32_t* cap, int n, int p,
                         int32_t* r, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_rank_predict, dim3((n + 255) / 256), dim3(256), 0, s, params, nrm, X, cap, n, p, r);
}

}  // namespace farcur
