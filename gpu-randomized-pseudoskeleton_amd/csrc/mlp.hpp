// Rank-predictor MLP (8 -> 64 -> 64 -> 1, GELU) launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace farcur {

constexpr int kMlpFeatures = 8;
constexpr int kMlpHidden = 64;
constexpr int kMlpParams = kMlpHidden * kMlpFeatures + kMlpHidden + kMlpHidden * kMlpHidden + kMlpHidden +
                           kMlpHidden + 1;  // 4801

// Input standardisation and output scaling: the network sees
// (x - xm) / xs and predicts (log r - ym) / ys.
struct MlpNorm {
    float xm[kMlpFeatures];
    float xs[kMlpFeatures];
    float ym, ys;
};

size_t mlp_tile_lds_bytes();
void launch_mlp_train_tiles(const float* params, const MlpNorm& nrm, const float* X, const float* y,
                            const int32_t* idx, int nbatch, float* partial, float* ploss, hipStream_t s);
void launch_mlp_reduce(const float* partial, const float* ploss, int nbatch, float* grad, float* loss,
                       hipStream_t s);
void launch_adam(float* params, const float* grad, float* m, float* v, int step, float lr, hipStream_t s);
void launch_mlp_predict(const float* params, const MlpNorm& nrm, const float* X, int n, float* y, hipStream_t s);
void launch_rank_predict(const float* params, const MlpNorm& nrm, const float* X, const int32_t* cap, int n, int p,
                         int32_t* r, hipStream_t s);

}  // namespace farcur
