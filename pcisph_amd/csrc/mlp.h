// mlp.h - the warm-start network: 8 inputs -> 64 SiLU -> 64 SiLU -> 1, BF16 weights
// and activations on the matrix cores (v_mfma_f32_16x16x16_bf16), FP32 accumulation,
// FP32 master weights for training.  No counterpart in the reference.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pcisph {

constexpr int kMlpIn = 8;
constexpr int kMlpHidden = 64;
// parameter vector layout (FP32): W1[64][8], b1[64], W2[64][64], b2[64], W3[64], b3
constexpr int kMlpW1 = 0;
constexpr int kMlpB1 = kMlpW1 + kMlpHidden * kMlpIn;
constexpr int kMlpW2 = kMlpB1 + kMlpHidden;
constexpr int kMlpB2 = kMlpW2 + kMlpHidden * kMlpHidden;
constexpr int kMlpW3 = kMlpB2 + kMlpHidden;
constexpr int kMlpB3 = kMlpW3 + kMlpHidden;
constexpr int kMlpParams = kMlpB3 + 1;  // 4801

struct MlpLaunch {
    hipStream_t stream;
};

// y[i] = net(x[i]) for n rows of x (row-major float[n][8]).
void mlp_forward(const MlpLaunch& L, const float* params, const float* x, float* y, uint32_t n);
// Sum over rows of the gradient of 0.5*(net(x)-t)^2 / n into grad[kMlpParams]
// (zeroed here) and the loss into loss[0].
void mlp_gradients(const MlpLaunch& L, const float* params, const float* x, const float* t, uint32_t n,
                   float* grad, float* loss);
// One Adam step (FP32) on params with moments m, v; `step` counts from 1.
void mlp_adam(const MlpLaunch& L, float* params, const float* grad, float* m, float* v, uint32_t step, float lr,
              float beta1, float beta2, float eps);

// Warm start: y[i] = net(feat[i]), the change of the particle's pressure from the
// previous step's converged value, in units of the scene's p_ref, for the rows i of
// the ranges in `rows` (device memory: {count, tiles, then per range {begin, end,
// first tile}}, ranges non-empty and increasing, first tile = the sum of
// ceil((end - begin) / 16) over the ranges before it, tiles = that sum over all).
// The launch is sized for `max_rows` rows, at least the ranges' total.
template <typename R>
void mlp_warm_start(const MlpLaunch& L, const float* params, const float* feat, const uint32_t* rows,
                    uint32_t max_rows, R* y);
// Appends (feat[i], p[i] / p_ref - feat[i][2]) -- the converged pressure's change from
// the previous step's, whose p / p_ref is input 2 -- for i = phase, phase + stride, ...
// to the sample store at slots base + i / stride (those below cap).
template <typename R>
void mlp_collect(const MlpLaunch& L, const float* feat, const R* p, uint32_t n, uint32_t stride, uint32_t phase,
                 float p_ref, float* store_x, float* store_t, uint64_t base, uint64_t cap);
// batch rows drawn from the store by a counter-based hash of (seed, step, row).
void mlp_gather_batch(const MlpLaunch& L, const float* store_x, const float* store_t, uint64_t count,
                      uint64_t seed, uint32_t step, uint32_t batch, float* x, float* t);

}  // namespace pcisph
