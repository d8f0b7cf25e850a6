// Rank model (MLP) training and gradients on the GPU.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace iegpu {

// `steps` Adam steps of 64 samples (drawn with replacement) on the mean
// squared error of the standardized rank; model (kMlpLen floats, device)
// holds the initial parameters and normalization and receives the trained
// parameters; loss (device, 1 float) the mean loss of the last 64 steps.
void launch_mlp_train(const float* X, const float* y, int n, float* model, int steps, float lr,
                      uint32_t seed, float* loss, hipStream_t st);
// Gradient of the mean loss over all n samples with respect to the trained
// parameters (kMlpParams floats, blob order W1 .. b3) and the loss.
void launch_mlp_grad(const float* X, const float* y, int n, const float* model, float* grad,
                     float* loss, hipStream_t st);

}  // namespace iegpu
