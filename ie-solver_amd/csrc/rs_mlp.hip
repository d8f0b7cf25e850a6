// Rank model: forward and backward passes of the MLP kFeat -> kHid (tanh)
// -> kHid (tanh) -> 1 over batches of 64 samples, one sample per lane,
// with the batch's activations and deltas in LDS for the weight-gradient
// sums; Adam (Kingma & Ba 2015) updates the parameters in LDS.  One
// workgroup does the whole training (the network is 1.4k parameters).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "rs_device.hpp"
#include "rs_mlp.hpp"

namespace iegpu {

namespace {

constexpr int kB = 64;        // samples per step, one per lane
constexpr int kLd = kHid + 1;  // padded LDS rows

struct MlpLds {
  float w[kMlpLen];
  float h1[kB][kLd], h2[kB][kLd], d1[kB][kLd], d2[kB][kLd];
  float x[kB][kFeat];
  float dout[kB];
  float red[kB];
  float g[kMlpParams];       // gradient sums of the step (or of all batches)
  float m1[kMlpParams], m2[kMlpParams];  // Adam moments
};

__device__ inline uint32_t hash32(uint32_t a) {
  a ^= a >> 16;
  a *= 0x7FEB352Du;
  a ^= a >> 15;
  a *= 0x846CA68Bu;
  a ^= a >> 16;
  return a;
}

// Forward and backward pass of sample idx (idx < 0: none) for this thread,
// then the batch's gradient sums added to L.g (thread t owns parameters
// t, t + 64, ...); returns the batch's sum of squared errors (in all
// threads).  scale multiplies d loss / d out.
__device__ float mlp_batch(MlpLds& L, const float* X, const float* Y, int idx, float scale) {
  const int tid = threadIdx.x;
  const float* w = L.w;
  // Activations and deltas live in this thread's LDS rows (no register
  // arrays indexed at run time).
  float* x = L.x[tid];
  float* h1 = L.h1[tid];
  float* h2 = L.h2[tid];
  float* d1 = L.d1[tid];
  float* d2 = L.d2[tid];
  for (int q = 0; q < kFeat; q++)
    x[q] = idx >= 0 ? (X[(int64_t)idx * kFeat + q] - w[kMlpMu + q]) / w[kMlpSd + q] : 0.0f;
  for (int j = 0; j < kHid; j++) {
    float z = w[kMlpB1 + j];
    for (int q = 0; q < kFeat; q++) z += w[kMlpW1 + q * kHid + j] * x[q];
    h1[j] = tanhf(z);
  }
  float out = w[kMlpB3];
  for (int j = 0; j < kHid; j++) {
    float z = w[kMlpB2 + j];
    for (int i = 0; i < kHid; i++) z += w[kMlpW2 + i * kHid + j] * h1[i];
    const float h = tanhf(z);
    h2[j] = h;
    out += w[kMlpW3 + j] * h;
  }
  const float e = idx >= 0 ? out - (Y[idx] - w[kMlpYMu]) / w[kMlpYSd] : 0.0f;
  const float dout = 2 * e * scale;
  for (int j = 0; j < kHid; j++) d2[j] = dout * w[kMlpW3 + j] * (1 - h2[j] * h2[j]);
  for (int i = 0; i < kHid; i++) {
    float dh = 0;
    for (int j = 0; j < kHid; j++) dh += w[kMlpW2 + i * kHid + j] * d2[j];
    d1[i] = dh * (1 - h1[i] * h1[i]);
  }
  L.dout[tid] = dout;
  L.red[tid] = e * e;
  __syncthreads();
  for (int p = tid; p < kMlpParams; p += kB) {
    const int o = p + kMlpW1;  // blob offset
    float sum = 0;
    if (o < kMlpB1) {
      const int q = (o - kMlpW1) / kHid, j = (o - kMlpW1) % kHid;
      for (int t = 0; t < kB; t++) sum += L.x[t][q] * L.d1[t][j];
    } else if (o < kMlpW2) {
      const int j = o - kMlpB1;
      for (int t = 0; t < kB; t++) sum += L.d1[t][j];
    } else if (o < kMlpB2) {
      const int i = (o - kMlpW2) / kHid, j = (o - kMlpW2) % kHid;
      for (int t = 0; t < kB; t++) sum += L.h1[t][i] * L.d2[t][j];
    } else if (o < kMlpW3) {
      const int j = o - kMlpB2;
      for (int t = 0; t < kB; t++) sum += L.d2[t][j];
    } else if (o < kMlpB3) {
      const int j = o - kMlpW3;
      for (int t = 0; t < kB; t++) sum += L.dout[t] * L.h2[t][j];
    } else {
      for (int t = 0; t < kB; t++) sum += L.dout[t];
    }
    L.g[p] += sum;
  }
  for (int o = kB / 2; o > 0; o >>= 1) {
    if (tid < o) L.red[tid] += L.red[tid + o];
    __syncthreads();
  }
  const float sse = L.red[0];
  __syncthreads();
  return sse;
}

__global__ __launch_bounds__(kB) void k_mlp_train(const float* X, const float* Y, int n,
                                                  float* model, int steps, float lr,
                                                  uint32_t seed, float* loss) {
  __shared__ MlpLds L;
  const int tid = threadIdx.x;
  for (int p = tid; p < kMlpLen; p += kB) L.w[p] = model[p];
  for (int p = tid; p < kMlpParams; p += kB) L.m1[p] = L.m2[p] = 0;
  const float b1 = 0.9f, b2 = 0.999f;
  float b1t = 1, b2t = 1, tail = 0;
  __syncthreads();
  for (int step = 0; step < steps; step++) {
    for (int p = tid; p < kMlpParams; p += kB) L.g[p] = 0;
    const int idx = (int)(hash32(seed ^ hash32((uint32_t)step * kB + tid)) % (uint32_t)n);
    const float sse = mlp_batch(L, X, Y, idx, 1.0f / kB);
    if (step >= steps - 64) tail += sse / kB;
    b1t *= b1;
    b2t *= b2;
    for (int p = tid; p < kMlpParams; p += kB) {
      const float gp = L.g[p];
      const float a = b1 * L.m1[p] + (1 - b1) * gp, v = b2 * L.m2[p] + (1 - b2) * gp * gp;
      L.m1[p] = a;
      L.m2[p] = v;
      L.w[kMlpW1 + p] -= lr * (a / (1 - b1t)) / (sqrtf(v / (1 - b2t)) + 1e-8f);
    }
    __syncthreads();
  }
  for (int p = tid; p < kMlpLen; p += kB) model[p] = L.w[p];
  if (tid == 0) *loss = tail / (float)(steps < 64 ? steps : 64);
}

__global__ __launch_bounds__(kB) void k_mlp_grad(const float* X, const float* Y, int n,
                                                 const float* model, float* grad, float* loss) {
  __shared__ MlpLds L;
  const int tid = threadIdx.x;
  for (int p = tid; p < kMlpLen; p += kB) L.w[p] = model[p];
  for (int p = tid; p < kMlpParams; p += kB) L.g[p] = 0;
  __syncthreads();
  float sse = 0;
  for (int b0 = 0; b0 < n; b0 += kB) {
    const int idx = b0 + tid < n ? b0 + tid : -1;
    sse += mlp_batch(L, X, Y, idx, 1.0f / (float)n);
  }
  for (int p = tid; p < kMlpParams; p += kB) grad[p] = L.g[p];
  if (tid == 0) *loss = sse / (float)n;
}

}  // namespace

void launch_mlp_train(const float* X, const float* y, int n, float* model, int steps, float lr,
                      uint32_t seed, float* loss, hipStream_t st) {
  k_mlp_train<<<1, kB, 0, st>>>(X, y, n, model, steps, lr, seed, loss);
}

void launch_mlp_grad(const float* X, const float* y, int n, const float* model, float* grad,
                     float* loss, hipStream_t st) {
  k_mlp_grad<<<1, kB, 0, st>>>(X, y, n, model, grad, loss);
}

}  // namespace iegpu
