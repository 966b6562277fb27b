// Training kernels of the T5 DNN (SDL/NeuralNetwork.h:116-162 architecture:
// Linear(38,32) ReLU Linear(32,32) ReLU Linear(32,1), sigmoid applied by the
// loss).  Parameters are one packed float array in the order of the chain's
// parameter files: bias_0[32], wgtT_0[38][32], bias_2[32], wgtT_2[32][32],
// bias_4[1], wgtT_4[32][1].
//
// Forward and the activation gradients run one thread per sample in FP32
// with the inputs accumulated in index order, as the in-chain inference does.
// Weight gradients run one thread per parameter, summing over the batch in
// sample order, so a gradient does not depend on scheduling.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/lst_amd.h"

namespace {

constexpr int kIn = 38, kH = 32;
constexpr int kB0 = 0, kW0 = kB0 + kH, kB2 = kW0 + kIn * kH, kW2 = kB2 + kH, kB4 = kW2 + kH * kH, kW4 = kB4 + 1;
static_assert(kW4 + kH == LST_T5DNN_NPARAMS, "packed parameter layout");

__global__ __launch_bounds__(256) void k_mlp_forward(const float* __restrict__ x, const float* __restrict__ p,
                                                     float* __restrict__ h1, float* __restrict__ h2,
                                                     float* __restrict__ logit, uint32_t n) {
  for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < n; b += gridDim.x * blockDim.x) {
    float in[kIn], a1[kH];
#pragma unroll
    for (int i = 0; i < kIn; i++) in[i] = x[(size_t)kIn * b + i];
#pragma unroll
    for (int o = 0; o < kH; o++) {
      float acc = 0.f;
#pragma unroll
      for (int i = 0; i < kIn; i++) acc += in[i] * p[kW0 + kH * i + o];
      acc += p[kB0 + o];
      a1[o] = acc > 0.f ? acc : 0.f;
      h1[(size_t)kH * b + o] = a1[o];
    }
    float out = 0.f;
#pragma unroll
    for (int o = 0; o < kH; o++) {
      float acc = 0.f;
#pragma unroll
      for (int i = 0; i < kH; i++) acc += a1[i] * p[kW2 + kH * i + o];
      acc += p[kB2 + o];
      const float a2 = acc > 0.f ? acc : 0.f;
      h2[(size_t)kH * b + o] = a2;
      out += a2 * p[kW4 + o];
    }
    logit[b] = out + p[kB4];
  }
}

// dL/dh2 and dL/dh1 (pre-activation) per sample from dL/dlogit
__global__ __launch_bounds__(256) void k_mlp_backward_act(const float* __restrict__ p, const float* __restrict__ h1,
                                                          const float* __restrict__ h2, const float* __restrict__ dlogit,
                                                          float* __restrict__ d2, float* __restrict__ d1, uint32_t n) {
  for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < n; b += gridDim.x * blockDim.x) {
    const float g = dlogit[b];
    float g2[kH];
#pragma unroll
    for (int i = 0; i < kH; i++) {
      g2[i] = h2[(size_t)kH * b + i] > 0.f ? g * p[kW4 + i] : 0.f;
      d2[(size_t)kH * b + i] = g2[i];
    }
#pragma unroll
    for (int j = 0; j < kH; j++) {
      float acc = 0.f;
#pragma unroll
      for (int i = 0; i < kH; i++) acc += g2[i] * p[kW2 + kH * j + i];
      d1[(size_t)kH * b + j] = h1[(size_t)kH * b + j] > 0.f ? acc : 0.f;
    }
  }
}

// one thread per parameter: the batch sum of its gradient, in sample order
__global__ __launch_bounds__(256) void k_mlp_backward_w(const float* __restrict__ x, const float* __restrict__ h1,
                                                        const float* __restrict__ h2, const float* __restrict__ dlogit,
                                                        const float* __restrict__ d2, const float* __restrict__ d1,
                                                        float* __restrict__ grad, uint32_t n) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= LST_T5DNN_NPARAMS) return;
  float acc = 0.f;
  if (q < kW0) {  // bias_0[o]
    for (uint32_t b = 0; b < n; b++) acc += d1[(size_t)kH * b + (q - kB0)];
  } else if (q < kB2) {  // wgtT_0[i][o]
    const int i = (q - kW0) / kH, o = (q - kW0) % kH;
    for (uint32_t b = 0; b < n; b++) acc += x[(size_t)kIn * b + i] * d1[(size_t)kH * b + o];
  } else if (q < kW2) {  // bias_2[o]
    for (uint32_t b = 0; b < n; b++) acc += d2[(size_t)kH * b + (q - kB2)];
  } else if (q < kB4) {  // wgtT_2[i][o]
    const int i = (q - kW2) / kH, o = (q - kW2) % kH;
    for (uint32_t b = 0; b < n; b++) acc += h1[(size_t)kH * b + i] * d2[(size_t)kH * b + o];
  } else if (q < kW4) {  // bias_4
    for (uint32_t b = 0; b < n; b++) acc += dlogit[b];
  } else {  // wgtT_4[i]
    const int i = q - kW4;
    for (uint32_t b = 0; b < n; b++) acc += h2[(size_t)kH * b + i] * dlogit[b];
  }
  grad[q] = acc;
}

inline unsigned grid_for(uint32_t n) {
  unsigned g = (n + 255) / 256;
  return g == 0 ? 1 : (g > 4096 ? 4096 : g);
}

}  // namespace

extern "C" {

lst_status lst_t5dnn_forward(const float* x, const float* params, float* h1, float* h2, float* logit, uint32_t n,
                             void* stream) {
  if (!x || !params || !h1 || !h2 || !logit) return LST_E_INPUT;
  hipLaunchKernelGGL(k_mlp_forward, dim3(grid_for(n)), dim3(256), 0, static_cast<hipStream_t>(stream), x, params,
                     h1, h2, logit, n);
  return hipGetLastError() == hipSuccess ? LST_OK : LST_E_DEVICE;
}

lst_status lst_t5dnn_backward(const float* x, const float* params, const float* h1, const float* h2,
                              const float* dlogit, float* d2, float* d1, float* grad, uint32_t n, void* stream) {
  if (!x || !params || !h1 || !h2 || !dlogit || !d2 || !d1 || !grad) return LST_E_INPUT;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_mlp_backward_act, dim3(grid_for(n)), dim3(256), 0, s, params, h1, h2, dlogit, d2, d1, n);
  hipLaunchKernelGGL(k_mlp_backward_w, dim3((LST_T5DNN_NPARAMS + 255) / 256), dim3(256), 0, s, x, h1, h2, dlogit, d2,
                     d1, grad, n);
  return hipGetLastError() == hipSuccess ? LST_OK : LST_E_DEVICE;
}

}  // extern "C"
