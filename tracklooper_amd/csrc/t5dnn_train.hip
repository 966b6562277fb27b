// Training kernels of the T5 DNN (SDL/NeuralNetwork.h:116-162 architecture:
// Linear(38,32) ReLU Linear(32,32) ReLU Linear(32,1), sigmoid applied by the
// loss).  Parameters are one packed float array in the order of the chain's
// parameter files: bias_0[32], wgtT_0[38][32], bias_2[32], wgtT_2[32][32],
// bias_4[1], wgtT_4[32][1].
//
// Two paths:
//  * FP32 (lst_t5dnn_forward / lst_t5dnn_backward, behind the autograd
//    function of the Python driver): forward and activation gradients one
//    thread per sample, accumulated in index order as the in-chain inference
//    does; weight gradients per chunk of samples staged in LDS, then the chunk
//    partials summed in a fixed order (k_sum_chunks), so a gradient does not
//    depend on scheduling.
//  * BF16 on the matrix cores (lst_t5dnn_train_step_bf16): forward, binary
//    cross-entropy and backward fused per chunk of 64 samples, every GEMM a
//    v_mfma_f32_16x16x16_bf16 with BF16 operands and FP32 accumulation (the
//    32 -> 1 output layer and the element-wise steps in FP32), the chunk
//    partials summed in the fixed order as above.  lst_t5dnn_train_step_bf16x3 is
//    the same step with every operand split into two BF16 terms and three
//    MFMAs per product, which brings the GEMMs to about FP32 accuracy.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "../../include/lst_amd.h"

namespace {

constexpr int kIn = 38, kH = 32;
constexpr int kB0 = 0, kW0 = kB0 + kH, kB2 = kW0 + kIn * kH, kW2 = kB2 + kH, kB4 = kW2 + kH * kH, kW4 = kB4 + 1;
constexpr int kNP = LST_T5DNN_NPARAMS;
static_assert(kW4 + kH == kNP, "packed parameter layout");
constexpr int kPart = kNP + 1;  // a chunk's partial: the gradient, then the loss

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

// FP32 weight gradients of one chunk of kWRows samples (work-group blockIdx.x),
// the chunk's rows staged in LDS, every parameter summed over the chunk in
// sample order; partial q of chunk c at part[c * kNP + q]
constexpr int kWRows = 64;
__global__ __launch_bounds__(256) void k_mlp_wgrad_partial(const float* __restrict__ x, const float* __restrict__ h1,
                                                           const float* __restrict__ h2,
                                                           const float* __restrict__ dlogit,
                                                           const float* __restrict__ d2, const float* __restrict__ d1,
                                                           float* __restrict__ part, uint32_t n) {
  __shared__ float sX[kWRows][kIn], sH1[kWRows][kH], sH2[kWRows][kH], sD1[kWRows][kH], sD2[kWRows][kH];
  __shared__ float sG[kWRows];
  const uint32_t r0 = blockIdx.x * kWRows;
  const uint32_t rows = n - r0 < (uint32_t)kWRows ? n - r0 : (uint32_t)kWRows;
  for (uint32_t t = threadIdx.x; t < rows * kIn; t += blockDim.x) sX[t / kIn][t % kIn] = x[(size_t)kIn * r0 + t];
  for (uint32_t t = threadIdx.x; t < rows * kH; t += blockDim.x) {
    const size_t g = (size_t)kH * r0 + t;
    sH1[t / kH][t % kH] = h1[g];
    sH2[t / kH][t % kH] = h2[g];
    sD1[t / kH][t % kH] = d1[g];
    sD2[t / kH][t % kH] = d2[g];
  }
  for (uint32_t t = threadIdx.x; t < rows; t += blockDim.x) sG[t] = dlogit[r0 + t];
  __syncthreads();
  for (int q = threadIdx.x; q < kNP; q += blockDim.x) {
    float acc = 0.f;
    if (q < kW0) {  // bias_0[o]
      for (uint32_t b = 0; b < rows; b++) acc += sD1[b][q - kB0];
    } else if (q < kB2) {  // wgtT_0[i][o]
      const int i = (q - kW0) / kH, o = (q - kW0) % kH;
      for (uint32_t b = 0; b < rows; b++) acc += sX[b][i] * sD1[b][o];
    } else if (q < kW2) {  // bias_2[o]
      for (uint32_t b = 0; b < rows; b++) acc += sD2[b][q - kB2];
    } else if (q < kB4) {  // wgtT_2[i][o]
      const int i = (q - kW2) / kH, o = (q - kW2) % kH;
      for (uint32_t b = 0; b < rows; b++) acc += sH1[b][i] * sD2[b][o];
    } else if (q < kW4) {  // bias_4
      for (uint32_t b = 0; b < rows; b++) acc += sG[b];
    } else {  // wgtT_4[i]
      const int i = q - kW4;
      for (uint32_t b = 0; b < rows; b++) acc += sH2[b][i] * sG[b];
    }
    part[(size_t)blockIdx.x * kNP + q] = acc;
  }
}

// Sum of the chunk partials: column q < width of out = sum over chunks c of
// part[c * stride + q], columns q >= n1 going to out2[q - n1] (when out2 is
// given).  A work-group takes 64 columns; wave w adds the chunks of its quarter
// [w * per, (w + 1) * per) in chunk order, then the four wave sums are added in
// wave order: a fixed order, so the result does not depend on scheduling.
__global__ __launch_bounds__(256) void k_sum_chunks(const float* __restrict__ part, uint32_t chunks, uint32_t stride,
                                                    uint32_t width, float* __restrict__ out, uint32_t n1,
                                                    float* __restrict__ out2) {
  __shared__ float sum[4][64];
  const uint32_t w = threadIdx.x >> 6, l = threadIdx.x & 63, q = blockIdx.x * 64 + l;
  const uint32_t per = (chunks + 3) / 4, c0 = w * per, c1 = c0 + per < chunks ? c0 + per : chunks;
  float acc = 0.f;
  if (q < width) {
    uint32_t c = c0;
    for (; c + 8 <= c1; c += 8) {  // eight loads in flight, added in order
      float v[8];
#pragma unroll
      for (int k = 0; k < 8; k++) v[k] = part[(size_t)(c + k) * stride + q];
#pragma unroll
      for (int k = 0; k < 8; k++) acc += v[k];
    }
    for (; c < c1; c++) acc += part[(size_t)c * stride + q];
  }
  sum[w][l] = acc;
  __syncthreads();
  if (w == 0 && q < width) {
    const float t = ((sum[0][l] + sum[1][l]) + sum[2][l]) + sum[3][l];
    if (q < n1) out[q] = t;
    else if (out2) out2[q - n1] = t;
  }
}

// ------------------------------------------------------------------ BF16 MFMA
using bf16x4 = __attribute__((ext_vector_type(4))) short;
using f32x4 = __attribute__((ext_vector_type(4))) float;

// float -> bfloat16, round to nearest even (the inputs are finite)
__device__ __forceinline__ short f2bf(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}
__device__ __forceinline__ float bf2f(short s) { return __uint_as_float(((uint32_t)(uint16_t)s) << 16); }

// D = A B + C for 16x16 tiles.  Lane l holds A[l % 16][4 (l / 16) + e],
// B[4 (l / 16) + e][l % 16] and D[4 (l / 16) + e][l % 16], e = 0..3.
__device__ __forceinline__ f32x4 mma16(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// An operand of kT BF16 terms: kT = 1 is the value rounded to BF16; kT = 2
// adds the BF16 rounding of the remainder (hi + lo carries 16 mantissa bits).
// The product of two split operands takes three MFMAs, lo*hi + hi*lo + hi*hi
// (the lo*lo term is below FP32 rounding) -- the "bf16x3" scheme of Henry,
// Tang and Heinecke, "Leveraging the bfloat16 Artificial Intelligence Datatype
// For Higher-Precision Computations" (ARITH 2019).
template <int kT>
__device__ __forceinline__ void put_split(short* hi, short* lo, float v) {
  const short h = f2bf(v);
  *hi = h;
  if constexpr (kT == 2) *lo = f2bf(v - bf2f(h));
}
template <int kT>
__device__ __forceinline__ f32x4 mma_split(const bf16x4 (&a)[kT], const bf16x4 (&b)[kT], f32x4 c) {
  if constexpr (kT == 2) {
    c = mma16(a[1], b[0], c);
    c = mma16(a[0], b[1], c);
  }
  return mma16(a[0], b[0], c);
}

constexpr int kK1 = 48;              // the 38 inputs padded to three K steps of 16
constexpr int kStepWaves = 4;        // one 16-sample tile per wave
constexpr int kStepRows = 16 * kStepWaves;

// One chunk of kStepRows samples per work-group: forward, loss, backward, and the chunk's gradient partial
// (of the mean loss over n samples) at part[c * kPart], its loss sum at part[c * kPart + kNP].  Every GEMM
// operand has kT BF16 terms (put_split); the 32 -> 1 output layer and the element-wise steps are FP32.
template <int kT>
__global__ __launch_bounds__(256) void k_t5dnn_step(const float* __restrict__ x, const float* __restrict__ y,
                                                    const float* __restrict__ p, float* __restrict__ part,
                                                    uint32_t n, float invN) {
  __shared__ short sW0[kT][kK1][kH];         // layer 1 B operand [in][out], inputs 38..47 zero
  __shared__ short sW2[kT][kH][kH];          // layer 2 B operand [in][out]
  __shared__ short sW2t[kT][kH][kH];         // dH1 = dH2 W2^T B operand [out][in]
  __shared__ float sB0[kH], sB2[kH], sW4[kH];
  __shared__ short sX[kT][kStepWaves][16][kK1];
  __shared__ float sH1[kStepWaves][16][kH], sH2[kStepWaves][16][kH];
  __shared__ short sA[kT][kStepWaves][16][kH];   // H1 (layer 2 A), then dH2
  __shared__ short sD1[kT][kStepWaves][16][kH];  // dH1
  __shared__ float sDl[kStepWaves][16], sLoss[kStepWaves][16];
  __shared__ float sP[kStepWaves][kPart];        // per-wave partials, summed in wave order at the end
  constexpr int L = kT - 1;                      // index of the lo term (the hi term when kT = 1)

  const int w = threadIdx.x / 64, l = threadIdx.x & 63, m = l & 15, kg = l >> 4;
  for (int t = threadIdx.x; t < kK1 * kH; t += blockDim.x) {
    const int i = t / kH, o = t % kH;
    put_split<kT>(&sW0[0][i][o], &sW0[L][i][o], i < kIn ? p[kW0 + kH * i + o] : 0.f);
  }
  for (int t = threadIdx.x; t < kH * kH; t += blockDim.x) {
    const int i = t / kH, o = t % kH;
    const float v = p[kW2 + kH * i + o];
    put_split<kT>(&sW2[0][i][o], &sW2[L][i][o], v);
    put_split<kT>(&sW2t[0][o][i], &sW2t[L][o][i], v);
  }
  if (threadIdx.x < kH) {
    sB0[threadIdx.x] = p[kB0 + threadIdx.x];
    sB2[threadIdx.x] = p[kB2 + threadIdx.x];
    sW4[threadIdx.x] = p[kW4 + threadIdx.x];
  }
  const float b4 = p[kB4];
  const uint32_t r0 = blockIdx.x * kStepRows + 16 * w;  // this wave's first sample
  for (int t = l; t < 16 * kK1; t += 64) {
    const int r = t / kK1, i = t % kK1;
    const uint32_t g = r0 + r;
    put_split<kT>(&sX[0][w][r][i], &sX[L][w][r][i], (i < kIn && g < n) ? x[(size_t)kIn * g + i] : 0.f);
  }
  __syncthreads();

  // layer 1: H1 = relu(X W0 + b0), two 16-column halves, three K steps each
  for (int nh = 0; nh < 2; nh++) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int ks = 0; ks < 3; ks++) {
      bf16x4 a[kT], b[kT];
      for (int q = 0; q < kT; q++)
        for (int e = 0; e < 4; e++) {
          a[q][e] = sX[q][w][m][16 * ks + 4 * kg + e];
          b[q][e] = sW0[q][16 * ks + 4 * kg + e][16 * nh + m];
        }
      acc = mma_split<kT>(a, b, acc);
    }
    for (int e = 0; e < 4; e++) {
      const int r = 4 * kg + e, c = 16 * nh + m;
      float h = acc[e] + sB0[c];
      h = h > 0.f ? h : 0.f;
      sH1[w][r][c] = h;
      put_split<kT>(&sA[0][w][r][c], &sA[L][w][r][c], h);
    }
  }
  __syncthreads();
  // layer 2: H2 = relu(H1 W2 + b2)
  for (int nh = 0; nh < 2; nh++) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int ks = 0; ks < 2; ks++) {
      bf16x4 a[kT], b[kT];
      for (int q = 0; q < kT; q++)
        for (int e = 0; e < 4; e++) {
          a[q][e] = sA[q][w][m][16 * ks + 4 * kg + e];
          b[q][e] = sW2[q][16 * ks + 4 * kg + e][16 * nh + m];
        }
      acc = mma_split<kT>(a, b, acc);
    }
    for (int e = 0; e < 4; e++) {
      const int r = 4 * kg + e, c = 16 * nh + m;
      const float h = acc[e] + sB2[c];
      sH2[w][r][c] = h > 0.f ? h : 0.f;
    }
  }
  __syncthreads();
  // output layer, loss and dL/dlogit in FP32 (lanes 0..15: one sample each)
  if (l < 16) {
    const uint32_t g = r0 + l;
    float z = b4;
    for (int k = 0; k < kH; k++) z += sH2[w][l][k] * sW4[k];
    float dl = 0.f, loss = 0.f;
    if (g < n) {
      const float t = y[g];
      const float az = z < 0.f ? -z : z;
      loss = ((z > 0.f ? z : 0.f) - z * t + log1pf(expf(-az))) * invN;  // stable BCE with logits, of the mean
      const float s = z >= 0.f ? 1.f / (1.f + expf(-z)) : expf(z) / (1.f + expf(z));
      dl = (s - t) * invN;
    }
    sDl[w][l] = dl;
    sLoss[w][l] = loss;
  }
  __syncthreads();
  // dH2 = dlogit W4^T masked by H2 > 0 (split for the GEMMs); output-layer and layer-2 bias partials
  for (int t = l; t < 16 * kH; t += 64) {
    const int r = t / kH, k = t % kH;
    put_split<kT>(&sA[0][w][r][k], &sA[L][w][r][k], sH2[w][r][k] > 0.f ? sDl[w][r] * sW4[k] : 0.f);
  }
  float* P = sP[w];
  if (l < kH) {  // column l: sums over the tile's 16 samples in order
    float db2 = 0.f, dw4 = 0.f;
    for (int r = 0; r < 16; r++) {
      const float d2 = sH2[w][r][l] > 0.f ? sDl[w][r] * sW4[l] : 0.f;
      db2 += d2;
      dw4 += sH2[w][r][l] * sDl[w][r];
    }
    P[kB2 + l] = db2;
    P[kW4 + l] = dw4;
  }
  if (l == 0) {
    float db4 = 0.f, loss = 0.f;
    for (int r = 0; r < 16; r++) {
      db4 += sDl[w][r];
      loss += sLoss[w][r];
    }
    P[kB4] = db4;
    P[kNP] = loss;
  }
  __syncthreads();
  // dH1 = (dH2 W2^T) masked by H1 > 0
  for (int nh = 0; nh < 2; nh++) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int ks = 0; ks < 2; ks++) {
      bf16x4 a[kT], b[kT];
      for (int q = 0; q < kT; q++)
        for (int e = 0; e < 4; e++) {
          a[q][e] = sA[q][w][m][16 * ks + 4 * kg + e];
          b[q][e] = sW2t[q][16 * ks + 4 * kg + e][16 * nh + m];
        }
      acc = mma_split<kT>(a, b, acc);
    }
    for (int e = 0; e < 4; e++) {
      const int r = 4 * kg + e, c = 16 * nh + m;
      put_split<kT>(&sD1[0][w][r][c], &sD1[L][w][r][c], sH1[w][r][c] > 0.f ? acc[e] : 0.f);
    }
  }
  __syncthreads();
  if (l < kH) {  // layer-1 bias partial: column sums of dH1 (the split values the GEMM uses)
    float db0 = 0.f;
    for (int r = 0; r < 16; r++) {
      float d = bf2f(sD1[0][w][r][l]);
      if constexpr (kT == 2) d += bf2f(sD1[1][w][r][l]);
      db0 += d;
    }
    P[kB0 + l] = db0;
  }
  // dW2 = H1^T dH2 (in x out) and dW0 = X^T dH1 (inputs x hidden), K = the tile's 16 samples
  for (int mh = 0; mh < 2; mh++)
    for (int nh = 0; nh < 2; nh++) {
      bf16x4 a[kT], b[kT];
      for (int e = 0; e < 4; e++) {
        short hl[2];
        put_split<kT>(&hl[0], &hl[1], sH1[w][4 * kg + e][16 * mh + m]);
        for (int q = 0; q < kT; q++) {
          a[q][e] = hl[q];
          b[q][e] = sA[q][w][4 * kg + e][16 * nh + m];
        }
      }
      const f32x4 acc = mma_split<kT>(a, b, f32x4{0.f, 0.f, 0.f, 0.f});
      for (int e = 0; e < 4; e++) P[kW2 + kH * (16 * mh + 4 * kg + e) + 16 * nh + m] = acc[e];
    }
  for (int mh = 0; mh < 3; mh++)
    for (int nh = 0; nh < 2; nh++) {
      bf16x4 a[kT], b[kT];
      for (int q = 0; q < kT; q++)
        for (int e = 0; e < 4; e++) {
          a[q][e] = sX[q][w][4 * kg + e][16 * mh + m];
          b[q][e] = sD1[q][w][4 * kg + e][16 * nh + m];
        }
      const f32x4 acc = mma_split<kT>(a, b, f32x4{0.f, 0.f, 0.f, 0.f});
      for (int e = 0; e < 4; e++) {
        const int i = 16 * mh + 4 * kg + e;
        if (i < kIn) P[kW0 + kH * i + 16 * nh + m] = acc[e];
      }
    }
  __syncthreads();
  // the work-group's partial: the four waves' partials summed in wave order
  float* out = part + (size_t)blockIdx.x * kPart;
  for (int q = threadIdx.x; q < kPart; q += blockDim.x) out[q] = ((sP[0][q] + sP[1][q]) + sP[2][q]) + sP[3][q];
}

// Adam (Kingma and Ba 2015) on the packed parameters, step t >= 1
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                                              float* __restrict__ v, uint32_t np, float lr, float b1, float b2,
                                              float eps, float c1, float c2) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= np) return;
  const float mq = b1 * m[q] + (1.f - b1) * g[q];
  const float vq = b2 * v[q] + (1.f - b2) * g[q] * g[q];
  m[q] = mq;
  v[q] = vq;
  p[q] -= lr * (mq / c1) / (sqrtf(vq / c2) + eps);
}

inline unsigned grid_for(uint32_t n) {
  unsigned g = (n + 255) / 256;
  return g == 0 ? 1 : (g > 4096 ? 4096 : g);
}

// Per-device workspace of the allocating entry points (grown on demand, kept until the process ends).
constexpr int kMaxDevices = 64;
struct Workspace {
  std::mutex mu;
  float* ptr[kMaxDevices] = {};
  size_t floats[kMaxDevices] = {};
};
Workspace& workspace() {
  static Workspace w;
  return w;
}
float* workspace_for(size_t floats) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
  Workspace& w = workspace();
  std::lock_guard<std::mutex> lock(w.mu);
  if (w.floats[dev] < floats) {
    if (w.ptr[dev]) {
      if (hipDeviceSynchronize() != hipSuccess) return nullptr;
      (void)hipFree(w.ptr[dev]);
      w.ptr[dev] = nullptr;
      w.floats[dev] = 0;
    }
    if (hipMalloc(&w.ptr[dev], floats * sizeof(float)) != hipSuccess) return nullptr;
    w.floats[dev] = floats;
  }
  return w.ptr[dev];
}

// Scratch of the no-allocation BF16 step: the chunk partials of up to
// kScratchChunks chunks (kScratchChunks * kStepRows samples per call).  Host
// code uses the device address from hipGetSymbolAddress, never the name.
constexpr uint32_t kScratchChunks = 512;
static __device__ __align__(256) float scratch[kScratchChunks * kPart];
static_assert(kScratchChunks * kStepRows == LST_T5DNN_NOALLOC_MAX_ROWS, "scratch holds the declared maximum");
struct ScratchAddr {
  std::mutex mu;
  float* ptr[kMaxDevices] = {};
};
ScratchAddr& scratch_addr() {
  static ScratchAddr s;
  return s;
}
// this device's copy of `scratch` (looked up once per device, with the device current)
float* scratch_for_device() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
  ScratchAddr& s = scratch_addr();
  std::lock_guard<std::mutex> lock(s.mu);
  if (!s.ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
    s.ptr[dev] = static_cast<float*>(p);
  }
  return s.ptr[dev];
}

lst_status step_bf16(const float* x, const float* y, const float* params, float* grad, float* loss, uint32_t n,
                     float* part, hipStream_t s, int terms) {
  const uint32_t chunks = (n + kStepRows - 1) / kStepRows;
  if (chunks == 0) return LST_OK;
  if (terms == 2) {
    hipLaunchKernelGGL(k_t5dnn_step<2>, dim3(chunks), dim3(256), 0, s, x, y, params, part, n, 1.f / (float)n);
  } else {
    hipLaunchKernelGGL(k_t5dnn_step<1>, dim3(chunks), dim3(256), 0, s, x, y, params, part, n, 1.f / (float)n);
  }
  hipLaunchKernelGGL(k_sum_chunks, dim3((kPart + 63) / 64), dim3(256), 0, s, part, chunks, (uint32_t)kPart,
                     (uint32_t)kPart, grad, (uint32_t)kNP, loss);
  return hipGetLastError() == hipSuccess ? LST_OK : LST_E_DEVICE;
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
  const uint32_t chunks = (n + kWRows - 1) / kWRows;
  hipLaunchKernelGGL(k_mlp_backward_act, dim3(grid_for(n)), dim3(256), 0, s, params, h1, h2, dlogit, d2, d1, n);
  if (chunks == 0) {
    hipLaunchKernelGGL(k_sum_chunks, dim3((kNP + 63) / 64), dim3(256), 0, s, grad, 0u, 0u, (uint32_t)kNP, grad,
                       (uint32_t)kNP, nullptr);
    return hipGetLastError() == hipSuccess ? LST_OK : LST_E_DEVICE;
  }
  float* part = workspace_for((size_t)chunks * kNP);
  if (!part) return LST_E_DEVICE;
  hipLaunchKernelGGL(k_mlp_wgrad_partial, dim3(chunks), dim3(256), 0, s, x, h1, h2, dlogit, d2, d1, part, n);
  hipLaunchKernelGGL(k_sum_chunks, dim3((kNP + 63) / 64), dim3(256), 0, s, part, chunks, (uint32_t)kNP,
                     (uint32_t)kNP, grad, (uint32_t)kNP, nullptr);
  return hipGetLastError() == hipSuccess ? LST_OK : LST_E_DEVICE;
}

lst_status lst_t5dnn_train_step_bf16(const float* x, const float* y, const float* params, float* grad, float* loss,
                                     uint32_t n, void* stream) {
  if (!x || !y || !params || !grad) return LST_E_INPUT;
  const uint32_t chunks = (n + kStepRows - 1) / kStepRows;
  float* part = workspace_for((size_t)(chunks ? chunks : 1) * kPart);
  if (!part) return LST_E_DEVICE;
  return step_bf16(x, y, params, grad, loss, n, part, static_cast<hipStream_t>(stream), 1);
}

lst_status lst_t5dnn_train_step_bf16x3(const float* x, const float* y, const float* params, float* grad, float* loss,
                                       uint32_t n, void* stream) {
  if (!x || !y || !params || !grad) return LST_E_INPUT;
  const uint32_t chunks = (n + kStepRows - 1) / kStepRows;
  float* part = workspace_for((size_t)(chunks ? chunks : 1) * kPart);
  if (!part) return LST_E_DEVICE;
  return step_bf16(x, y, params, grad, loss, n, part, static_cast<hipStream_t>(stream), 2);
}

lst_status lst_t5dnn_train_step_bf16_noalloc(const float* x, const float* y, const float* params, float* grad,
                                             float* loss, uint32_t n, void* stream) {
  if (!x || !y || !params || !grad) return LST_E_INPUT;
  if (n > LST_T5DNN_NOALLOC_MAX_ROWS) return LST_E_CAPACITY;
  float* part = scratch_for_device();
  if (!part) return LST_E_DEVICE;
  return step_bf16(x, y, params, grad, loss, n, part, static_cast<hipStream_t>(stream), 1);
}

lst_status lst_t5dnn_adam(float* params, const float* grad, float* m, float* v, uint32_t step, float lr, float beta1,
                          float beta2, float eps, void* stream) {
  if (!params || !grad || !m || !v || step == 0) return LST_E_INPUT;
  const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3((kNP + 255) / 256), dim3(256), 0, static_cast<hipStream_t>(stream), params, grad,
                     m, v, (uint32_t)kNP, lr, beta1, beta2, eps, c1, c2);
  return hipGetLastError() == hipSuccess ? LST_OK : LST_E_DEVICE;
}

}  // extern "C"
