// rank_mlp.hip -- rank-slab sizes for the batched ACA from a small MLP.
//
// Per admissible leaf, 8 geometric features (tree level, log n1, log n2,
// diam ratio, distance / max diam, log10 eps, kappa, log10 max diam) go
// through an 8-64-64-1 ReLU network predicting log(1 + ACA rank).  Weights
// arrive in FP32 and are rounded to BF16; activations are rounded to BF16
// between layers, products accumulate in FP32 ("BF16 inference").  The slab
// is ceil(pad * (exp(pred) - 1)) + margin, clipped to [1, maxk].  ACA keeps
// going past an under-sized slab (the leaf is rerun with a larger one), so
// the prediction changes memory and time, never the assembled matrix.
#include <cmath>

#include "ahm_internal.h"

namespace ahm {

constexpr int kF = 8, kH = 64;
// packed FP32 parameters: W1 (kF x kH, row = input), b1, W2 (kH x kH), b2,
// W3 (kH), b3, feature mean (kF), feature std (kF)
constexpr int kParams = kF * kH + kH + kH * kH + kH + kH + 1 + 2 * kF;

__device__ __forceinline__ float bf16_round(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
  return __uint_as_float(u & 0xFFFF0000u);
}

__global__ __launch_bounds__(256) void rank_mlp_kernel(const float* __restrict__ feats, uint32_t n,
                                                       const float* __restrict__ prm, float pad,
                                                       uint32_t margin, const uint32_t* __restrict__ maxk,
                                                       uint32_t* __restrict__ hint,
                                                       float* __restrict__ pred) {
  __shared__ float w1[kF * kH], w2[kH * kH], w3[kH], bb1[kH], bb2[kH], mu[kF], sd[kF];
  __shared__ float bb3;
  const float* W1 = prm;
  const float* B1 = W1 + kF * kH;
  const float* W2 = B1 + kH;
  const float* B2 = W2 + kH * kH;
  const float* W3 = B2 + kH;
  const float* B3 = W3 + kH;
  const float* MU = B3 + 1;
  const float* SD = MU + kF;
  for (int i = threadIdx.x; i < kF * kH; i += blockDim.x) w1[i] = bf16_round(W1[i]);
  for (int i = threadIdx.x; i < kH * kH; i += blockDim.x) w2[i] = bf16_round(W2[i]);
  for (int i = threadIdx.x; i < kH; i += blockDim.x) {
    w3[i] = bf16_round(W3[i]);
    bb1[i] = B1[i];
    bb2[i] = B2[i];
  }
  if (threadIdx.x < kF) {
    mu[threadIdx.x] = MU[threadIdx.x];
    sd[threadIdx.x] = SD[threadIdx.x];
  }
  if (threadIdx.x == 0) bb3 = B3[0];
  __syncthreads();
  const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n) return;
  float x[kF], h1[kH];
  for (int f = 0; f < kF; ++f) x[f] = bf16_round((feats[(size_t)l * kF + f] - mu[f]) / sd[f]);
  for (int j = 0; j < kH; ++j) {
    float s = bb1[j];
    for (int f = 0; f < kF; ++f) s = fmaf(x[f], w1[f * kH + j], s);
    h1[j] = bf16_round(fmaxf(s, 0.f));
  }
  float out = bb3;
  for (int j = 0; j < kH; ++j) {
    float s = bb2[j];
    for (int i = 0; i < kH; ++i) s = fmaf(h1[i], w2[i * kH + j], s);
    out = fmaf(bf16_round(fmaxf(s, 0.f)), w3[j], out);
  }
  if (pred) pred[l] = out;
  const float r = fmaxf(expf(out) - 1.f, 0.f);
  uint32_t k = (uint32_t)ceilf(pad * r) + margin;
  if (k < 1) k = 1;
  if (maxk && k > maxk[l]) k = maxk[l];
  hint[l] = k;
}

}  // namespace ahm

extern "C" int ahm_rank_predict(ahm_ctx ctx, const float* feats, uint32_t n, const float* params,
                                float pad, uint32_t margin, const uint32_t* maxk, uint32_t* hint,
                                float* pred) {
  using namespace ahm;
  if (!ctx || !feats || !params || !hint) return fail(AHM_ERR_INVALID, "ahm_rank_predict: NULL argument");
  if (n == 0) return AHM_OK;
  AHM_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  DeviceBuf df, dp, dm, dh, dpr;
  AHM_TRY(df.alloc((size_t)n * kF * 4));
  AHM_TRY(dp.alloc((size_t)kParams * 4));
  AHM_TRY(dh.alloc((size_t)n * 4));
  if (maxk) AHM_TRY(dm.alloc((size_t)n * 4));
  if (pred) AHM_TRY(dpr.alloc((size_t)n * 4));
  AHM_HIP(hipMemcpyAsync(df.p, feats, (size_t)n * kF * 4, hipMemcpyHostToDevice, s));
  AHM_HIP(hipMemcpyAsync(dp.p, params, (size_t)kParams * 4, hipMemcpyHostToDevice, s));
  if (maxk) AHM_HIP(hipMemcpyAsync(dm.p, maxk, (size_t)n * 4, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(rank_mlp_kernel, dim3((n + 255) / 256), dim3(256), 0, s, (const float*)df.p, n,
                     (const float*)dp.p, pad, margin, maxk ? (const uint32_t*)dm.p : nullptr,
                     (uint32_t*)dh.p, pred ? (float*)dpr.p : nullptr);
  AHM_HIP(hipGetLastError());
  AHM_HIP(hipMemcpyAsync(hint, dh.p, (size_t)n * 4, hipMemcpyDeviceToHost, s));
  if (pred) AHM_HIP(hipMemcpyAsync(pred, dpr.p, (size_t)n * 4, hipMemcpyDeviceToHost, s));
  AHM_HIP(hipStreamSynchronize(s));
  return AHM_OK;
}
