// Shared definitions for the fea_amd HIP library (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fea_amd.h"

namespace fea {

// thread-local error message, exposed through fea_last_error()
void set_error(const char *fmt, ...);

#define FEA_CHECK_ARG(cond, ...)                \
  do {                                          \
    if (!(cond)) {                              \
      ::fea::set_error(__VA_ARGS__);            \
      return FEA_ERR_ARG;                       \
    }                                           \
  } while (0)

#define FEA_CHECK_LAUNCH(name)                                                   \
  do {                                                                           \
    hipError_t e_ = hipGetLastError();                                           \
    if (e_ != hipSuccess) {                                                      \
      ::fea::set_error("%s: launch failed: %s", name, hipGetErrorString(e_));    \
      return FEA_ERR_HIP;                                                        \
    }                                                                            \
  } while (0)

#define FEA_HIP(call, name)                                                      \
  do {                                                                           \
    hipError_t e_ = (call);                                                      \
    if (e_ != hipSuccess) {                                                      \
      ::fea::set_error("%s: %s", name, hipGetErrorString(e_));                   \
      return FEA_ERR_HIP;                                                        \
    }                                                                            \
  } while (0)

constexpr int kBlock = 256;   // 4 wave64 per workgroup

inline int blocks_for(int64_t n, int per_block = kBlock) {
  return (int)((n + per_block - 1) / per_block);
}

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// ---------------------------------------------------------------- device math
// The element-force math (these helpers, the kinematics, stress and shape-gradient helpers and
// the per-element force routine) fuses a * b + c into an fma only inside one expression
// (`#pragma clang fp contract(on)` in each body): the same source then rounds alike in every
// kernel that inlines it.  Under the default cross-statement contraction the backend picks
// which product of a * b + c * d to fuse by the surrounding kernel's code, and two kernels
// computing one element's forces could differ in the last float bit.

template <typename T>
__device__ __forceinline__ T det3(const T a[9]) {
  #pragma clang fp contract(on)
  return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
         a[2] * (a[3] * a[7] - a[4] * a[6]);
}

// inverse of a row-major 3x3 given its determinant
template <typename T>
__device__ __forceinline__ void inv3(const T a[9], T d, T out[9]) {
  #pragma clang fp contract(on)
  T id = T(1) / d;
  out[0] = (a[4] * a[8] - a[5] * a[7]) * id;
  out[1] = (a[2] * a[7] - a[1] * a[8]) * id;
  out[2] = (a[1] * a[5] - a[2] * a[4]) * id;
  out[3] = (a[5] * a[6] - a[3] * a[8]) * id;
  out[4] = (a[0] * a[8] - a[2] * a[6]) * id;
  out[5] = (a[2] * a[3] - a[0] * a[5]) * id;
  out[6] = (a[3] * a[7] - a[4] * a[6]) * id;
  out[7] = (a[1] * a[6] - a[0] * a[7]) * id;
  out[8] = (a[0] * a[4] - a[1] * a[3]) * id;
}

// det(I + H) - 1 without cancellation: tr H + (sum of principal 2-minors) + det H
template <typename T>
__device__ __forceinline__ T detm1_of_grad(const T h[9]) {
  #pragma clang fp contract(on)
  T tr = h[0] + h[4] + h[8];
  T m2 = (h[0] * h[4] - h[1] * h[3]) + (h[4] * h[8] - h[5] * h[7]) + (h[0] * h[8] - h[2] * h[6]);
  return tr + m2 + det3(h);
}

// Block-wide float64 sum for blockDim.x == kBlock; result valid in thread 0.
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double red[kBlock / 64];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    for (int w = 0; w < kBlock / 64; ++w) s += red[w];
  }
  __syncthreads();
  return s;
}

// exclusive scan of in[0..n) into out[0..n] (out[n] = the total) by one workgroup of 1024
// threads: each thread sums a contiguous chunk, the 1024 chunk sums are scanned with wave
// shuffles (16 waves) and a scan of the 16 wave totals, and each thread writes its chunk's
// prefixes.  Integer and exact.  in == out is allowed (each thread reads an entry of its own
// chunk before it overwrites it; out[n] lies outside every chunk).
__device__ inline void scan_1024(int n, const int32_t *in, int32_t *out) {
  __shared__ int32_t wsum[16];
  const int t = threadIdx.x;
  const int lane = t & 63, wv = t >> 6;
  const int chunk = (n + 1023) / 1024;
  const int b0 = t * chunk, b1 = min(n, b0 + chunk);
  int32_t s = 0;
  for (int b = b0; b < b1; ++b) s += in[b];
  int32_t v = s;  // inclusive scan within the wave
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int32_t u = __shfl_up(v, off, 64);
    if (lane >= off) v += u;
  }
  if (lane == 63) wsum[wv] = v;
  __syncthreads();
  if (wv == 0) {
    int32_t w = lane < 16 ? wsum[lane] : 0;
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) {
      const int32_t u = __shfl_up(w, off, 64);
      if (lane >= off) w += u;
    }
    if (lane < 16) wsum[lane] = w;  // inclusive over the waves
  }
  __syncthreads();
  const int32_t incl = v + (wv > 0 ? wsum[wv - 1] : 0);
  if (t == 1023) out[n] = incl;
  int32_t acc = incl - s;
  for (int b = b0; b < b1; ++b) {
    const int32_t v = in[b];
    out[b] = acc;
    acc += v;
  }
}

}  // namespace fea
