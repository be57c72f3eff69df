// Shared device helpers for the MaoMi optical-stage HIP library (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <cstdint>
#include <cstdio>

#define MAOMI_API extern "C" __attribute__((visibility("default")))

// Error codes returned by every C-ABI entry point (0 = success; positive
// values are hipError_t codes passed through from the runtime).
enum MaomiStatus : int {
  MAOMI_OK = 0,
  MAOMI_EINVAL = -1,      // an argument failed validation (see maomi_last_error)
  MAOMI_EUNSUPPORTED = -2 // a grid size or mode the kernels do not implement
};

// Records a human-readable message for the last failing call on this thread.
void maomi_set_error(const char* fmt, ...);

#define MAOMI_CHECK_LAUNCH()                                              \
  do {                                                                    \
    hipError_t e_ = hipGetLastError();                                    \
    if (e_ != hipSuccess) {                                               \
      maomi_set_error("%s: launch failed: %s", __func__, hipGetErrorString(e_)); \
      return (int)e_;                                                     \
    }                                                                     \
  } while (0)

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {  // a * conj(b)
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// Block-wide sum of a double; every thread receives the total.  `red` needs
// blockDim.x / 64 doubles of LDS.
__device__ __forceinline__ double block_sum_double(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double t = 0.0;
  const int nw = (blockDim.x + 63) >> 6;
  for (int i = 0; i < nw; ++i) t += red[i];
  __syncthreads();
  return t;
}
