// Shared device/host helpers for the gfx950 FPI + EFE kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

namespace pmh {

// pymdp/maths.py:15  MINVAL = float64 eps, applied to FP32 data as the FP32
// value nearest it (log_stable, stable_xlogx clips).
constexpr float kMinVal = 2.220446049250313e-16f;

constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// Error plumbing (host side)
// ---------------------------------------------------------------------------
void set_error(const std::string& msg);
const char* get_error();

#define PMH_HIP_CHECK(expr, prefix)                                                    \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess) {                                                            \
      ::pmh::set_error(std::string(prefix) + #expr + ": " + hipGetErrorString(_e));    \
      return 2;                                                                        \
    }                                                                                  \
  } while (0)

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ float log_stable(float x) { return logf(fmaxf(x, kMinVal)); }

// xlogy(x, clip(x, MINVAL)); exactly 0 for x == 0 (pymdp/maths.py:17-30).
__device__ __forceinline__ float stable_xlogx(float x) {
  return x == 0.0f ? 0.0f : x * logf(fmaxf(x, kMinVal));
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, kWave));
  return v;
}

// Sum over the low `log2w` lane bits (groups of 2^log2w adjacent lanes).
__device__ __forceinline__ float group_sum_low(float v, int log2w) {
  for (int b = 0; b < log2w; ++b) v += __shfl_xor(v, 1 << b, kWave);
  return v;
}

// Sum over lane bits [log2w, 6): lanes that share (lane & (2^log2w - 1)).
__device__ __forceinline__ float group_sum_high(float v, int log2w) {
  for (int b = log2w; b < 6; ++b) v += __shfl_xor(v, 1 << b, kWave);
  return v;
}

__host__ __device__ __forceinline__ int ceil_log2(int n) {
  int l = 0;
  while ((1 << l) < n) ++l;
  return l;
}

}  // namespace pmh
