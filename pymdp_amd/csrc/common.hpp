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

// ---------------------------------------------------------------------------
// ln(x) of a float, evaluated to FP64 accuracy (~1e-16 relative).
// x = 2^e * m with m in [1, 2); m = c_i * (1 + r), c_i = 1 + i/128 (i = top 7
// mantissa bits), so ln x = e ln2 + ln c_i + ln(1 + r) with 0 <= r < 2^-7 and
// ln(1 + r) from its Taylor series (degree 7: truncation < 2^-59).  The two
// 128-entry tables (ln c_i, 1/c_i) are computed on the host in double and
// copied to constant memory once per device (ensure_log_table).
// ---------------------------------------------------------------------------
static __constant__ double g_logtab[256];  // [0,128): ln(1+i/128); [128,256): 1/(1+i/128)

// Per translation unit (internal linkage: each TU has its own g_logtab and
// its own "initialised" flags).
namespace {
inline int ensure_log_table() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 2;
  static bool done[64] = {false};
  if (dev < 0 || dev >= 64) return 2;
  if (done[dev]) return 0;
  double h[256];
  for (int i = 0; i < 128; ++i) {
    const double c = 1.0 + i / 128.0;
    h[i] = __builtin_log(c);
    h[128 + i] = 1.0 / c;
  }
  if (hipMemcpyToSymbol(HIP_SYMBOL(g_logtab), h, sizeof(h)) != hipSuccess) return 2;
  done[dev] = true;
  return 0;
}
}  // namespace

// x must be a positive normal float (callers clip at kMinVal first).
__device__ __forceinline__ double log_f64(float x) {
  const unsigned bits = __float_as_uint(x);
  const int e = (int)(bits >> 23) - 127;
  const unsigned mb = bits & 0x7FFFFFu;
  const int i = (int)(mb >> 16);
  const double m = 1.0 + (double)mb * 0x1p-23;
  const double c = 1.0 + (double)i * 0x1p-7;
  const double r = (m - c) * g_logtab[128 + i];
  const double p =
      r * (1.0 + r * (-0.5 + r * (1.0 / 3.0 + r * (-0.25 + r * (0.2 + r * (-1.0 / 6.0 + r * (1.0 / 7.0)))))));
  return (double)e * 0.69314718055994530942 + g_logtab[i] + p;
}

// ln(x) of a positive normal double to ~1e-16 relative, with the two tables
// given by pointer (e.g. staged in LDS): tab[0,128) = ln(1 + i/128),
// tab[128,256) = 1/(1 + i/128).  Degree-6 series, truncation < 2^-49 / 7.
__device__ __forceinline__ double log_f64d(double x, const double* tab) {
  const long long bits = __double_as_longlong(x);
  const int e = (int)((bits >> 52) & 0x7FF) - 1023;
  const long long mb = bits & 0xFFFFFFFFFFFFFll;
  const int i = (int)(mb >> 45);
  const double m = __longlong_as_double(mb | 0x3FF0000000000000ll);
  const double c = 1.0 + (double)i * 0x1p-7;
  const double r = (m - c) * tab[128 + i];
  const double p = r * (1.0 + r * (-0.5 + r * (1.0 / 3.0 + r * (-0.25 + r * (0.2 + r * (-1.0 / 6.0))))));
  return (double)e * 0.69314718055994530942 + tab[i] + p;
}

// ln(x) of a positive normal double with an absolute error below ~1.5e-7
// at any magnitude: the exponent is split off exactly (e ln 2 in FP64) and
// only ln of the mantissa m in [1, 2) comes from the hardware FP32 log2
// (v_log_f32, about 1 ulp of a result in [0, 1)) of m rounded to float
// (relative 6e-8).  A few instructions and no table, for sums of q ln q
// weighted by probabilities (error <= 1.5e-7 per unit of weight).
__device__ __forceinline__ double ln_fast(double x) {
  const long long bits = __double_as_longlong(x);
  const int e = (int)((bits >> 52) & 0x7FF) - 1023;
  const double m = __longlong_as_double((bits & 0xFFFFFFFFFFFFFll) | 0x3FF0000000000000ll);
  return ((double)e + (double)__builtin_amdgcn_logf((float)m)) * 0.69314718055994530942;
}

// log_stable in FP64 accuracy: ln(max(x, MINVAL)).
__device__ __forceinline__ double log_stable_f64(float x) { return log_f64(fmaxf(x, kMinVal)); }

// x ln x with the reference's clip (0 for x == 0), in FP64 accuracy.
__device__ __forceinline__ double xlogx_f64(float x) { return x == 0.0f ? 0.0 : (double)x * log_stable_f64(x); }

// q ln q of an FP64 value; the log argument is rounded to float first
// (relative 6e-8, i.e. an absolute 6e-8 on ln q), the product stays FP64.
__device__ __forceinline__ double xlogx_d(double q) {
  return q == 0.0 ? 0.0 : q * log_stable_f64((float)q);
}

// Workgroup barrier that orders LDS accesses only (s_waitcnt lgkmcnt(0);
// s_barrier): global loads issued before it stay in flight, unlike
// __syncthreads(), whose workgroup-scope fences also wait for them.  Use it
// only where no thread reads global memory another thread of the
// workgroup wrote since the last full barrier.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const T u = __shfl_xor(v, o, kWave);
    v = u > v ? u : v;
  }
  return v;
}

// Sum over the low `log2w` lane bits (groups of 2^log2w adjacent lanes).
template <typename T>
__device__ __forceinline__ T group_sum_low(T v, int log2w) {
  for (int b = 0; b < log2w; ++b) v += __shfl_xor(v, 1 << b, kWave);
  return v;
}

// DPP cross-lane reductions of doubles (VALU moves between lanes instead of
// ds_bpermute round trips through the LDS unit).  Source lanes are always in
// the destination's own aligned group, so every lane of each reduced group
// must be active; the wave-wide forms need the whole wave active.
//   0xB1: quad_perm [1,0,3,2] (lane ^ 1)     0x4E: quad_perm [2,3,0,1] (lane ^ 2)
//   0x141: row_half_mirror (lane i <-> 7 - i within 8)
//   0x140: row_mirror (lane i <-> 15 - i within 16)
// After the xor-1 and xor-2 steps every aligned group of 4 is uniform, so a
// half-row mirror pairs each lane with the other group of 4 of its 8, and a
// row mirror each lane with the other group of 8 of its 16.
template <int Ctrl>
__device__ __forceinline__ double dpp_mov_f64(double v) {
  const long long bits = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)bits, Ctrl, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(bits >> 32), Ctrl, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ double readlane_f64(double v, int l) {
  const long long bits = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)bits, l);
  const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum of the whole wave in every lane: rows of 16 by DPP, then the four row
// sums read from lanes 0, 16, 32, 48 (fixed order).
__device__ __forceinline__ double wave_sum_dpp(double v) {
  v += dpp_mov_f64<0xB1>(v);
  v += dpp_mov_f64<0x4E>(v);
  v += dpp_mov_f64<0x141>(v);
  v += dpp_mov_f64<0x140>(v);
  return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}

__device__ __forceinline__ double wave_max_dpp(double v) {
  v = fmax(v, dpp_mov_f64<0xB1>(v));
  v = fmax(v, dpp_mov_f64<0x4E>(v));
  v = fmax(v, dpp_mov_f64<0x141>(v));
  v = fmax(v, dpp_mov_f64<0x140>(v));
  return fmax(fmax(readlane_f64(v, 0), readlane_f64(v, 16)), fmax(readlane_f64(v, 32), readlane_f64(v, 48)));
}

// group_sum_low for doubles: groups of up to 16 lanes by DPP, wider groups
// finish with shuffles.  Every lane of a group ends with the group's sum.
__device__ __forceinline__ double group_sum_low_dpp(double v, int log2w) {
  if (log2w > 0) v += dpp_mov_f64<0xB1>(v);
  if (log2w > 1) v += dpp_mov_f64<0x4E>(v);
  if (log2w > 2) v += dpp_mov_f64<0x141>(v);
  if (log2w > 3) v += dpp_mov_f64<0x140>(v);
  for (int b = 4; b < log2w; ++b) v += __shfl_xor(v, 1 << b, kWave);
  return v;
}

// Sum over lane bits [log2w, 6): lanes that share (lane & (2^log2w - 1)).
template <typename T>
__device__ __forceinline__ T group_sum_high(T v, int log2w) {
  for (int b = log2w; b < 6; ++b) v += __shfl_xor(v, 1 << b, kWave);
  return v;
}

__host__ __device__ __forceinline__ int ceil_log2(int n) {
  int l = 0;
  while ((1 << l) < n) ++l;
  return l;
}

}  // namespace pmh
