// Dirichlet-parameter learning of the observation and transition models
// (pymdp/learning.py:11-234, driven by Agent.infer_parameters,
// pymdp/agent.py:416-620): for every agent b and every column c of a count
// tensor p [R][N] (R = the normalised axis: outcomes of A_m, next states of
// B_f; N = every other axis, row-major),
//
//   q[b][r][c] = p[b][r][c] + lr[b] * sum_t rowd[b][t][r] * prod_j col_j[b][t][c_j]
//   e[b][r][c] = max(q, MINVAL) / sum_r' max(q[b][r'][c], MINVAL)
//
// rowd is the observation (A) or the next-state belief (B); col_j are the
// beliefs of the dependency factors at t and, for B, the one-hot action of
// the factor at t (zero when the action is out of range or, in the agent's
// masked form, when any factor's action at t is negative).
//
// One fused pass, HBM-bound (p read once, q and e written once: 12 bytes per
// element).  Up to 64 rows: a workgroup stages a [rows x 128 columns] tile
// in LDS with 16-byte loads, a thread per column adds the update and forms
// the column sum, and the tile leaves as 16-byte stores of q and e.  More
// rows: a thread owns one column, holds 64 rows of it in registers at a time
// and re-reads its own q for e.  The streamed tensors are loaded and stored
// nontemporally: with ordinary 16-byte accesses the L2 fetched ~1.9x and
// wrote back 1.5x the tensor bytes (EA request counters), nontemporal
// accesses move exactly the tensor bytes and run 1.7x faster.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>

#include "../../include/pymdp_hip.h"
#include "common.hpp"

namespace pmh {
namespace lrn {

constexpr int kNT = 256;
constexpr int kMaxC = 8;      // column factors (dependencies + action)
constexpr int kRowsReg = 64;  // rows of a column held in registers

struct Args {
  const float* p;
  long long p_bs;  // batch stride of p (0: shared by every agent)
  float* q;
  float* e;
  long long N;
  int R, T, nc, nblk;  // nblk = column blocks per agent
  int n[kMaxC];
  const float* col[kMaxC];
  long long col_bs[kMaxC], col_ts[kMaxC];
  const float* rowd;
  long long row_bs, row_ts;
  int has_act;     // 1: column factor nc - 1 is onehot(act[b][t][act_f])
  const int* act;
  long long act_bs, act_ts;
  int act_F, act_f, mask_all;
  const float* lr;
  float lr_s;
};

// weight of column (cj) at step t: prod_j col_j[t][c_j] times the action
// indicator (0 when masked)
__device__ __forceinline__ double col_weight(const Args& a, int b, int t, const int* cj, int nca) {
  if (a.has_act) {
    const int* at = a.act + b * a.act_bs + t * a.act_ts;
    bool ok = at[a.act_f] == cj[a.nc - 1];
    if (a.mask_all)
      for (int f = 0; f < a.act_F; ++f) ok = ok && at[f] >= 0;
    if (!ok) return 0.0;
  }
  double w = 1.0;
#pragma unroll
  for (int j = 0; j < kMaxC; ++j)
    if (j < nca) w *= (double)a.col[j][b * a.col_bs[j] + t * a.col_ts[j] + cj[j]];
  return w;
}

__device__ __forceinline__ void col_index(const Args& a, long long c, int* cj) {
  long long rest = c;
#pragma unroll
  for (int j = kMaxC - 1; j >= 0; --j) {
    cj[j] = 0;
    if (j < a.nc) {
      cj[j] = (int)(rest % a.n[j]);
      rest /= a.n[j];
    }
  }
}

constexpr int kTC = 128;  // columns of an LDS tile

// Tile kernel for R <= RMAX.  VEC: N, the batch stride and every base
// pointer are multiples of 4 floats, so rows move as float4.
typedef float v4f __attribute__((ext_vector_type(4)));

template <int RMAX, bool VEC>
__global__ __launch_bounds__(kNT) void dirichlet_tile_kernel(Args a) {
  __shared__ float s_t[RMAX * kTC];
  __shared__ float s_sum[2][kTC];
  const int tid = threadIdx.x;
  const int b = blockIdx.x / a.nblk;
  const long long c0 = (long long)(blockIdx.x - b * a.nblk) * kTC;
  const int ncol = (int)min((long long)kTC, a.N - c0);
  const int R = a.R;
  const float* pp = a.p + b * a.p_bs + c0;
  const long long ob = (long long)b * R * a.N + c0;
  if (VEC && ncol == kTC) {
    constexpr int NB = RMAX * kTC / 4 / kNT;
    // every thread issues all NB loads before using any, unconditionally, so
    // they are in flight together (tile rows past R repeat row R - 1: in
    // bounds, and those LDS rows are never read)
    v4f v[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) {
      const int e = (tid + k * kNT) * 4, r = min(e / kTC, R - 1), cl = e % kTC;
      v[k] = __builtin_nontemporal_load(reinterpret_cast<const v4f*>(pp + (long long)r * a.N + cl));
    }
#pragma unroll
    for (int k = 0; k < NB; ++k) *reinterpret_cast<v4f*>(s_t + (tid + k * kNT) * 4) = v[k];
  } else {
    for (int e = tid; e < R * kTC; e += kNT) {
      const int r = e / kTC, cl = e - r * kTC;
      if (cl < ncol) s_t[e] = __builtin_nontemporal_load(pp + (long long)r * a.N + cl);
    }
  }
  __syncthreads();
  {
    const int cl = tid % kTC, half = tid / kTC;
    float part = 0.0f;
    if (cl < ncol) {
      int cj[kMaxC];
      col_index(a, c0 + cl, cj);
      const int nca = a.has_act ? a.nc - 1 : a.nc;
      const double lr = a.lr ? (double)a.lr[b] : (double)a.lr_s;
      if (a.T == 1) {
        const double w = lr * col_weight(a, b, 0, cj, nca);
        const float* rd = a.rowd + b * a.row_bs;
        if (w != 0.0)
          for (int r = half; r < R; r += 2) s_t[r * kTC + cl] = (float)fma((double)rd[r], w, (double)s_t[r * kTC + cl]);
      } else {
        for (int r = half; r < R; r += 2) {
          double acc = 0.0;
          for (int t = 0; t < a.T; ++t) {
            const double w = col_weight(a, b, t, cj, nca);
            if (w != 0.0) acc = fma((double)a.rowd[b * a.row_bs + t * a.row_ts + r], w, acc);
          }
          s_t[r * kTC + cl] = (float)fma(lr, acc, (double)s_t[r * kTC + cl]);
        }
      }
      for (int r = half; r < R; r += 2) part += fmaxf(s_t[r * kTC + cl], kMinVal);
    }
    s_sum[half][cl] = part;
  }
  __syncthreads();
  float* qq = a.q + ob;
  float* ee = a.e + ob;
  if (VEC && ncol == kTC) {
    constexpr int NB = RMAX * kTC / 4 / kNT;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
      const int e = (tid + k * kNT) * 4, r = e / kTC, cl = e - r * kTC;
      if (r < R) {
        const v4f q = *reinterpret_cast<const v4f*>(s_t + e);
        v4f w;
        w.x = fmaxf(q.x, kMinVal) / (s_sum[0][cl] + s_sum[1][cl]);
        w.y = fmaxf(q.y, kMinVal) / (s_sum[0][cl + 1] + s_sum[1][cl + 1]);
        w.z = fmaxf(q.z, kMinVal) / (s_sum[0][cl + 2] + s_sum[1][cl + 2]);
        w.w = fmaxf(q.w, kMinVal) / (s_sum[0][cl + 3] + s_sum[1][cl + 3]);
        __builtin_nontemporal_store(q, reinterpret_cast<v4f*>(qq + (long long)r * a.N + cl));
        __builtin_nontemporal_store(w, reinterpret_cast<v4f*>(ee + (long long)r * a.N + cl));
      }
    }
  } else {
    for (int e = tid; e < R * kTC; e += kNT) {
      const int r = e / kTC, cl = e - r * kTC;
      if (cl < ncol) {
        const float q = s_t[e];
        __builtin_nontemporal_store(q, qq + (long long)r * a.N + cl);
        __builtin_nontemporal_store(fmaxf(q, kMinVal) / (s_sum[0][cl] + s_sum[1][cl]), ee + (long long)r * a.N + cl);
      }
    }
  }
}

__global__ __launch_bounds__(kNT) void dirichlet_kernel(Args a) {
  const int b = blockIdx.x / a.nblk;
  const long long c = (long long)(blockIdx.x - b * a.nblk) * kNT + threadIdx.x;
  if (c >= a.N) return;  // no barriers below
  int cj[kMaxC];
  col_index(a, c, cj);
  const double lr = a.lr ? (double)a.lr[b] : (double)a.lr_s;
  const float* pp = a.p + b * a.p_bs + c;
  const long long ob = (long long)b * a.R * a.N + c;
  float* qq = a.q + ob;
  float* ee = a.e + ob;
  const int nca = a.has_act ? a.nc - 1 : a.nc;  // belief factors
  double sum = 0.0;
  double acc[kRowsReg];
  for (int r0 = 0; r0 < a.R; r0 += kRowsReg) {
    const int nr = min(kRowsReg, a.R - r0);
#pragma unroll
    for (int r = 0; r < kRowsReg; ++r) acc[r] = 0.0;
    for (int t = 0; t < a.T; ++t) {
      const double w = col_weight(a, b, t, cj, nca);
      if (w == 0.0) continue;
      const float* rd = a.rowd + b * a.row_bs + t * a.row_ts + r0;
#pragma unroll
      for (int r = 0; r < kRowsReg; ++r)
        if (r < nr) acc[r] = fma((double)rd[r], w, acc[r]);
    }
#pragma unroll
    for (int r = 0; r < kRowsReg; ++r)
      if (r < nr) {
        const long long o = (long long)(r0 + r) * a.N;
        const float qv = (float)fma(lr, acc[r], (double)pp[o]);
        qq[o] = qv;
        acc[r] = (double)fmaxf(qv, kMinVal);
        sum += acc[r];
      }
  }
  const double inv = 1.0 / sum;
  if (a.R <= kRowsReg) {
#pragma unroll
    for (int r = 0; r < kRowsReg; ++r)
      if (r < a.R) ee[(long long)r * a.N] = (float)(acc[r] * inv);
  } else {
    // q of the earlier chunks was written by this thread: read it back
    for (int r = 0; r < a.R; ++r) {
      const long long o = (long long)r * a.N;
      ee[o] = (float)((double)fmaxf(qq[o], kMinVal) * inv);
    }
  }
}

static int einval(const std::string& m) {
  set_error("learning: " + m);
  return PYMDP_HIP_EINVAL;
}

static bool al16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

static int launch(Args& a, int64_t batch, hipStream_t s) {
  if (batch <= 0) return einval("batch must be positive");
  if (a.N <= 0 || a.R <= 0) return PYMDP_HIP_OK;
  if (a.R <= 64) {
    const long long nblk = (a.N + kTC - 1) / kTC;
    if (nblk * batch > INT32_MAX) return einval("tensor too large for one launch");
    a.nblk = (int)nblk;
    const bool vec = a.N % 4 == 0 && a.p_bs % 4 == 0 && al16(a.p) && al16(a.q) && al16(a.e);
    const dim3 g((unsigned)(nblk * batch));
    if (a.R <= 32 && vec)
      hipLaunchKernelGGL((dirichlet_tile_kernel<32, true>), g, dim3(kNT), 0, s, a);
    else if (a.R <= 32)
      hipLaunchKernelGGL((dirichlet_tile_kernel<32, false>), g, dim3(kNT), 0, s, a);
    else if (vec)
      hipLaunchKernelGGL((dirichlet_tile_kernel<64, true>), g, dim3(kNT), 0, s, a);
    else
      hipLaunchKernelGGL((dirichlet_tile_kernel<64, false>), g, dim3(kNT), 0, s, a);
    PMH_HIP_CHECK(hipGetLastError(), "learning: ");
    return PYMDP_HIP_OK;
  }
  const long long nblk = (a.N + kNT - 1) / kNT;
  if (nblk * batch > INT32_MAX) return einval("tensor too large for one launch");
  a.nblk = (int)nblk;
  hipLaunchKernelGGL(dirichlet_kernel, dim3((unsigned)(nblk * batch)), dim3(kNT), 0, s, a);
  PMH_HIP_CHECK(hipGetLastError(), "learning: ");
  return PYMDP_HIP_OK;
}

}  // namespace lrn
}  // namespace pmh

using namespace pmh;

extern "C" int pymdp_hip_update_pA(const float* pA, int64_t pA_batch_stride, float* qA_out, float* EA_out,
                                   int64_t batch, int64_t T, int64_t O, int32_t n_deps, const int64_t* dep_sizes,
                                   const float* obs, int64_t obs_batch_stride, const float* const* qs_deps,
                                   const int64_t* qs_batch_stride, const float* lr, float lr_scalar, void* stream) {
  using namespace pmh::lrn;
  set_error("");
  if (!pA || !qA_out || !EA_out || !obs) return einval("pA, qA_out, EA_out and obs must be non-NULL");
  if (T < 0 || O <= 0 || O > INT32_MAX) return einval("invalid T or O");
  if (n_deps < 0 || n_deps > kMaxC) return einval("A dependency count must be in [0, 8]");
  if (n_deps > 0 && (!dep_sizes || !qs_deps || !qs_batch_stride)) return einval("dependency arrays are NULL");
  Args a{};
  a.p = pA;
  a.p_bs = pA_batch_stride;
  a.q = qA_out;
  a.e = EA_out;
  a.R = (int)O;
  a.T = (int)T;
  a.nc = n_deps;
  long long N = 1;
  for (int j = 0; j < n_deps; ++j) {
    if (dep_sizes[j] <= 0 || dep_sizes[j] > INT32_MAX) return einval("dependency sizes must be positive");
    if (!qs_deps[j]) return einval("qs_deps[" + std::to_string(j) + "] is NULL");
    a.n[j] = (int)dep_sizes[j];
    a.col[j] = qs_deps[j];
    a.col_bs[j] = qs_batch_stride[j];
    a.col_ts[j] = dep_sizes[j];
    N *= dep_sizes[j];
  }
  a.N = N;
  a.rowd = obs;
  a.row_bs = obs_batch_stride;
  a.row_ts = O;
  a.lr = lr;
  a.lr_s = lr_scalar;
  return launch(a, batch, static_cast<hipStream_t>(stream));
}

extern "C" int pymdp_hip_update_pB(const float* pB, int64_t pB_batch_stride, float* qB_out, float* EB_out,
                                   int64_t batch, int64_t T1, int64_t S, int32_t n_deps, const int64_t* dep_sizes,
                                   int64_t U, const float* qs_next, int64_t qs_next_batch_stride,
                                   int64_t qs_next_t_stride, const float* const* qs_deps,
                                   const int64_t* qs_deps_batch_stride, const int64_t* qs_deps_t_stride,
                                   const int32_t* actions, int64_t actions_batch_stride, int32_t num_factors,
                                   int32_t factor, int32_t mask_all, const float* lr, float lr_scalar,
                                   void* stream) {
  using namespace pmh::lrn;
  set_error("");
  if (!pB || !qB_out || !EB_out) return einval("pB, qB_out and EB_out must be non-NULL");
  if (T1 < 0 || S <= 0 || S > INT32_MAX || U <= 0 || U > INT32_MAX) return einval("invalid T1, S or U");
  if (n_deps < 1 || n_deps > kMaxC - 1) return einval("B dependency count must be in [1, 7]");
  if (!dep_sizes || !qs_deps || !qs_deps_batch_stride || !qs_deps_t_stride) return einval("dependency arrays are NULL");
  if (T1 > 0 && (!qs_next || !actions)) return einval("qs_next and actions must be non-NULL");
  if (num_factors < 1 || factor < 0 || factor >= num_factors) return einval("factor out of range");
  Args a{};
  a.p = pB;
  a.p_bs = pB_batch_stride;
  a.q = qB_out;
  a.e = EB_out;
  a.R = (int)S;
  a.T = (int)T1;
  a.nc = n_deps + 1;
  long long N = U;
  for (int j = 0; j < n_deps; ++j) {
    if (dep_sizes[j] <= 0 || dep_sizes[j] > INT32_MAX) return einval("dependency sizes must be positive");
    if (T1 > 0 && !qs_deps[j]) return einval("qs_deps[" + std::to_string(j) + "] is NULL");
    a.n[j] = (int)dep_sizes[j];
    a.col[j] = qs_deps[j];
    a.col_bs[j] = qs_deps_batch_stride[j];
    a.col_ts[j] = qs_deps_t_stride[j];
    N *= dep_sizes[j];
  }
  a.n[n_deps] = (int)U;
  a.N = N;
  a.rowd = qs_next;
  a.row_bs = qs_next_batch_stride;
  a.row_ts = qs_next_t_stride;
  a.has_act = 1;
  a.act = actions;
  a.act_bs = actions_batch_stride;
  a.act_ts = num_factors;
  a.act_F = num_factors;
  a.act_f = factor;
  a.mask_all = mask_all ? 1 : 0;
  a.lr = lr;
  a.lr_s = lr_scalar;
  return launch(a, batch, static_cast<hipStream_t>(stream));
}
