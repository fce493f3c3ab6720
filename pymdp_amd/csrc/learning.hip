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
// One fused pass: a thread owns one column, holds up to 64 rows of it in
// registers, and reads p once and writes q and e once (12 bytes per element:
// HBM-bound).  Rows beyond 64 are written in chunks and re-read for e.
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

__global__ __launch_bounds__(kNT) void dirichlet_kernel(Args a) {
  const int b = blockIdx.x / a.nblk;
  const long long c = (long long)(blockIdx.x - b * a.nblk) * kNT + threadIdx.x;
  if (c >= a.N) return;  // no barriers below
  int cj[kMaxC];
  {
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
      double w = 1.0;
      if (a.has_act) {
        const int* at = a.act + b * a.act_bs + t * a.act_ts;
        bool ok = at[a.act_f] == cj[a.nc - 1];
        if (a.mask_all)
          for (int f = 0; f < a.act_F; ++f) ok = ok && at[f] >= 0;
        if (!ok) continue;
      }
#pragma unroll
      for (int j = 0; j < kMaxC; ++j)
        if (j < nca) w *= (double)a.col[j][b * a.col_bs[j] + t * a.col_ts[j] + cj[j]];
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

static int launch(Args& a, int64_t batch, hipStream_t s) {
  if (batch <= 0) return einval("batch must be positive");
  if (a.N <= 0 || a.R <= 0) return PYMDP_HIP_OK;
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
