// Sequence inference over a window of T time steps: marginal message passing
// (MMP) and variational message passing (VMP), pymdp/algos.py:157-500 with
// the masks and action conditioning of pymdp/inference.py:56-271, for
// factor-local transitions (B_dependencies[f] == [f]).
//
// Per iteration, for every step t and factor f:
//   ln_A[t][f]  = sum over modalities of the leave-one-out contraction of
//                 ll_m[t] with q[t] (pymdp/algos.py:30-51)
//   fwd[t][f]   = ln prior_f at a chain start, else
//                 MMP: ln (B_f[:, :, u_{t-1}] q_f[t-1])        VMP: ln B_f[:, :, u_{t-1}] q_f[t-1]
//                 (MMP: times 1/2 when the transition t -> t+1 is valid)
//   bwd[t][f]   = 0 at a chain end, else
//                 MMP: 1/2 ln (q_f[t+1] B~)  with B~ = B[:, :, u_t] normalised over its last axis
//                 VMP: q_f[t+1] ln B[:, :, u_t]
//   ln q'       = ln q + tau (ln_A - ln q + bwd + fwd);  q = softmax(ln q' - mean ln q')
// with q starting uniform.  A chain starts at t = 0 and, when the window
// has a valid-step count (pymdp/inference.py:100-126), after every invalid
// transition (a negative action, or a step before the valid suffix; masked
// steps carry no observation).  Without it every transition counts and a
// negative action selects the neutral transition, uniform over next states
// (pymdp/inference.py:128-171).
//
// One workgroup per agent keeps the whole window in LDS (q, ln q, messages
// and the log-likelihoods of every step, FP64); B is read from global memory
// (L2-resident: a few KB per agent).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>

#include "../../include/pymdp_hip.h"
#include "common.hpp"

namespace pmh {
namespace seq {

constexpr int kNT = 256;
constexpr int kNW = kNT / kWave;
constexpr int kMaxF = 8, kMaxM = 8, kMaxR = 4, kMaxT = 64, kMaxS = 64;
constexpr int kMaxTS = 1024;  // T * sum S
constexpr int kMaxTK = 4096;  // T * sum_m K_m

struct Desc {
  int F, M, T, totS, vmp, num_iter, obs_index, has_actions;
  float tau;
  int S[kMaxF], S_off[kMaxF + 1], U[kMaxF];
  int O[kMaxM], R[kMaxM], dep[kMaxM][kMaxR], dim[kMaxM][kMaxR];
  int K[kMaxM], K_off[kMaxM + 1];
  long long obs_off[kMaxM + 1], obs_stride;
  const float* A[kMaxM];
  long long A_bs[kMaxM];
  const float* B[kMaxF];
  long long B_bs[kMaxF];
  const float* prior[kMaxF];
  long long prior_bs[kMaxF];
  const void* obs;
  const int* actions;
  const int* valid_steps;
  float* out;
};

__device__ __forceinline__ double ln_clip(double x) { return log(fmax(x, (double)kMinVal)); }

// B_f[s, s2, u] of the transition conditioned on action u; an action outside
// [0, U_f) is the neutral transition 1 / S_f
__device__ __forceinline__ double bval(const float* Bf, int Sf, int Uf, int s, int s2, int u) {
  return (u >= 0 && u < Uf) ? (double)Bf[((long long)s * Sf + s2) * Uf + u] : 1.0 / Sf;
}

template <bool VMP>
__global__ __launch_bounds__(kNT) void seq_kernel(Desc d) {
  __shared__ double s_q[kMaxTS], s_lnq[kMaxTS], s_la[kMaxTS], s_fwd[kMaxTS], s_bwd[kMaxTS], s_rs[kMaxTS];
  __shared__ double s_ll[kMaxTK];
  __shared__ double s_lp[kMaxTS / 2];
  __shared__ int s_act[kMaxT * kMaxF];
  __shared__ unsigned char s_ov[kMaxT], s_tv[kMaxT];
  const int b = blockIdx.x, tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const int T = d.T, totS = d.totS, F = d.F;
  const int TS = T * totS;

  // masks (pymdp/inference.py:100-126; all valid without a valid-step
  // count) and actions
  for (int i = tid; i < (T - 1) * F; i += kNT) s_act[i] = d.has_actions ? d.actions[(long long)b * (T - 1) * F + i] : 0;
  __syncthreads();
  for (int t = tid; t < T; t += kNT) {
    bool ov = true, tv = t < T - 1;
    if (d.valid_steps) {
      const int k_valid = min(max(d.valid_steps[b], 1), T);
      ov = t >= T - k_valid;
      tv = tv && ov;
      for (int f = 0; f < F && tv; ++f) tv = s_act[t * F + f] >= 0;
    }
    s_ov[t] = ov;
    s_tv[t] = tv;
  }
  for (int f = 0; f < F; ++f) {
    const float* p = d.prior[f] + b * d.prior_bs[f];
    for (int i = tid; i < d.S[f]; i += kNT) s_lp[d.S_off[f] + i] = ln_clip((double)p[i]);
  }
  for (int i = tid; i < TS; i += kNT) {
    int f = 0;
    const int s = i % totS;
    while (s >= d.S_off[f + 1]) ++f;
    s_q[i] = 1.0 / d.S[f];
  }
  __syncthreads();
  // log-likelihood of every step, masked by obs validity
  const int TK = T * d.K_off[d.M];
  for (int i = tid; i < TK; i += kNT) {
    const int t = i / d.K_off[d.M], r = i - t * d.K_off[d.M];
    int m = 0;
    while (r >= d.K_off[m + 1]) ++m;
    const int k = r - d.K_off[m];
    const float* Am = d.A[m] + b * d.A_bs[m];
    double ll = 0.0;
    if (s_ov[t]) {
      if (d.obs_index) {
        const int o = static_cast<const int*>(d.obs)[((long long)b * T + t) * d.M + m];
        if (o >= 0 && o < d.O[m]) ll = ln_clip((double)Am[(long long)o * d.K[m] + k]);
      } else {
        const float* od = static_cast<const float*>(d.obs) + b * d.obs_stride + (long long)t * d.obs_off[d.M] +
                          d.obs_off[m];
        for (int o = 0; o < d.O[m]; ++o)
          if (od[o] != 0.0f) ll += (double)od[o] * ln_clip((double)Am[(long long)o * d.K[m] + k]);
      }
    }
    s_ll[i] = ll;
  }
  // MMP: row sums of B_f[:, :, u_t] over the last axis, for B~
  if (!VMP) {
    for (int i = tid; i < (T - 1) * totS; i += kNT) {
      const int t = i / totS, s = i - t * totS;
      int f = 0;
      while (s >= d.S_off[f + 1]) ++f;
      const int Sf = d.S[f], sl = s - d.S_off[f], u = s_act[t * F + f];
      const float* Bf = d.B[f] + b * d.B_bs[f];
      double rs = 0.0;
      for (int s2 = 0; s2 < Sf; ++s2) rs += bval(Bf, Sf, d.U[f], sl, s2, u);
      s_rs[i] = fmax(rs, (double)kMinVal);
    }
  }
  __syncthreads();

  for (int it = 0; it < d.num_iter; ++it) {
    // ln q, messages, leave-one-out contractions: one thread per (t, s)
    for (int i = tid; i < TS; i += kNT) {
      const int t = i / totS, s = i - t * totS;
      int f = 0;
      while (s >= d.S_off[f + 1]) ++f;
      const int Sf = d.S[f], sl = s - d.S_off[f], Uf = d.U[f];
      const float* Bf = d.B[f] + b * d.B_bs[f];
      s_lnq[i] = ln_clip(s_q[i]);
      // forward
      const bool start = t == 0 || !s_tv[t - 1];
      double fw;
      if (start) {
        fw = s_lp[s];
      } else {
        const int u = s_act[(t - 1) * F + f];
        const double* qp = s_q + (t - 1) * totS + d.S_off[f];
        double acc = 0.0;
        for (int s2 = 0; s2 < Sf; ++s2) {
          const double bv = bval(Bf, Sf, Uf, sl, s2, u);
          acc += VMP ? ln_clip(bv) * qp[s2] : bv * qp[s2];
        }
        fw = VMP ? acc : ln_clip(acc);
      }
      if (!VMP && t < T - 1 && s_tv[t]) fw *= 0.5;
      s_fwd[i] = fw;
      // backward
      double bw = 0.0;
      if (t < T - 1 && s_tv[t]) {
        const int u = s_act[t * F + f];
        const double* qn = s_q + (t + 1) * totS + d.S_off[f];
        double acc = 0.0;
        for (int s2 = 0; s2 < Sf; ++s2) {
          const double bv = bval(Bf, Sf, Uf, s2, sl, u);
          acc += VMP ? qn[s2] * ln_clip(bv) : qn[s2] * (bv / s_rs[t * totS + d.S_off[f] + s2]);
        }
        bw = VMP ? acc : 0.5 * ln_clip(acc);
      }
      s_bwd[i] = bw;
      // leave-one-out contraction of every modality that depends on f
      double la = 0.0;
      const double* qt = s_q + t * totS;
      for (int m = 0; m < d.M; ++m) {
        const int R = d.R[m];
        int j = -1;
        for (int jj = 0; jj < R; ++jj)
          if (d.dep[m][jj] == f) j = jj;
        if (j < 0) continue;
        const double* ll = s_ll + t * d.K_off[d.M] + d.K_off[m];
        const int n_other = d.K[m] / d.dim[m][j];
        double acc = 0.0;
        for (int c = 0; c < n_other; ++c) {
          // mixed-radix digits of c over the dims other than j, with s at j
          int rest = c, k = 0, stride = 1;
          double w = 1.0;
          for (int jj = R - 1; jj >= 0; --jj) {
            int digit;
            if (jj == j) {
              digit = sl;
            } else {
              digit = rest % d.dim[m][jj];
              rest /= d.dim[m][jj];
              w *= qt[d.S_off[d.dep[m][jj]] + digit];
            }
            k += digit * stride;
            stride *= d.dim[m][jj];
          }
          acc += ll[k] * w;
        }
        la += acc;
      }
      s_la[i] = la;
    }
    __syncthreads();
    // mirror-descent step and softmax: one wave per (t, f) row
    for (int row = wave; row < T * F; row += kNW) {
      const int t = row / F, f = row - t * F, Sf = d.S[f];
      const int i = t * totS + d.S_off[f] + lane;
      double l = 0.0;
      if (lane < Sf) l = s_lnq[i] + (double)d.tau * (s_la[i] - s_lnq[i] + s_bwd[i] + s_fwd[i]);
      const double mean = wave_sum(lane < Sf ? l : 0.0) / Sf;
      const double x = lane < Sf ? l - mean : -INFINITY;
      const double mx = wave_max(x);
      const double e = lane < Sf ? exp(x - mx) : 0.0;
      const double se = wave_sum(e);
      if (lane < Sf) s_q[i] = e / se;
    }
    __syncthreads();
  }
  float* out = d.out + (long long)b * TS;
  for (int i = tid; i < TS; i += kNT) out[i] = (float)s_q[i];
}

static int einval(const std::string& m) {
  set_error("seq_ffi: " + m);
  return PYMDP_HIP_EINVAL;
}

}  // namespace seq
}  // namespace pmh

using namespace pmh;

extern "C" int pymdp_hip_sequence_infer(int32_t method, const float* const* A_ptrs, const int64_t* A_batch_stride,
                                        const float* const* B_ptrs, const int64_t* B_batch_stride, const void* obs,
                                        int32_t obs_is_index, const float* const* prior_ptrs,
                                        const int64_t* prior_batch_stride, const int32_t* actions,
                                        const int32_t* valid_steps, float* q_out, int64_t batch, int64_t T,
                                        const int64_t* S, const int64_t* U, int64_t F, const int64_t* O, int64_t M,
                                        const int64_t* obs_offsets, const int64_t* A_dep_flat,
                                        const int64_t* A_dep_offsets, int32_t num_iter, float tau, void* stream) {
  using namespace pmh::seq;
  set_error("");
  if (method != 0 && method != 1) return einval("method must be 0 (mmp) or 1 (vmp)");
  if (!A_ptrs || !A_batch_stride || !B_ptrs || !B_batch_stride || !obs || !prior_ptrs || !prior_batch_stride ||
      !q_out || !S || !U || !O || !A_dep_flat || !A_dep_offsets)
    return einval("NULL argument");
  if (!obs_is_index && !obs_offsets) return einval("obs_offsets is NULL");
  if (batch <= 0 || batch > INT32_MAX) return einval("batch must be in [1, 2^31)");
  if (T < 1 || T > kMaxT) return einval("window length T must be in [1, 64]");
  if (F < 1 || F > kMaxF) return einval("number of factors must be in [1, 8]");
  if (M < 1 || M > kMaxM) return einval("number of modalities must be in [1, 8]");
  if (num_iter < 0) return einval("num_iter must be >= 0");
  Desc d{};
  d.F = (int)F;
  d.M = (int)M;
  d.T = (int)T;
  d.vmp = method;
  d.num_iter = num_iter;
  d.tau = tau;
  d.obs_index = obs_is_index ? 1 : 0;
  d.has_actions = actions != nullptr;
  d.S_off[0] = 0;
  for (int f = 0; f < F; ++f) {
    if (S[f] < 1 || S[f] > kMaxS) return einval("state counts must be in [1, 64]");
    if (U[f] < 1) return einval("control counts must be positive");
    if (!actions && T > 1 && U[f] != 1)
      return einval("actions are required when a transition tensor has more than one control state");
    d.S[f] = (int)S[f];
    d.U[f] = (int)U[f];
    d.S_off[f + 1] = d.S_off[f] + d.S[f];
    d.B[f] = B_ptrs[f];
    d.B_bs[f] = B_batch_stride[f];
    d.prior[f] = prior_ptrs[f];
    d.prior_bs[f] = prior_batch_stride[f];
  }
  d.totS = d.S_off[F];
  if ((long long)T * d.totS > kMaxTS) return einval("T * sum(S) exceeds 1024");
  if (A_dep_offsets[0] != 0) return einval("A_dep_offsets[0] must be 0");
  d.K_off[0] = 0;
  for (int m = 0; m < M; ++m) {
    const int64_t r0 = A_dep_offsets[m], r1 = A_dep_offsets[m + 1];
    if (r1 <= r0 || r1 - r0 > kMaxR) return einval("modality dependency count must be in [1, 4]");
    d.R[m] = (int)(r1 - r0);
    long long K = 1;
    for (int j = 0; j < d.R[m]; ++j) {
      const int64_t f = A_dep_flat[r0 + j];
      if (f < 0 || f >= F) return einval("A dependency out of range");
      d.dep[m][j] = (int)f;
      d.dim[m][j] = (int)S[f];
      K *= S[f];
    }
    if (O[m] < 1) return einval("O must be positive");
    if ((long long)d.K_off[m] + K > kMaxTK) return einval("joint state count too large");
    d.O[m] = (int)O[m];
    d.K[m] = (int)K;
    d.K_off[m + 1] = d.K_off[m] + (int)K;
    d.A[m] = A_ptrs[m];
    d.A_bs[m] = A_batch_stride[m];
  }
  if ((long long)T * d.K_off[M] > kMaxTK) return einval("T * sum_m prod S[deps] exceeds 4096");
  for (int m = 0; m <= M; ++m) d.obs_off[m] = obs_is_index ? m : obs_offsets[m];
  d.obs_stride = obs_is_index ? T * M : T * obs_offsets[M];
  d.obs = obs;
  d.actions = actions;
  d.valid_steps = valid_steps;
  d.out = q_out;
  const dim3 g((unsigned)batch);
  if (method == 1)
    hipLaunchKernelGGL(seq_kernel<true>, g, dim3(kNT), 0, static_cast<hipStream_t>(stream), d);
  else
    hipLaunchKernelGGL(seq_kernel<false>, g, dim3(kNT), 0, static_cast<hipStream_t>(stream), d);
  PMH_HIP_CHECK(hipGetLastError(), "seq_ffi: ");
  return PYMDP_HIP_OK;
}
