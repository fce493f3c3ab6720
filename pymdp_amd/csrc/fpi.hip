// Fused mean-field fixed-point iteration (FPI) for gfx950.
//
// Algorithm of record: pymdp/algos.py:88-155 (run_factorized_fpi) with the
// log-likelihood of pymdp/maths.py:214-287.  One workgroup (4 wave64s) per
// agent runs the whole thing: log-likelihood build from A and the
// observation, num_iter Jacobi sweeps, final softmax.
//
// Design (MI355X-first, not a translation of the reference CUDA kernel):
//  * Modalities with the same dependency tuple are summed into one group
//    log-likelihood LL_g (linearity of the leave-one-out contraction), so a
//    sweep costs one contraction per distinct tuple, not per modality.
//  * The leave-one-factor-out contraction of a rank-R group is done as R
//    "stages".  Stage k views the partially contracted tensor X_k as a
//    [rows x n] matrix whose columns are the dependency at position
//    c = R-1-k; it produces both the marginal of that factor
//    M_c[s] = sum_r w_c[r] X_k[r, s]  (w_c = Kronecker product of the q of
//    the leading factors) and the next tensor X_{k+1}[r] = X_k[r,:] . q_c.
//    Cost per sweep ~ 2K instead of R*K.
//  * Lane mapping: a wave64 is split into 64/npad rows of npad lanes
//    (npad = next pow2 >= n).  Row sums are butterflies over the low lane
//    bits; the marginal is a per-lane register accumulator reduced once per
//    stage over the high lane bits and across waves through LDS.
//  * Stage 0 input (the big LL) stays in VGPRs for the whole kernel when it
//    fits (JMAX slots of one element per lane); otherwise it streams from a
//    per-agent global workspace (L2 / MALL resident).  Everything smaller
//    (q, log_q, Kronecker weights, X_{k>=1}) lives in LDS.
//  * Observation rows with zero weight are skipped: for one-hot observations
//    the build reads one row of A per modality (O_m times less HBM traffic),
//    and the skipped terms are exact zeros (0 * finite log).
#include "common.hpp"
#include "scratch.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/pymdp_hip.h"

namespace pmh {

constexpr int kMaxF = 16;
constexpr int kMaxM = 32;
constexpr int kMaxG = 8;
constexpr int kMaxR = PYMDP_HIP_MAX_DEP_RANK;
constexpr int kMaxDim = 64;            // per-stage column width handled by one wave row
constexpr int kLdsBudget = 64 * 1024;  // bytes of dynamic LDS per workgroup

struct FpiDesc {
  int F, M, G, iters;
  int totS, sumO, mode;  // mode 0: ll_flat + lp_flat (log prior); 1: A + obs + prior probs
  int obs_index;         // mode 1: obs holds int32 indices [batch, M] instead of distributions
  int lds_S, lds_O, lds_w, lds_x, lds_x2;
  int S[kMaxF];
  int S_off[kMaxF + 1];
  int mod_group[kMaxM];
  int O[kMaxM];
  int obs_off[kMaxM];
  long long K[kMaxM];
  long long ll_off[kMaxM];
  long long A_stride[kMaxM];
  long long ll_stride;
  long long ws_stride;
  long long prior_stride[kMaxF];
  int R[kMaxG];
  int fac[kMaxG][kMaxR];
  int dim[kMaxG][kMaxR];
  int wt_off[kMaxG][kMaxR];
  int jbase[kMaxG];
  int jcount[kMaxG];
  long long ws_off[kMaxG];
  int scr_global;        // JMAX == 0: weight tables and stage buffers in the workspace
  long long scr_ws_off;  //   at this offset of the agent's workspace row
};

struct FpiPtrs {
  const float* ll;     // mode 0
  const float* lp;     // mode 0: packed log prior [batch, totS]
  const void* obs;     // mode 1: float [batch, sumO] or int32 [batch, M]
  float* q;
  double* ws;          // global workspace (JMAX == 0 only)
  const float* A[kMaxM];
  const float* prior[kMaxF];  // mode 1: per-factor prior probabilities
  // optional timing trace of fpi_fast_kernel (pymdp_hip_fpi_set_trace):
  // per agent iters + 3 wall-clock ticks (100 MHz): start, log-likelihood
  // rows formed, end of each sweep, end
  long long* trace;
};

// Rows of a group's stage whose column position is c: prod_{j<c} dim[j].
__device__ __forceinline__ int stage_rows(const FpiDesc& d, int g, int c) {
  int r = 1;
  for (int j = 0; j < c; ++j) r *= d.dim[g][j];
  return r;
}

template <int NW>
__device__ void softmax_factors(const FpiDesc& d, const double* src, double* dst, int lane, int wave) {
  for (int f = wave; f < d.F; f += NW) {
    const int off = d.S_off[f];
    const int n = d.S[f];
    double m = -INFINITY;
    for (int i = lane; i < n; i += kWave) m = fmax(m, src[off + i]);
    m = wave_max_dpp(m);
    double s = 0.0;
    for (int i = lane; i < n; i += kWave) {
      const double e = exp(src[off + i] - m);
      dst[off + i] = e;
      s += e;
    }
    s = wave_sum_dpp(s);
    for (int i = lane; i < n; i += kWave) dst[off + i] = dst[off + i] / s;
  }
}

// Reduce the per-lane marginal accumulator of one stage and add it into
// log_q[f].  All lanes of the workgroup must call this.
template <int NW>
__device__ __forceinline__ void stage_finish(double acc, int lg, int n, int f, const FpiDesc& d, double* red,
                                             double* log_q, int tid, int lane, int wave) {
  acc = group_sum_high(acc, lg);
  const int col = lane & ((1 << lg) - 1);
  if ((lane >> lg) == 0 && col < n) red[wave * kWave + col] = acc;
  __syncthreads();
  if (tid < n) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) s += red[w * kWave + tid];
    log_q[d.S_off[f] + tid] += s;
  }
}

// The log-likelihood is built from the FP32 data with an FP64-accurate log
// and held in FP64, and every sweep accumulates in FP64: the mean-field
// sweep can oscillate (period-2 cycles of the Jacobi update), and there any
// FP32 rounding of the likelihood or of the iteration state is amplified
// sweep by sweep.
template <int JMAX, int NT>
__global__ __launch_bounds__(NT) void fpi_kernel(const FpiDesc d, const FpiPtrs p) {
  constexpr int NW = NT / kWave;
  extern __shared__ double smd[];
  const int b = blockIdx.x;
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int wave = tid >> 6;

  double* log_q = smd;
  double* qv = log_q + d.lds_S;
  double* lp = qv + d.lds_S;
  double* red0 = lp + d.lds_S;
  double* red1 = red0 + NW * kWave;
  float* obs = reinterpret_cast<float*>(red1 + NW * kWave);
  double* wbuf = (JMAX == 0 && d.scr_global) ? p.ws + (long long)b * d.ws_stride + d.scr_ws_off
                                             : reinterpret_cast<double*>(obs + ((d.lds_O + 1) & ~1));
  double* xa = wbuf + d.lds_w;
  double* xb = xa + d.lds_x;

  // ---- priors and observations --------------------------------------------
  if (d.mode) {
    for (int f = 0; f < d.F; ++f)
      for (int i = tid; i < d.S[f]; i += NT)
        lp[d.S_off[f] + i] = log_stable_f64(p.prior[f][(long long)b * d.prior_stride[f] + i]);
    if (d.obs_index) {
      // one_hot(index): an out-of-range index gives an all-zero row (jax.nn.one_hot).
      const int* oi = static_cast<const int*>(p.obs) + (long long)b * d.M;
      for (int m = 0; m < d.M; ++m)
        for (int o = tid; o < d.O[m]; o += NT) obs[d.obs_off[m] + o] = (oi[m] == o) ? 1.0f : 0.0f;
    } else {
      const float* of = static_cast<const float*>(p.obs) + (long long)b * d.sumO;
      for (int i = tid; i < d.sumO; i += NT) obs[i] = of[i];
    }
  } else {
    for (int i = tid; i < d.totS; i += NT) lp[i] = (double)p.lp[(long long)b * d.totS + i];
  }
  for (int i = tid; i < d.totS; i += NT) log_q[i] = 0.0;
  __syncthreads();

  // ---- log-likelihood build -------------------------------------------------
  double X0[JMAX > 0 ? JMAX : 1];
#pragma unroll
  for (int j = 0; j < (JMAX > 0 ? JMAX : 1); ++j) X0[j] = 0.0;

  if constexpr (JMAX > 0) {
    for (int m = 0; m < d.M; ++m) {
      const int g = d.mod_group[m];
      const int R = d.R[g];
      const int n = d.dim[g][R - 1];
      const int lg = ceil_log2(n);
      const int rpw_lg = 6 - lg;
      const int rows = stage_rows(d, g, R - 1);
      const int jb = d.jbase[g];
      const int je = jb + d.jcount[g];
      const int col = lane & ((1 << lg) - 1);
      const int lr = lane >> lg;
      if (d.mode) {
        const float* Am = p.A[m] + (long long)b * d.A_stride[m];
        const long long K = d.K[m];
        for (int o = 0; o < d.O[m]; ++o) {
          const float w = obs[d.obs_off[m] + o];
          if (w == 0.0f) continue;  // uniform across the workgroup
          const float* row = Am + (long long)o * K;
          // ll_m = sum_o obs[o] * log_stable(A[o]), FP64-accurate log
          constexpr int CH = JMAX < 16 ? JMAX : 16;
#pragma unroll
          for (int j0 = 0; j0 < JMAX; j0 += CH) {
            if (j0 + CH <= jb || j0 >= je) continue;  // uniform
            float t[CH];
#pragma unroll
            for (int i = 0; i < CH; ++i) {
              const int j = j0 + i;
              // every load of the chunk unconditional at a clamped row and
              // column (slots outside [jb, je) included), masked at use
              const int r = (((j - jb) * NW + wave) << rpw_lg) + lr;
              t[i] = __builtin_nontemporal_load(row + (long long)min(max(r, 0), rows - 1) * n + min(col, n - 1));
            }
#pragma unroll
            for (int i = 0; i < CH; ++i) {
              const int j = j0 + i;
              if (j >= jb && j < je) {
                const int r = (((j - jb) * NW + wave) << rpw_lg) + lr;
                const double v = fma((double)w, log_stable_f64(t[i]), X0[j]);
                X0[j] = (r < rows && col < n) ? v : X0[j];
              }
            }
          }
        }
      } else {
        const float* llm = p.ll + (long long)b * d.ll_stride + d.ll_off[m];
#pragma unroll
        for (int j = 0; j < JMAX; ++j) {
          const int r = (((j - jb) * NW + wave) << rpw_lg) + lr;
          const float v = llm[(long long)min(max(r, 0), rows - 1) * n + min(col, n - 1)];
          X0[j] += (j >= jb && j < je && r < rows && col < n) ? (double)v : 0.0;
        }
      }
    }
  } else {
    // Streaming fallback: LL_g (FP64) in natural layout in the per-agent workspace.
    double* ws = p.ws + (long long)b * d.ws_stride;
    for (int g = 0; g < d.G; ++g) {
      long long Kg = 1;
      for (int j = 0; j < d.R[g]; ++j) Kg *= d.dim[g][j];
      double* out = ws + d.ws_off[g];
      for (long long e = tid; e < Kg; e += NT) {
        double acc = 0.0;
        for (int m = 0; m < d.M; ++m) {
          if (d.mod_group[m] != g) continue;
          if (d.mode) {
            const float* Am = p.A[m] + (long long)b * d.A_stride[m];
            for (int o = 0; o < d.O[m]; ++o) {
              const float w = obs[d.obs_off[m] + o];
              if (w == 0.0f) continue;
              acc = fma((double)w, log_stable_f64(Am[(long long)o * d.K[m] + e]), acc);
            }
          } else {
            acc += (double)p.ll[(long long)b * d.ll_stride + d.ll_off[m] + e];
          }
        }
        out[e] = acc;
      }
    }
    __threadfence_block();
    __syncthreads();
  }

  // ---- Jacobi sweeps ---------------------------------------------------------
  for (int it = 0; it < d.iters; ++it) {
    softmax_factors<NW>(d, log_q, qv, lane, wave);
    __syncthreads();
    for (int i = tid; i < d.totS; i += NT) log_q[i] = lp[i];

    for (int g = 0; g < d.G; ++g) {
      const int R = d.R[g];
      // Kronecker weight tables w_c for c = 1..R-1 (w_c has prod_{j<c} dim[j] entries).
      for (int c = 1; c < R; ++c) {
        double* wt = wbuf + d.wt_off[g][c];
        const int rows = stage_rows(d, g, c);
        const int dprev = d.dim[g][c - 1];
        const double* qprev = qv + d.S_off[d.fac[g][c - 1]];
        const double* wprev = (c > 1) ? wbuf + d.wt_off[g][c - 1] : nullptr;
        for (int i = tid; i < rows; i += NT) {
          const int hi = i / dprev;
          const int lo = i - hi * dprev;
          wt[i] = (wprev ? wprev[hi] : 1.0) * qprev[lo];
        }
        __syncthreads();
      }

      double* src = nullptr;
      for (int k = 0; k < R; ++k) {
        const int c = R - 1 - k;
        const int n = d.dim[g][c];
        const int f = d.fac[g][c];
        const int lg = ceil_log2(n);
        const int rpw_lg = 6 - lg;
        const int rows = stage_rows(d, g, c);
        const double* w = (c > 0) ? wbuf + d.wt_off[g][c] : nullptr;
        double* dst = (c > 0) ? ((k & 1) ? xb : xa) : nullptr;
        double* red = (k & 1) ? red1 : red0;
        const int col = lane & ((1 << lg) - 1);
        const int lr = lane >> lg;
        const double qcol = (col < n) ? qv[d.S_off[f] + col] : 0.0;
        double acc = 0.0;

        auto body = [&](double x, int r) {
          const bool rv = r < rows;
          const double wr = rv ? (w ? w[r] : 1.0) : 0.0;
          acc = fma(wr, x, acc);
          double v = group_sum_low(x * qcol, lg);
          if (dst && rv && col == 0) dst[r] = v;
        };

        if (k == 0) {
          if constexpr (JMAX > 0) {
            const int jb = d.jbase[g];
            const int je = jb + d.jcount[g];
#pragma unroll
            for (int j = 0; j < JMAX; ++j) {
              if (j >= jb && j < je) {
                const int r = (((j - jb) * NW + wave) << rpw_lg) + lr;
                body(X0[j], r);
              }
            }
          } else {
            const double* ws = p.ws + (long long)b * d.ws_stride + d.ws_off[g];
            const int steps = (rows + (NW << rpw_lg) - 1) / (NW << rpw_lg);
            for (int s = 0; s < steps; ++s) {
              const int r = ((s * NW + wave) << rpw_lg) + lr;
              const double x = (r < rows && col < n) ? ws[(long long)r * n + col] : 0.0;
              body(x, r);
            }
          }
        } else {
          const int steps = (rows + (NW << rpw_lg) - 1) / (NW << rpw_lg);
          for (int s = 0; s < steps; ++s) {
            const int r = ((s * NW + wave) << rpw_lg) + lr;
            const double x = (r < rows && col < n) ? src[r * n + col] : 0.0;
            body(x, r);
          }
        }
        stage_finish<NW>(acc, lg, n, f, d, red, log_q, tid, lane, wave);
        src = dst;
      }
      __syncthreads();
    }
  }
  __syncthreads();

  // ---- output: q = softmax(log_q) --------------------------------------------
  softmax_factors<NW>(d, log_q, qv, lane, wave);
  __syncthreads();
  for (int i = tid; i < d.totS; i += NT) p.q[(long long)b * d.totS + i] = (float)qv[i];
}

// ===========================================================================
// Single-dependency-tuple fast path: every modality depends on the same
// factors (the reference's default, fully connected A_dependencies, and all
// of the plain workloads).  Row-per-lane layout: lane t owns rows
// r = t + 256 i of the stage-0 matrix [rows0 x n] and keeps them in VGPRs
// (FP64 sums of FP64-accurate logs of the FP32 data); stage 0 then needs
// no cross-lane traffic per element (row sums and weighted column sums are
// both lane-local FP64 FMAs) and one LDS reduction per sweep.  The smaller
// stages run from LDS with thread-per-row sums and thread-per-(column,
// chunk) weighted sums, each with a single reduction.
// ===========================================================================
// X [rows x n] in LDS -> dst[r] = X[r,:] . qc (if dst), log_q_f += sum_r w[r] X[r,:]
template <int kFastNT>
__device__ void lds_stage(const double* X, int rows, int n, const double* qc, const double* w, double* dst,
                          double* red, double* log_q_f, int tid, int lane, int wave) {
  if (dst) {
    for (int r = tid; r < rows; r += kFastNT) {
      double t = 0.0;
      for (int s = 0; s < n; ++s) t = fma(X[r * n + s], qc[s], t);
      dst[r] = t;
    }
  }
  const int lg = ceil_log2(n);
  const int s = tid & ((1 << lg) - 1);
  const int nch = kFastNT >> lg;
  double acc = 0.0;
  if (s < n) {
    for (int r = tid >> lg; r < rows; r += nch) acc = fma(w ? w[r] : 1.0, X[r * n + s], acc);
  }
  acc = group_sum_high(acc, lg);
  if ((lane >> lg) == 0 && s < n) red[wave * kWave + s] = acc;
  __syncthreads();
  if (tid < n) {
    double m = 0.0;
#pragma unroll
    for (int v = 0; v < kFastNT / kWave; ++v) m += red[v * kWave + tid];
    log_q_f[tid] += m;
  }
}

// TRACE: the launch records the timing trace (a separate instantiation, so
// that ordinary launches carry no trace code in the sweep loop).
template <int N0, int RPL, int NT, bool TRACE = false>
__global__ __launch_bounds__(NT) void fpi_fast_kernel(const FpiDesc d, const FpiPtrs p) {
  constexpr int kFastNW = NT / kWave;
  constexpr int LGN = N0 == 4 ? 2 : N0 == 8 ? 3 : 4;  // log2(N0)
  extern __shared__ double smd[];
  const int b = blockIdx.x;
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int wave = tid >> 6;
  // timing trace (diagnostic), written with relaxed atomic stores
  long long* const tr = TRACE ? p.trace + (long long)b * (d.iters + 3) : nullptr;
  auto stamp = [&](int i) {
    if (TRACE && tid == 0)
      __hip_atomic_store(tr + i, (long long)wall_clock64(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  };
  stamp(0);

  double* ltab = smd;  // ln / reciprocal tables for log_f64d
  double* log_q = ltab + 256;
  double* qv = log_q + d.lds_S;
  double* lp = qv + d.lds_S;
  double* wbuf = lp + d.lds_S;
  double* xa = wbuf + d.lds_w;
  double* xb = xa + d.lds_x;
  double* part = xb + d.lds_x2;
  double* red0 = part + N0 * NT;
  double* red1 = red0 + kFastNW * kWave;
  float* obs = reinterpret_cast<float*>(red1 + kFastNW * kWave);

  if (d.mode) {
    for (int f = 0; f < d.F; ++f)
      for (int i = tid; i < d.S[f]; i += NT)
        lp[d.S_off[f] + i] = log_stable_f64(p.prior[f][(long long)b * d.prior_stride[f] + i]);
    if (d.obs_index) {
      const int* oi = static_cast<const int*>(p.obs) + (long long)b * d.M;
      for (int m = 0; m < d.M; ++m)
        for (int o = tid; o < d.O[m]; o += NT) obs[d.obs_off[m] + o] = (oi[m] == o) ? 1.0f : 0.0f;
    } else {
      const float* of = static_cast<const float*>(p.obs) + (long long)b * d.sumO;
      for (int i = tid; i < d.sumO; i += NT) obs[i] = of[i];
    }
  } else {
    for (int i = tid; i < d.totS; i += NT) lp[i] = (double)p.lp[(long long)b * d.totS + i];
  }
  for (int i = tid; i < d.totS; i += NT) log_q[i] = 0.0;
  for (int i = tid; i < 256; i += NT) ltab[i] = g_logtab[i];
  __syncthreads();

  const int R = d.R[0];
  const int n = d.dim[0][R - 1];
  int rows0 = 1;
  for (int j = 0; j < R - 1; ++j) rows0 *= d.dim[0][j];
  const int dprev = (R >= 2) ? d.dim[0][R - 2] : 1;

  // ---- log-likelihood rows (FP64-accurate log per term, FP64 accumulation) --
  double X0[RPL][N0];
#pragma unroll
  for (int i = 0; i < RPL; ++i)
#pragma unroll
    for (int s = 0; s < N0; ++s) X0[i][s] = 0.0;
  if (d.mode && d.obs_index) {
    // One-hot observations: LL[s] = sum_m ln max(A_m[o_m, s], MINVAL)
    //                              = sum_c ln prod_{m in chunk c} max(A_m[o_m, s], MINVAL)
    // over chunks of kProdChunk modalities.  A chunk's FP64 product is at
    // least MINVAL^16 = 2^-832, a normal double (MINVAL^20 = 2^-1040 would
    // already be subnormal, where log_f64d is not valid), so each state costs
    // one FP64 log per 16 modalities instead of one per modality.
    constexpr int kProdChunk = 16;
    const int* oi = static_cast<const int*>(p.obs) + (long long)b * d.M;
#pragma unroll
    for (int i = 0; i < RPL; ++i) {
      const int r = tid + NT * i;
      for (int m0 = 0; m0 < d.M; m0 += kProdChunk) {
        double pr[N0];
#pragma unroll
        for (int s = 0; s < N0; ++s) pr[s] = 1.0;
        const int m1 = min(d.M, m0 + kProdChunk);
        for (int m = m0; m < m1; ++m) {
          const int o = oi[m];
          if (o < 0 || o >= d.O[m]) continue;  // all-zero one_hot row contributes 0
          const float* row = p.A[m] + (long long)b * d.A_stride[m] + (long long)o * d.K[m] + (long long)r * n;
          float t[N0];
          if (N0 % 4 == 0 && n == N0) {
#pragma unroll
            for (int s4 = 0; s4 < N0; s4 += 4) {
              typedef float v4f __attribute__((ext_vector_type(4)));
              v4f v = {1.f, 1.f, 1.f, 1.f};
              if (r < rows0) v = __builtin_nontemporal_load(reinterpret_cast<const v4f*>(row + s4));
              t[s4] = v.x;
              t[s4 + 1] = v.y;
              t[s4 + 2] = v.z;
              t[s4 + 3] = v.w;
            }
          } else {
#pragma unroll
            for (int s = 0; s < N0; ++s) t[s] = (r < rows0 && s < n) ? __builtin_nontemporal_load(row + s) : 1.0f;
          }
#pragma unroll
          for (int s = 0; s < N0; ++s) pr[s] *= (double)fmaxf(t[s], kMinVal);
        }
#pragma unroll
        for (int s = 0; s < N0; ++s) X0[i][s] += (r < rows0 && s < n) ? log_f64d(pr[s], ltab) : 0.0;
      }
    }
  } else {
    for (int m = 0; m < d.M; ++m) {
      const long long K = d.K[m];
      const int nO = d.mode ? d.O[m] : 1;
      for (int o = 0; o < nO; ++o) {
        const float w = d.mode ? obs[d.obs_off[m] + o] : 1.0f;
        if (w == 0.0f) continue;  // uniform
        const float* row = d.mode ? p.A[m] + (long long)b * d.A_stride[m] + (long long)o * K
                                  : p.ll + (long long)b * d.ll_stride + d.ll_off[m];
#pragma unroll
        for (int i = 0; i < RPL; ++i) {
          const int r = tid + NT * i;
#pragma unroll
          for (int s = 0; s < N0; ++s) {
            if (r < rows0 && s < n) {
              const float a = __builtin_nontemporal_load(row + (long long)r * n + s);
              X0[i][s] += d.mode ? (double)w * log_f64d((double)fmaxf(a, kMinVal), ltab) : (double)a;
            }
          }
        }
      }
    }
  }

  if (TRACE) {
    __syncthreads();  // (trace only) the rows' loads have landed
    stamp(1);
  }
  // ---- Jacobi sweeps ---------------------------------------------------------
  const int f_last = d.fac[0][R - 1];
  const int f_prev = (R >= 2) ? d.fac[0][R - 2] : 0;
  int hi[RPL], lo[RPL];  // stage-0 row r = hi * dim[R-2] + lo
#pragma unroll
  for (int i = 0; i < RPL; ++i) {
    const int r = tid + NT * i;
    hi[i] = r / dprev;
    lo[i] = r - hi[i] * dprev;
  }
  for (int it = 0; it < d.iters; ++it) {
    softmax_factors<kFastNW>(d, log_q, qv, lane, wave);
    __syncthreads();
    for (int i = tid; i < d.totS; i += NT) log_q[i] = lp[i];
    // W_c = q_{d0} x ... x q_{d_{c-1}} for c = 2 .. R-2 (W_1 is q_{d0} itself)
    for (int c = 2; c <= R - 2; ++c) {
      double* wt = wbuf + d.wt_off[0][c];
      const int rows = stage_rows(d, 0, c);
      const int dp = d.dim[0][c - 1];
      const double* qp = qv + d.S_off[d.fac[0][c - 1]];
      const double* wp = (c > 2) ? wbuf + d.wt_off[0][c - 1] : qv + d.S_off[d.fac[0][0]];
      for (int i = tid; i < rows; i += NT) {
        const int hi = i / dp;
        wt[i] = wp[hi] * qp[i - hi * dp];
      }
      __syncthreads();
    }
    const double* Wp = (R >= 4) ? wbuf + d.wt_off[0][R - 2] : qv + d.S_off[d.fac[0][0]];  // W_{R-2}
    const double* qprev = qv + d.S_off[f_prev];
    // ---- stage 0 from registers: row sums tr[i] = X0[i,:] . q_last and
    //      weighted column sums part[s] = sum_i w_i X0[i, s]
    double wv[RPL], tr[RPL];
#pragma unroll
    for (int i = 0; i < RPL; ++i) {
      const int r = tid + NT * i;
      double w = 0.0;
      if (r < rows0) w = (R == 1) ? 1.0 : (R == 2) ? qprev[r] : Wp[hi[i]] * qprev[lo[i]];
      wv[i] = w;
      tr[i] = 0.0;
    }
#pragma unroll
    for (int s = 0; s < N0; ++s) {
      const double qs_ = (s < n) ? qv[d.S_off[f_last] + s] : 0.0;
      double a = 0.0;
#pragma unroll
      for (int i = 0; i < RPL; ++i) {
        const double x = X0[i][s];
        tr[i] = fma(x, qs_, tr[i]);
        a = fma(wv[i], x, a);
      }
      part[s * NT + tid] = a;
    }
    if (R >= 2) {
#pragma unroll
      for (int i = 0; i < RPL; ++i) {
        const int r = tid + NT * i;
        if (r < rows0) xa[r] = tr[i];
      }
    }
    __syncthreads();
    {
      // column sums of part [NT x N0]: thread (column, chunk), one butterfly,
      // then across waves
      const int s = tid & (N0 - 1);
      const int nch = NT >> LGN;
      double v = 0.0;
      for (int r = tid >> LGN; r < NT; r += nch) v += part[s * NT + r];
      v = group_sum_high(v, LGN);
      if ((lane >> LGN) == 0) red0[wave * N0 + s] = v;
      __syncthreads();
      if (tid < n) {
        double m = 0.0;
#pragma unroll
        for (int w2 = 0; w2 < kFastNW; ++w2) m += red0[w2 * N0 + tid];
        log_q[d.S_off[f_last] + tid] += m;
      }
    }
    // ---- stages 1 .. R-1 from LDS
    double* src = xa;
    for (int k = 1; k < R; ++k) {
      const int c = R - 1 - k;
      const int nk = d.dim[0][c];
      const int f = d.fac[0][c];
      const int rows = stage_rows(d, 0, c);
      const double* w = (c == 0) ? nullptr : (c == 1) ? qv + d.S_off[d.fac[0][0]] : wbuf + d.wt_off[0][c];
      double* dst = (c > 0) ? ((k & 1) ? xb : xa) : nullptr;
      lds_stage<NT>(src, rows, nk, qv + d.S_off[f], w, dst, (k & 1) ? red1 : red0, log_q + d.S_off[f], tid, lane,
                    wave);
      src = dst;
    }
    __syncthreads();
    stamp(2 + it);
  }

  softmax_factors<kFastNW>(d, log_q, qv, lane, wave);
  __syncthreads();
  for (int i = tid; i < d.totS; i += NT) p.q[(long long)b * d.totS + i] = (float)qv[i];
  stamp(2 + d.iters);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
namespace {

thread_local std::string g_err;

int einval(const char* prefix, const std::string& msg) {
  set_error(std::string(prefix) + msg);
  return PYMDP_HIP_EINVAL;
}

}  // namespace

void set_error(const std::string& msg) { g_err = msg; }
const char* get_error() { return g_err.c_str(); }

// Validate the reference's FPI attributes (pymdp/ffi/fpi/fpi_entry.h:24-82)
// and build the launch descriptor.  ll_sizes[m] is the per-agent element count
// of modality m's log-likelihood.
static int build_fpi_desc(FpiDesc* d, const int64_t* S, int64_t F, int64_t M, const int64_t* lp_offsets,
                          const int64_t* A_dep_flat, const int64_t* A_dep_offsets, int32_t num_iter,
                          int* jslots_out) {
  const char* pfx = "fpi_ffi: ";
  std::memset(d, 0, sizeof(*d));
  if (F <= 0 || M <= 0) return einval(pfx, "invalid F=" + std::to_string(F) + " or M=" + std::to_string(M));
  if (F > kMaxF) return einval(pfx, "F=" + std::to_string(F) + " exceeds the supported " + std::to_string(kMaxF));
  if (M > kMaxM) return einval(pfx, "M=" + std::to_string(M) + " exceeds the supported " + std::to_string(kMaxM));
  if (num_iter < 0) return einval(pfx, "num_iter = " + std::to_string(num_iter) + ", must be non-negative");
  d->F = (int)F;
  d->M = (int)M;
  if (lp_offsets[0] != 0) return einval(pfx, "lp_offsets[0] must be 0");
  for (int64_t f = 0; f < F; ++f) {
    if (S[f] <= 0) return einval(pfx, "S[" + std::to_string(f) + "] must be positive");
    if (lp_offsets[f + 1] - lp_offsets[f] != S[f])
      return einval(pfx, "lp_offsets not consistent with S at factor " + std::to_string(f));
    d->S[f] = (int)S[f];
    d->S_off[f] = (int)lp_offsets[f];
  }
  d->S_off[F] = (int)lp_offsets[F];
  d->totS = (int)lp_offsets[F];

  // Group modalities by identical dependency tuple.
  std::vector<std::vector<int64_t>> tuples;
  if (A_dep_offsets[0] != 0) return einval(pfx, "A_dep_offsets[0] must be 0");
  for (int64_t m = 0; m < M; ++m) {
    const int64_t a = A_dep_offsets[m], e = A_dep_offsets[m + 1];
    if (e < a) return einval(pfx, "A_dep_offsets must be monotonic");
    const int64_t rank = e - a;
    if (rank < 1 || rank > kMaxR) return einval(pfx, "A_dependencies[m] rank must be in [1, 8]");
    std::vector<int64_t> t(A_dep_flat + a, A_dep_flat + e);
    for (int64_t i = 0; i < rank; ++i) {
      if (t[i] < 0 || t[i] >= F) return einval(pfx, "A_dependencies factor index out of range");
      for (int64_t j = 0; j < i; ++j)
        if (t[j] == t[i])
          return einval(pfx, "modality " + std::to_string(m) + " has duplicate factor in A_dependencies");
      if (S[t[i]] > kMaxDim)
        return einval(pfx, "state dimension " + std::to_string(S[t[i]]) + " of a dependency exceeds " +
                               std::to_string(kMaxDim));
    }
    int g = -1;
    for (size_t k = 0; k < tuples.size(); ++k)
      if (tuples[k] == t) g = (int)k;
    if (g < 0) {
      if ((int)tuples.size() == kMaxG) return einval(pfx, "more than 8 distinct A_dependencies tuples");
      tuples.push_back(t);
      g = (int)tuples.size() - 1;
    }
    d->mod_group[m] = g;
    long long K = 1;
    for (int64_t i = 0; i < rank; ++i) K *= S[t[i]];
    d->K[m] = K;
  }
  d->G = (int)tuples.size();
  int jslots = 0;
  long long ws = 0;
  int lds_w = 1, lds_x = 1;
  for (int g = 0; g < d->G; ++g) {
    const auto& t = tuples[g];
    const int R = (int)t.size();
    d->R[g] = R;
    for (int i = 0; i < R; ++i) {
      d->fac[g][i] = (int)t[i];
      d->dim[g][i] = (int)S[t[i]];
    }
    int off = 0, rows = 1;
    for (int c = 1; c < R; ++c) {
      rows *= d->dim[g][c - 1];
      d->wt_off[g][c] = off;
      off += rows;
    }
    lds_w = std::max(lds_w, off);
    lds_x = std::max(lds_x, rows);  // rows of stage 0 = prod_{j<R-1} dim[j]
    const int n = d->dim[g][R - 1];
    d->jbase[g] = 0;
    d->jcount[g] = 0;
    d->ws_off[g] = ws;
    ws += (long long)rows * n;
  }
  d->ws_stride = (ws + 63) / 64 * 64;
  d->lds_S = (d->totS + 3) / 4 * 4;
  d->lds_w = (lds_w + 3) / 4 * 4;
  d->lds_x = (lds_x + 3) / 4 * 4;
  d->iters = num_iter;
  *jslots_out = jslots;
  return 0;
}

static size_t lds_bytes(const FpiDesc& d, int NW) {
  return sizeof(double) * (size_t)(3 * d.lds_S + 2 * NW * kWave + (d.scr_global ? 0 : d.lds_w + 2 * d.lds_x)) +
         sizeof(float) * (size_t)((d.lds_O + 1) & ~1);
}

// Register slots per lane of stage 0 when the workgroup has NW waves.
static int assign_slots(FpiDesc* d, int NW) {
  int jslots = 0;
  for (int g = 0; g < d->G; ++g) {
    const int R = d->R[g];
    int rows = 1;
    for (int c = 0; c < R - 1; ++c) rows *= d->dim[g][c];
    const int rpw = kWave >> ceil_log2(d->dim[g][R - 1]);
    const int steps = (rows + NW * rpw - 1) / (NW * rpw);
    d->jbase[g] = jslots;
    d->jcount[g] = steps;
    jslots += steps;
  }
  return jslots;
}

// Fast-path eligibility and launch configurations.  Candidates are
// (N0 = padded innermost width, RPL = rows per lane, NT = workgroup size) with
// RPL * N0 <= 64 FP64 values per lane and NT in {64, 128, 256, 512}; the
// default rule is in launch_fpi, the autotuner times every candidate once
// per shape (see autotune_fast).
struct FastChoice {
  int N0, RPL, NT;
};

#define PMH_FAST_INSTANCES(X)                                                            \
  X(4, 1, 64) X(4, 2, 64) X(4, 4, 64) X(4, 1, 128) X(4, 2, 128) X(4, 4, 128)                 \
  X(8, 1, 64) X(8, 2, 64) X(8, 4, 64) X(8, 1, 128) X(8, 2, 128) X(8, 4, 128)                 \
  X(16, 1, 64) X(16, 2, 64) X(16, 1, 128) X(16, 2, 128)                                      \
  X(4, 1, 256) X(4, 2, 256) X(4, 4, 256) X(4, 8, 256) X(4, 16, 256) X(4, 4, 512) X(4, 8, 512) \
  X(8, 1, 256) X(8, 2, 256) X(8, 4, 256) X(8, 8, 256) X(8, 2, 512) X(8, 4, 512)              \
  X(16, 1, 256) X(16, 2, 256) X(16, 4, 256) X(16, 1, 512) X(16, 2, 512)

static bool has_fast_instance(int n0, int rpl, int nt) {
#define PMH_HAS(a, b, c) if (n0 == a && rpl == b && nt == c) return true;
  PMH_FAST_INSTANCES(PMH_HAS)
#undef PMH_HAS
  return false;
}

static bool fast_shape(FpiDesc* d, int* n0_out, long long* rows0_out) {
  if (d->G != 1) return false;
  const int R = d->R[0];
  for (int j = 0; j < R; ++j)
    if (d->dim[0][j] > kMaxDim) return false;
  const int n = d->dim[0][R - 1];
  if (n > 16) return false;
  int n0 = 4;
  while (n0 < n) n0 *= 2;
  long long rows0 = 1;
  for (int j = 0; j < R - 1; ++j) rows0 *= d->dim[0][j];
  long long w = 2;
  if (R >= 4) {
    long long rr = 1;
    for (int j = 0; j < R - 2; ++j) rr *= d->dim[0][j];
    w = d->wt_off[0][R - 2] + rr;
  }
  d->lds_w = (int)((w + 1) / 2 * 2);
  d->lds_x = (int)((rows0 + 1) / 2 * 2);
  const long long x2 = (R >= 2) ? rows0 / d->dim[0][R - 2] : 1;
  d->lds_x2 = (int)((x2 + 1) / 2 * 2);
  *n0_out = n0;
  *rows0_out = rows0;
  return true;
}

static size_t fast_lds(const FpiDesc& d, int n0, int nt) {
  return sizeof(double) * (size_t)(256 + 3 * d.lds_S + d.lds_w + d.lds_x + d.lds_x2 + n0 * nt + 2 * nt) +
         sizeof(float) * (size_t)d.lds_O;
}

static std::vector<FastChoice> fast_candidates(const FpiDesc& d, int n0, long long rows0) {
  std::vector<FastChoice> c;
  for (int nt : {64, 128, 256, 512}) {
    int rpl = 1;
    while ((long long)rpl * nt < rows0) rpl *= 2;
    if (rpl * n0 > 64 || !has_fast_instance(n0, rpl, nt)) continue;
    if (fast_lds(d, n0, nt) > (size_t)kLdsBudget) continue;
    c.push_back(FastChoice{n0, rpl, nt});
  }
  return c;
}

// Name of the kernel instance of the most recent FPI launch on this thread.
thread_local char g_last_launch[96] = "";

static int launch_fast(const FpiDesc& d, const FpiPtrs& p, const FastChoice& fc, int64_t batch, hipStream_t stream) {
  const dim3 grid((unsigned)batch);
  snprintf(g_last_launch, sizeof(g_last_launch), "fpi_fast_kernel<%d,%d,%d>", fc.N0, fc.RPL, fc.NT);
  const size_t lds = fast_lds(d, fc.N0, fc.NT);
#define PMH_FAST(n0, rpl, nt)                                                                    \
  if (fc.N0 == n0 && fc.RPL == rpl && fc.NT == nt) {                                             \
    if (p.trace)                                                                                 \
      hipLaunchKernelGGL((fpi_fast_kernel<n0, rpl, nt, true>), grid, dim3(nt), lds, stream, d, p); \
    else                                                                                         \
      hipLaunchKernelGGL((fpi_fast_kernel<n0, rpl, nt>), grid, dim3(nt), lds, stream, d, p);     \
    return 0;                                                                                    \
  }
  PMH_FAST_INSTANCES(PMH_FAST)
#undef PMH_FAST
  return -1;
}

// Per-shape choice cache (key: FNV-1a of the descriptor's shape fields and
// the batch size).  With PYMDP_HIP_AUTOTUNE=1 the first call of a shape on a
// stream that is not capturing times each candidate (3 launches after one
// warm-up, hipEvents on the caller's stream) and keeps the fastest; otherwise
// the default rule picks.
std::mutex g_tune_mu;
std::map<uint64_t, FastChoice> g_tune;

static uint64_t shape_key(const FpiDesc& d, int64_t batch) {
  uint64_t h = 14695981039346656037ull;
  auto mix = [&](long long v) {
    for (int i = 0; i < 8; ++i) {
      h ^= (uint64_t)((v >> (8 * i)) & 0xff);
      h *= 1099511628211ull;
    }
  };
  mix(d.G);
  mix(d.R[0]);
  for (int j = 0; j < d.R[0]; ++j) mix(d.dim[0][j]);
  mix(d.M);
  mix(d.F);
  mix(d.mode);
  mix(d.obs_index);
  mix(d.iters);
  mix(batch);
  return h;
}

static bool autotune_enabled() {
  const char* e = getenv("PYMDP_HIP_AUTOTUNE");
  return e && e[0] == '1';
}

static int autotune_fast(const FpiDesc& d, const FpiPtrs& p, const std::vector<FastChoice>& cands, int64_t batch,
                         hipStream_t stream, FastChoice* best) {
  hipEvent_t e0, e1;
  PMH_HIP_CHECK(hipEventCreate(&e0), "fpi_ffi: ");
  PMH_HIP_CHECK(hipEventCreate(&e1), "fpi_ffi: ");
  float best_ms = 1e30f;
  for (const FastChoice& fc : cands) {
    if (launch_fast(d, p, fc, batch, stream)) continue;
    PMH_HIP_CHECK(hipEventRecord(e0, stream), "fpi_ffi: ");
    for (int r = 0; r < 3; ++r) launch_fast(d, p, fc, batch, stream);
    PMH_HIP_CHECK(hipEventRecord(e1, stream), "fpi_ffi: ");
    PMH_HIP_CHECK(hipEventSynchronize(e1), "fpi_ffi: ");
    float ms = 0.f;
    PMH_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1), "fpi_ffi: ");
    if (ms < best_ms) {
      best_ms = ms;
      *best = fc;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return 0;
}

// fixed_ws / fixed_cap: a caller-provided workspace (the allocation-free
// entry point's fixed scratch) instead of the library's stream scratch;
// shapes that need more than fixed_cap bytes are rejected.
// Diagnostic timing trace of the fast FPI kernel (pymdp_hip_fpi_set_trace):
// later launches of fpi_fast_kernel record per-agent timestamps into it.
static long long* g_fpi_trace = nullptr;

static int launch_fpi(const FpiDesc& d0, FpiPtrs p, int /*unused*/, int64_t batch, hipStream_t stream,
                      double* fixed_ws = nullptr, size_t fixed_cap = 0) {
  const char* pfx = "fpi_ffi: ";
  FpiDesc d = d0;
  p.trace = g_fpi_trace;
  if (batch == 0) return 0;
  if (ensure_log_table()) {
    set_error("fpi_ffi: cannot initialise the log table");
    return PYMDP_HIP_EDEVICE;
  }
  {
    FpiDesc df = d0;
    int n0 = 0;
    long long rows0 = 0;
    if (fast_shape(&df, &n0, &rows0)) {
      const std::vector<FastChoice> cands = fast_candidates(df, n0, rows0);
      if (!cands.empty()) {
        // default: the fewest of 256 / 512 threads with <= 64 FP64 row
        // values per lane (on 1024 agents of 16384 states the 256-thread,
        // 64-value variant took 0.35 ms against 0.54 ms for 512 threads with
        // 32 values each; on 64 agents of 12 states one or two waves per
        // agent were no faster than four, the sweeps' barrier-separated
        // phases set the time); the 64 / 128-thread variants are autotuner
        // candidates
        FastChoice fc = cands.back();
        for (const FastChoice& c : cands)
          if (c.NT >= 256 && c.RPL * c.N0 <= 64) {
            fc = c;
            break;
          }
        const uint64_t key = shape_key(df, batch);
        bool cached = false;
        {
          std::lock_guard<std::mutex> lk(g_tune_mu);
          auto it = g_tune.find(key);
          if (it != g_tune.end()) {
            fc = it->second;
            cached = true;
          }
        }
        if (!cached && !fixed_ws && autotune_enabled() && cands.size() > 1) {
          hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
          if (hipStreamIsCapturing(stream, &cs) == hipSuccess && cs == hipStreamCaptureStatusNone) {
            int rc = autotune_fast(df, p, cands, batch, stream, &fc);
            if (rc) return rc;
            std::lock_guard<std::mutex> lk(g_tune_mu);
            g_tune[key] = fc;
          }
        }
        if (launch_fast(df, p, fc, batch, stream) == 0) {
          PMH_HIP_CHECK(hipGetLastError(), pfx);
          return 0;
        }
      }
    }
  }
  // 4 waves when the log-likelihood fits 16 FP64 slots per lane, else 8 waves
  // (up to 32 slots), else the streaming variant.
  int NW = 4;
  d.scr_global = 0;
  int jslots = assign_slots(&d, 4);
  if (jslots > 16) {
    NW = 8;
    jslots = assign_slots(&d, 8);
  }
  if (lds_bytes(d, NW) > (size_t)kLdsBudget) {
    // weight tables and stage buffers move to the workspace (streaming variant)
    NW = 8;
    jslots = 33;
    d.scr_global = 1;
    d.scr_ws_off = d.ws_stride;
    d.ws_stride = (d.ws_stride + d.lds_w + 2LL * d.lds_x + 63) / 64 * 64;
  }
  const size_t lds = lds_bytes(d, NW);
  if (lds > (size_t)kLdsBudget)
    return einval(pfx, "model needs " + std::to_string(lds) + " bytes of LDS per agent (limit " +
                           std::to_string(kLdsBudget) + ")");
  const dim3 grid((unsigned)batch);
  snprintf(g_last_launch, sizeof(g_last_launch), "fpi_kernel<%d,%d>",
           NW == 4 ? (jslots <= 4 ? 4 : 16) : (jslots <= 32 ? 32 : 0), NW == 4 ? 256 : 512);
  if (NW == 4 && jslots <= 4) {
    hipLaunchKernelGGL((fpi_kernel<4, 256>), grid, dim3(256), lds, stream, d, p);
  } else if (NW == 4) {
    hipLaunchKernelGGL((fpi_kernel<16, 256>), grid, dim3(256), lds, stream, d, p);
  } else if (jslots <= 32) {
    hipLaunchKernelGGL((fpi_kernel<32, 512>), grid, dim3(512), lds, stream, d, p);
  } else {
    double* ws = nullptr;
    const size_t need = sizeof(double) * (size_t)d.ws_stride * (size_t)batch;
    if (fixed_ws) {
      if (need > fixed_cap)
        return einval(pfx, "this call needs " + std::to_string(need) + " bytes of scratch; the allocation-free entry "
                           "point holds at most " + std::to_string(fixed_cap) + " (use fewer agents per call)");
      ws = fixed_ws;
    } else {
      int rc = scratch_get(kPoolFpi, stream, need, reinterpret_cast<void**>(&ws), "fpi_ffi: ");
      if (rc) return rc;
    }
    p.ws = ws;
    hipLaunchKernelGGL((fpi_kernel<0, 512>), grid, dim3(512), lds, stream, d, p);
  }
  PMH_HIP_CHECK(hipGetLastError(), pfx);
  return 0;
}

}  // namespace pmh

using namespace pmh;

extern "C" const char* pymdp_hip_last_error(void) { return pmh::get_error(); }

extern "C" const char* pymdp_hip_version(void) { return "pymdp_hip 0.1.0 gfx950"; }

extern "C" const char* pymdp_hip_last_launch(void) { return pmh::g_last_launch; }

extern "C" int pymdp_hip_fpi(const float* ll_flat, const float* lp_flat, float* q_out, int64_t ll_count,
                             int64_t lp_count, int64_t q_count, const int64_t* S, int64_t F,
                             const int64_t* ll_offsets, int64_t M, const int64_t* lp_offsets,
                             const int64_t* A_dep_flat, const int64_t* A_dep_offsets, int32_t num_iter,
                             void* stream) {
  const char* pfx = "fpi_ffi: ";
  pmh::set_error("");
  if (num_iter <= 0)
    return einval(pfx, "num_iter = " + std::to_string(num_iter) + ", must be positive");
  FpiDesc d;
  int jslots = 0;
  int rc = build_fpi_desc(&d, S, F, M, lp_offsets, A_dep_flat, A_dep_offsets, num_iter, &jslots);
  if (rc) return rc;
  if (ll_offsets[0] != 0) return einval(pfx, "ll_offsets[0] must be 0");
  for (int64_t m = 0; m < M; ++m) {
    if (ll_offsets[m + 1] < ll_offsets[m]) return einval(pfx, "ll_offsets must be monotonic");
    if (ll_offsets[m + 1] - ll_offsets[m] != d.K[m])
      return einval(pfx, "ll_offsets not consistent with S[A_dependencies] at modality " + std::to_string(m));
    d.ll_off[m] = ll_offsets[m];
  }
  d.ll_stride = ll_offsets[M];
  const int64_t totS = d.totS;
  if (lp_count <= 0 || lp_count % totS != 0)
    return einval(pfx, "lp_flat size = " + std::to_string(lp_count) + " not divisible by total_S = " +
                           std::to_string(totS));
  const int64_t batch = lp_count / totS;
  if (ll_count != batch * d.ll_stride)
    return einval(pfx, "ll_flat size = " + std::to_string(ll_count) + ", expected " +
                           std::to_string(batch * d.ll_stride));
  if (q_count != batch * totS)
    return einval(pfx, "q_out size = " + std::to_string(q_count) + ", expected " + std::to_string(batch * totS));
  if (F == 1) d.iters = 1;
  d.mode = 0;
  FpiPtrs p{};
  p.ll = ll_flat;
  p.lp = lp_flat;
  p.q = q_out;
  return launch_fpi(d, p, jslots, batch, (hipStream_t)stream);
}


// ---------------------------------------------------------------------------
// Allocation-free variant of pymdp_hip_infer_states.  The streaming FPI
// variant keeps each agent's FP64 log-likelihood and stage buffers in device
// scratch; here that scratch is this fixed array (one copy per device, held
// until the process ends) instead of library-allocated memory.  Its device
// address is looked up once per device with hipGetSymbolAddress: in host
// code the array's name is not a device address.  The kernels write every
// workspace element before reading it, so nothing is cleared between calls.
// Calls that overlap on one device share the array and race.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

namespace {
constexpr int kNoAllocMaxDevices = 64;
std::mutex g_noalloc_mu;
double* g_noalloc_ptr[kNoAllocMaxDevices] = {};
}  // namespace

static int noalloc_scratch(double** out) {
  const char* pfx = "fpi_ffi: ";
  int dev = 0;
  PMH_HIP_CHECK(hipGetDevice(&dev), pfx);
  if (dev < 0 || dev >= kNoAllocMaxDevices) return einval(pfx, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      set_error(std::string(pfx) + "hipGetSymbolAddress(scratch): " + hipGetErrorString(e));
      return PYMDP_HIP_EDEVICE;
    }
    g_noalloc_ptr[dev] = static_cast<double*>(ptr);
  }
  *out = g_noalloc_ptr[dev];
  return 0;
}

static int infer_states_impl(const float* const* A_ptrs, const int64_t* A_batch_stride,
                                      const void* obs, int32_t obs_is_index,
                                      const float* const* prior_ptrs, const int64_t* prior_batch_stride,
                                      float* q_out, int64_t batch,
                                      const int64_t* S, int64_t F, const int64_t* O, int64_t M,
                                      const int64_t* obs_offsets, const int64_t* lp_offsets,
                                      const int64_t* A_dep_flat, const int64_t* A_dep_offsets, int32_t num_iter,
                                      int32_t method_exact, void* stream, double* fixed_ws, size_t fixed_cap) {
  const char* pfx = "fpi_ffi: ";
  pmh::set_error("");
  if (batch < 0) return einval(pfx, "batch must be non-negative");
  FpiDesc d;
  int jslots = 0;
  int rc = build_fpi_desc(&d, S, F, M, lp_offsets, A_dep_flat, A_dep_offsets, num_iter, &jslots);
  if (rc) return rc;
  if (obs_offsets[0] != 0) return einval(pfx, "obs_offsets[0] must be 0");
  for (int64_t m = 0; m < M; ++m) {
    if (O[m] <= 0) return einval(pfx, "O[" + std::to_string(m) + "] must be positive");
    if (obs_offsets[m + 1] - obs_offsets[m] != O[m]) return einval(pfx, "obs_offsets not consistent with O");
    if (A_ptrs[m] == nullptr) return einval(pfx, "A[" + std::to_string(m) + "] is null");
    if (A_batch_stride[m] < 0) return einval(pfx, "negative A batch stride");
    d.O[m] = (int)O[m];
    d.obs_off[m] = (int)obs_offsets[m];
    d.A_stride[m] = A_batch_stride[m];
  }
  d.sumO = (int)obs_offsets[M];
  d.lds_O = (d.sumO + 3) / 4 * 4;
  d.mode = 1;
  d.obs_index = obs_is_index ? 1 : 0;
  if (method_exact) d.iters = 1;
  if (F == 1) d.iters = 1;  // closed form, pymdp/algos.py:101-104
  FpiPtrs p{};
  p.obs = obs;
  p.q = q_out;
  for (int64_t m = 0; m < M; ++m) p.A[m] = A_ptrs[m];
  for (int64_t f = 0; f < F; ++f) {
    if (prior_ptrs[f] == nullptr) return einval(pfx, "prior[" + std::to_string(f) + "] is null");
    if (prior_batch_stride[f] < 0) return einval(pfx, "negative prior batch stride");
    p.prior[f] = prior_ptrs[f];
    d.prior_stride[f] = prior_batch_stride[f];
  }
  return launch_fpi(d, p, jslots, batch, (hipStream_t)stream, fixed_ws, fixed_cap);
}

extern "C" int pymdp_hip_fpi_set_trace(int64_t* trace) {
  g_fpi_trace = reinterpret_cast<long long*>(trace);
  return PYMDP_HIP_OK;
}

extern "C" int pymdp_hip_infer_states(const float* const* A_ptrs, const int64_t* A_batch_stride,
                                      const void* obs, int32_t obs_is_index,
                                      const float* const* prior_ptrs, const int64_t* prior_batch_stride,
                                      float* q_out, int64_t batch,
                                      const int64_t* S, int64_t F, const int64_t* O, int64_t M,
                                      const int64_t* obs_offsets, const int64_t* lp_offsets,
                                      const int64_t* A_dep_flat, const int64_t* A_dep_offsets, int32_t num_iter,
                                      int32_t method_exact, void* stream) {
  return infer_states_impl(A_ptrs, A_batch_stride, obs, obs_is_index, prior_ptrs, prior_batch_stride, q_out, batch, S, F, O, M, obs_offsets, lp_offsets, A_dep_flat, A_dep_offsets, num_iter, method_exact, stream, nullptr, 0);
}

extern "C" int pymdp_hip_infer_states_noalloc(const float* const* A_ptrs, const int64_t* A_batch_stride,
                                      const void* obs, int32_t obs_is_index,
                                      const float* const* prior_ptrs, const int64_t* prior_batch_stride,
                                      float* q_out, int64_t batch,
                                      const int64_t* S, int64_t F, const int64_t* O, int64_t M,
                                      const int64_t* obs_offsets, const int64_t* lp_offsets,
                                      const int64_t* A_dep_flat, const int64_t* A_dep_offsets, int32_t num_iter,
                                      int32_t method_exact, void* stream) {
  pmh::set_error("");
  double* ws = nullptr;
  const int rc = noalloc_scratch(&ws);
  if (rc) return rc;
  return infer_states_impl(A_ptrs, A_batch_stride, obs, obs_is_index, prior_ptrs, prior_batch_stride, q_out, batch, S, F, O, M, obs_offsets, lp_offsets, A_dep_flat, A_dep_offsets, num_iter, method_exact, stream, ws, kNoAllocScratchBytes);
}
