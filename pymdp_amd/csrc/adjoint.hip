// Hand-written adjoints (vector-Jacobian products) of the FPI and neg-EFE
// forward kernels, for model fitting through state and policy inference.
//
// FPI (pymdp/algos.py:88-155 with the index-observation log-likelihood of
// pymdp/maths.py:214-287): q = softmax(log_q^(N)), log_q^(i+1) = ln prior +
// M(q^(i)), q^(i) = softmax(log_q^(i)), log_q^(0) = 0, where M_f(q) is the
// leave-f-out contraction of each dependency group's log-likelihood LL_g.
//   unrolled  reverse sweep over the N iterations (the exact derivative of
//             what the forward computes; iterates recomputed in FP64);
//   implicit  at a converged fixed point q* = softmax(ln prior + M(q*)) the
//             adjoint v solves v = q_bar + (dF/dq)^T v (fixed-point
//             iteration); used when the forward residual
//             max|q^(N) - q^(N-1)| is below `implicit_tol`, with the unrolled
//             sweep as the fallback when the forward has not converged or the
//             adjoint iteration does not.
// Gradients w.r.t. A (observed rows only: d/dA ln max(A, MINVAL)) and the
// prior (d/dprior ln max(prior, MINVAL)).
//
// neg-EFE (pymdp/control.py:604-690, horizon 1): G_p = sum_m [H(qo_m) +
// sum_s (sum_o A ln A)[s] w(s) + qo_m . C_m] with w = (x) of B_f[:, :, u_f] qs_f
// over the modality's dependencies and qo_m = A_m w.  Adjoint w.r.t. qs, A,
// B and C given dL/dG.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pymdp_hip.h"
#include "common.hpp"
#include "scratch.hpp"

namespace pmh {
namespace adj {

constexpr int kNT = 256;
constexpr int kNW = kNT / kWave;
constexpr int kF = 8, kM = 8, kR = 3;
constexpr int kS = 256;      // total states of all factors
constexpr int kLL = 2048;    // sum of the dependency groups' joint state counts
constexpr int kEl = kLL / kNT;
constexpr int kQ1 = 2048;    // sum_f U_f * S_f (predicted states per action)
constexpr int kCO = 4096;    // history combinations x outcomes of one modality

struct Desc {
  int F, M, G, totS, iters, mode, P, flags;
  float tol;
  int S[kF], S_off[kF + 1], U[kF], q1_off[kF + 1];
  int O[kM], mod_group[kM], mR[kM], mdep[kM][kR], mK[kM], mC[kM], cstride[kM][kR];
  int gR[kM], gfac[kM][kR], gdim[kM][kR], gK[kM], gll[kM];
  long long A_stride[kM], B_stride[kF], C_stride[kM], prior_stride[kF];
  long long ws_stride;
};

struct Ptrs {
  const float* A[kM];
  const float* B[kF];
  const float* C[kM];
  const float* prior[kF];
  float* Abar[kM];
  float* Bbar[kF];
  float* Cbar[kM];
  const int* obs;
  const float* qbar;
  float* priorbar;
  float* qout;
  int* modeout;
  const float* qs;
  const float* Gbar;
  float* qsbar;
  const int* policy;  // [P, T, F] device
  double* ws;
};

__device__ __forceinline__ double lstab(double x) { return log(fmax(x, (double)kMinVal)); }

// softmax per factor (one wave per factor), src -> dst
__device__ void soft(const Desc& d, const double* src, double* dst, int lane, int wave) {
  for (int f = wave; f < d.F; f += kNW) {
    const int o = d.S_off[f], n = d.S[f];
    const double x = lane < n ? src[o + lane] : -INFINITY;
    const double mx = wave_max(x);
    const double e = lane < n ? exp(x - mx) : 0.0;
    const double s = wave_sum(e);
    if (lane < n) dst[o + lane] = e / s;
  }
}

// softmax backward: out = q * (v - <q, v>) per factor
__device__ void sbp(const Desc& d, const double* q, const double* v, double* out, int lane, int wave) {
  for (int f = wave; f < d.F; f += kNW) {
    const int o = d.S_off[f], n = d.S[f];
    const double qv = lane < n ? q[o + lane] * v[o + lane] : 0.0;
    const double dot = wave_sum(qv);
    if (lane < n) out[o + lane] = q[o + lane] * (v[o + lane] - dot);
  }
}

__device__ __forceinline__ int factor_of(const Desc& d, int i) {
  int f = 0;
  while (f + 1 < d.F && d.S_off[f + 1] <= i) ++f;
  return f;
}

// element index of group g from its multi-index
__device__ __forceinline__ int gidx(const Desc& d, int g, const int* s) {
  int e = 0;
  for (int j = 0; j < d.gR[g]; ++j) e = e * d.gdim[g][j] + s[j];
  return e;
}

constexpr int kGrpF = 8;  // lanes per bin of the leave-one-out contractions

// Partial (lanes part, part + kGrpF, ...) of, for bin i = (f, s_f):
//   lb == nullptr: M_f(q)[s_f] = sum_{g contains f} sum_{s: s_f} LL_g[s] prod_{j != axis f} q_{f_j}[s_j]
//   else (dF/dq)^T: sum_g sum_{a != b} sum_{s: s_b} LL_g[s] lb_{f_a}[s_a] prod_{j != a, b} q_{f_j}[s_j]
// gd holds the digits of every LL element (3 bytes each).
__device__ double bin_contract(const Desc& d, const double* LL, const unsigned char* gd, const double* q,
                               const double* lb, int i, int part) {
  const int f = factor_of(d, i);
  const int sf = i - d.S_off[f];
  double acc = 0.0;
  for (int g = 0; g < d.G; ++g) {
    const int R = d.gR[g];
    int b = -1;
    for (int j = 0; j < R; ++j)
      if (d.gfac[g][j] == f) b = j;
    if (b < 0) continue;
    const int nb = d.gdim[g][b];
    int stride = 1;
    for (int j = b + 1; j < R; ++j) stride *= d.gdim[g][j];
    const float inv = 1.0f / (float)stride;
    const int others = d.gK[g] / nb;
    int qoff[kR];
    for (int j = 0; j < R; ++j) qoff[j] = d.S_off[d.gfac[g][j]];
    for (int r = part; r < others; r += kGrpF) {
      const int hi = (int)(((float)r + 0.5f) * inv);  // r / stride (exact: r < 2^12)
      const int e = d.gll[g] + (hi * nb + sf) * stride + (r - hi * stride);
      const unsigned char* s = gd + 3 * e;
      const double x = LL[e];
      if (!lb) {
        double w = 1.0;
        for (int j = 0; j < R; ++j)
          if (j != b) w *= q[qoff[j] + s[j]];
        acc = fma(x, w, acc);
      } else {
        for (int a = 0; a < R; ++a) {
          if (a == b) continue;
          double w = lb[qoff[a] + s[a]];
          for (int j = 0; j < R; ++j)
            if (j != a && j != b) w *= q[qoff[j] + s[j]];
          acc = fma(x, w, acc);
        }
      }
    }
  }
  return acc;
}

// out[i] = base[i] (or 0) + bin_contract(i) for every bin, kGrpF lanes per bin.
__device__ void all_bins(const Desc& d, const double* LL, const unsigned char* gd, const double* q, const double* lb,
                         const double* base, double* out, int tid) {
  for (int b0 = 0; b0 < d.totS * kGrpF; b0 += kNT) {
    const int slot = b0 + tid;
    const int bin = slot / kGrpF, part = slot % kGrpF;
    double v = bin < d.totS ? bin_contract(d, LL, gd, q, lb, bin, part) : 0.0;
    v = group_sum_low(v, 3);
    if (bin < d.totS && part == 0) out[bin] = (base ? base[bin] : 0.0) + v;
  }
}

// d/dLL_g[s] of sum_f lb_f . M_f(q): sum_a lb_{f_a}[s_a] prod_{j != a} q_{f_j}[s_j]
__device__ __forceinline__ double ll_grad(const Desc& d, int g, const unsigned char* s, const double* q,
                                          const double* lb) {
  const int R = d.gR[g];
  double acc = 0.0;
  for (int a = 0; a < R; ++a) {
    double w = lb[d.S_off[d.gfac[g][a]] + s[a]];
    for (int j = 0; j < R; ++j)
      if (j != a) w *= q[d.S_off[d.gfac[g][j]] + s[j]];
    acc += w;
  }
  return acc;
}

__global__ __launch_bounds__(kNT) void fpi_adjoint_kernel(const Desc d, const Ptrs p) {
  __shared__ double s_LL[kLL];
  __shared__ double s_lp[kS], s_q[kS], s_lq[kS], s_lb[kS], s_v[kS], s_t[kS], s_lpb[kS];
  __shared__ unsigned char s_gd[3 * kLL];
  __shared__ int s_obs[kM];
  __shared__ int s_flag;
  const int b = blockIdx.x, tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid >> 6;
  double* qit = p.ws + (long long)b * d.ws_stride;  // [iters + 1][totS] forward iterates

  for (int f = 0; f < d.F; ++f)
    for (int i = tid; i < d.S[f]; i += kNT)
      s_lp[d.S_off[f] + i] = lstab((double)p.prior[f][(long long)b * d.prior_stride[f] + i]);
  if (tid < d.M) s_obs[tid] = p.obs[(long long)b * d.M + tid];
  for (int i = tid; i < d.totS; i += kNT) {
    s_lq[i] = 0.0;
    s_lpb[i] = 0.0;
  }
  __syncthreads();
  for (int g = 0; g < d.G; ++g)
    for (int e = tid; e < d.gK[g]; e += kNT) {
      double acc = 0.0;
      for (int m = 0; m < d.M; ++m) {
        if (d.mod_group[m] != g) continue;
        const int o = s_obs[m];
        if (o < 0 || o >= d.O[m]) continue;
        acc += lstab((double)p.A[m][(long long)b * d.A_stride[m] + (long long)o * d.gK[g] + e]);
      }
      s_LL[d.gll[g] + e] = acc;
      int ee = e;
      for (int j = d.gR[g] - 1; j >= 0; --j) {
        s_gd[3 * (d.gll[g] + e) + j] = (unsigned char)(ee % d.gdim[g][j]);
        ee /= d.gdim[g][j];
      }
    }
  __syncthreads();

  // forward iterates (FP64): q^(i) = softmax(log_q^(i)), log_q^(i+1) = lp + M(q^(i))
  for (int it = 0; it <= d.iters; ++it) {
    soft(d, s_lq, s_q, lane, wave);
    __syncthreads();
    for (int i = tid; i < d.totS; i += kNT) qit[(long long)it * d.totS + i] = s_q[i];
    if (it == d.iters) break;
    all_bins(d, s_LL, s_gd, s_q, nullptr, s_lp, s_t, tid);
    __syncthreads();
    for (int i = tid; i < d.totS; i += kNT) s_lq[i] = s_t[i];
    __syncthreads();
  }
  // s_q = q^(N); residual against q^(N-1)
  if (tid == 0) s_flag = 0;
  __syncthreads();
  bool use_implicit = false;
  if (d.mode == 1 && d.iters >= 2) {
    double r = 0.0;
    for (int i = tid; i < d.totS; i += kNT)
      r = fmax(r, fabs(s_q[i] - qit[(long long)(d.iters - 1) * d.totS + i]));
    r = wave_max(r);
    if (lane == 0 && r > d.tol) atomicOr(&s_flag, 1);
    __syncthreads();
    use_implicit = s_flag == 0;
  }
  for (int i = tid; i < d.totS; i += kNT) {
    if (p.qout) p.qout[(long long)b * d.totS + i] = (float)s_q[i];
    s_v[i] = (double)p.qbar[(long long)b * d.totS + i];  // q_bar
  }
  __syncthreads();

  double llb[kEl];
#pragma unroll
  for (int k = 0; k < kEl; ++k) llb[k] = 0.0;

  if (use_implicit) {
    // v = q_bar + J^T v at q* (s_q); s_t keeps q_bar
    for (int i = tid; i < d.totS; i += kNT) s_t[i] = s_v[i];
    bool ok = false;
    for (int k = 0; k < 200; ++k) {
      __syncthreads();
      sbp(d, s_q, s_v, s_lb, lane, wave);
      __syncthreads();
      if (tid == 0) s_flag = 0;
      __syncthreads();
      all_bins(d, s_LL, s_gd, s_q, s_lb, s_t, s_lq, tid);
      __syncthreads();
      double dv = 0.0, vm = 0.0;
      for (int i = tid; i < d.totS; i += kNT) {
        dv = fmax(dv, fabs(s_lq[i] - s_v[i]));
        vm = fmax(vm, fabs(s_lq[i]));
      }
      dv = wave_max(dv);
      vm = wave_max(vm);
      if (lane == 0 && dv > 1e-13 + 1e-11 * vm) atomicOr(&s_flag, 1);
      __syncthreads();
      for (int i = tid; i < d.totS; i += kNT) s_v[i] = s_lq[i];
      const bool more = s_flag != 0;
      __syncthreads();
      if (!more) {
        ok = true;
        break;
      }
    }
    if (ok) {
      sbp(d, s_q, s_v, s_lb, lane, wave);
      __syncthreads();
      for (int i = tid; i < d.totS; i += kNT) s_lpb[i] = s_lb[i];
      for (int g = 0; g < d.G; ++g)
#pragma unroll
        for (int k = 0; k < kEl; ++k) {
          const int e = tid + k * kNT - d.gll[g];
          if (e < 0 || e >= d.gK[g]) continue;
          llb[k] += ll_grad(d, g, s_gd + 3 * (tid + k * kNT), s_q, s_lb);
        }
    } else {
      use_implicit = false;
      for (int i = tid; i < d.totS; i += kNT) s_v[i] = s_t[i];
    }
    __syncthreads();
  }
  if (!use_implicit) {
    // unrolled reverse sweep; s_v holds q_bar of the final softmax
    sbp(d, s_q, s_v, s_lb, lane, wave);  // l_bar of log_q^(N)
    __syncthreads();
    for (int it = d.iters; it >= 1; --it) {
      // log_q^(it) = lp + M(q^(it-1))
      for (int i = tid; i < d.totS; i += kNT) {
        s_lpb[i] += s_lb[i];
        s_q[i] = qit[(long long)(it - 1) * d.totS + i];
      }
      __syncthreads();
      for (int g = 0; g < d.G; ++g)
#pragma unroll
        for (int k = 0; k < kEl; ++k) {
          const int e = tid + k * kNT - d.gll[g];
          if (e < 0 || e >= d.gK[g]) continue;
          llb[k] += ll_grad(d, g, s_gd + 3 * (tid + k * kNT), s_q, s_lb);
        }
      if (it == 1) break;  // q^(0) = softmax(0) is constant
      all_bins(d, s_LL, s_gd, s_q, s_lb, nullptr, s_v, tid);
      __syncthreads();
      sbp(d, s_q, s_v, s_lb, lane, wave);
      __syncthreads();
    }
  }
  __syncthreads();
  if (p.modeout && tid == 0) p.modeout[b] = use_implicit ? 1 : 0;
  // dA at the observed rows, dprior
  for (int g = 0; g < d.G; ++g)
#pragma unroll
    for (int k = 0; k < kEl; ++k) {
      const int e = tid + k * kNT - d.gll[g];
      if (e < 0 || e >= d.gK[g]) continue;
      for (int m = 0; m < d.M; ++m) {
        if (d.mod_group[m] != g) continue;
        const int o = s_obs[m];
        if (o < 0 || o >= d.O[m]) continue;
        const long long ix = (long long)b * d.A_stride[m] + (long long)o * d.gK[g] + e;
        const double a = (double)p.A[m][ix];
        p.Abar[m][ix] = a >= (double)kMinVal ? (float)(llb[k] / a) : 0.0f;
      }
    }
  for (int f = 0; f < d.F; ++f)
    for (int i = tid; i < d.S[f]; i += kNT) {
      const double pr = (double)p.prior[f][(long long)b * d.prior_stride[f] + i];
      p.priorbar[(long long)b * d.totS + d.S_off[f] + i] =
          pr >= (double)kMinVal ? (float)(s_lpb[d.S_off[f] + i] / pr) : 0.0f;
    }
}

// ---------------------------------------------------------------------------
// neg-EFE adjoint (horizon 1, factor-local B)
// ---------------------------------------------------------------------------
// LDS (doubles unless noted): q1, q1b [Q]; qs [totS]; HA [Kmax]; SB [Cmax];
// QO, GQ [COmax]; A of the current modality as floats [OKmax]; digit tables
// tables cof [Cmax][3] (uint16: offset in q1 of q1_{d_j}[u_j(c)]) and sd [Kmax][3]
// (bytes: s_j(k)).
constexpr int kGrp = 8;  // lanes per output bin of the dq1 reduction

template <int kMaxEl>
__global__ __launch_bounds__(kNT) void efe_adjoint_kernel(const Desc d, const Ptrs p, int Kmax, int Cmax, int COmax,
                                                          int OKmax, int Bmax) {
  extern __shared__ double sm[];
  const int Q = d.q1_off[d.F];
  double* q1 = sm;
  double* q1b = q1 + Q;
  double* qs = q1b + Q;
  double* HA = qs + d.totS;
  double* SB = HA + Kmax;
  double* QO = SB + Cmax;
  double* GQ = QO + COmax;
  float* As = reinterpret_cast<float*>(GQ + COmax);
  float* Bs = As + ((OKmax + 1) & ~1);          // B_f of one factor [Bmax] (Bmax = 0: read from HBM)
  double* Gb = reinterpret_cast<double*>(Bs + ((Bmax + 1) & ~1));  // dL/dG [P]
  unsigned short* cof = reinterpret_cast<unsigned short*>(Gb + d.P);  // q1 offset of u_j(c) [Cmax][3]
  unsigned char* sd = reinterpret_cast<unsigned char*>(cof + 3 * Cmax);
  unsigned char* pol = sd + 3 * Kmax;           // first actions [P][F]
  const int b = blockIdx.x, tid = threadIdx.x;
  const bool util = d.flags & PYMDP_HIP_FLAG_USE_UTILITY;
  const bool info = d.flags & PYMDP_HIP_FLAG_USE_STATES_INFO_GAIN;

  for (int i = tid; i < d.totS; i += kNT) qs[i] = (double)p.qs[(long long)b * d.totS + i];
  for (int i = tid; i < Q; i += kNT) q1b[i] = 0.0;
  for (int i = tid; i < d.P; i += kNT) Gb[i] = (double)p.Gbar[(long long)b * d.P + i];
  for (int i = tid; i < d.P * d.F; i += kNT) pol[i] = (unsigned char)p.policy[i];
  // predicted states per action: q1_f[u] = B_f[:, :, u] qs_f
  for (int f = 0; f < d.F; ++f) {
    const float* Bg = p.B[f] + (long long)b * d.B_stride[f];
    const int n = d.S[f], U = d.U[f];
    const float* Bf = Bg;
    if (Bmax) {
      __syncthreads();
      for (int e = tid; e < n * n * U; e += kNT) Bs[e] = Bg[e];
      Bf = Bs;
    }
    __syncthreads();
    for (int e = tid; e < U * n; e += kNT) {
      const int u = e / n, s2 = e - u * n;
      double acc = 0.0;
      for (int s = 0; s < n; ++s) acc = fma((double)Bf[(s2 * n + s) * U + u], qs[d.S_off[f] + s], acc);
      q1[d.q1_off[f] + e] = acc;
    }
  }

  for (int m = 0; m < d.M; ++m) {
    const int O = d.O[m], K = d.mK[m], C = d.mC[m], R = d.mR[m];
    const float* Am = p.A[m] + (long long)b * d.A_stride[m];
    const float* Cm = p.C[m] + (long long)b * d.C_stride[m];
    int qo_[kR], nj[kR];
    for (int j = 0; j < R; ++j) {
      qo_[j] = d.q1_off[d.mdep[m][j]];
      nj[j] = d.S[d.mdep[m][j]];
    }
    __syncthreads();
    for (int e = tid; e < O * K; e += kNT) As[e] = Am[e];
    for (int c = tid; c < C; c += kNT) {
      int cc = c;
      for (int j = R - 1; j >= 0; --j) {
        const int U = d.U[d.mdep[m][j]];
        cof[3 * c + j] = (unsigned short)(qo_[j] + (cc % U) * nj[j]);
        cc /= U;
      }
      double acc = 0.0;
      for (int q = 0; q < d.P; ++q) {
        int x = 0;
        for (int j = 0; j < R; ++j) x += (int)pol[q * d.F + d.mdep[m][j]] * d.cstride[m][j];
        if (x == c) acc += Gb[q];
      }
      SB[c] = acc;
    }
    for (int k = tid; k < K; k += kNT) {
      int kk = k;
      for (int j = R - 1; j >= 0; --j) {
        sd[3 * k + j] = (unsigned char)(kk % nj[j]);
        kk /= nj[j];
      }
    }
    __syncthreads();
    // W[c, k] = prod_j q1_{d_j}[u_j(c)][s_j(k)]
    auto W = [&](int c, int k, int skip) {
      double w = 1.0;
      for (int j = 0; j < R; ++j)
        if (j != skip) w *= q1[(int)cof[3 * c + j] + (int)sd[3 * k + j]];
      return w;
    };
    for (int k = tid; k < K; k += kNT) {
      double acc = 0.0;
      for (int o = 0; o < O; ++o) {
        const double a = (double)As[o * K + k];
        if (a != 0.0) acc += a * lstab(a);
      }
      HA[k] = acc;
    }
    // qo[c, o] = sum_k A[o, k] W[c, k]  (neighbouring lanes: same o, next c)
    for (int e = tid; e < C * O; e += kNT) {
      const int o = e / C, c = e - o * C;
      int cb[kR];
      for (int j = 0; j < R; ++j) cb[j] = cof[3 * c + j];
      const float* Ao = As + o * K;
      double a0 = 0.0, a1 = 0.0;
      int k = 0;
      for (; k + 1 < K; k += 2) {
        double w0 = 1.0, w1 = 1.0;
        for (int j = 0; j < R; ++j) {
          w0 *= q1[cb[j] + (int)sd[3 * k + j]];
          w1 *= q1[cb[j] + (int)sd[3 * k + 3 + j]];
        }
        a0 = fma((double)Ao[k], w0, a0);
        a1 = fma((double)Ao[k + 1], w1, a1);
      }
      if (k < K) {
        double w0 = 1.0;
        for (int j = 0; j < R; ++j) w0 *= q1[cb[j] + (int)sd[3 * k + j]];
        a0 = fma((double)Ao[k], w0, a0);
      }
      QO[c * O + o] = a0 + a1;
    }
    __syncthreads();
    for (int e = tid; e < C * O; e += kNT) {
      const int c = e / O, o = e - c * O;
      const double qo = QO[e];
      double g = 0.0;
      if (util) g += (double)Cm[o];
      if (info) g -= qo >= (double)kMinVal ? lstab(qo) + 1.0 : lstab(qo);
      GQ[e] = SB[c] * g;
    }
    __syncthreads();
    if (p.Cbar[m])
      for (int o = tid; o < O; o += kNT) {
        double acc = 0.0;
        if (util)
          for (int c = 0; c < C; ++c) acc = fma(SB[c], QO[c * O + o], acc);
        p.Cbar[m][(long long)b * d.C_stride[m] + o] = (float)acc;
      }
    // dA[o, k] = sum_c (GQ[c, o] + SB[c] (ln A + 1)) W[c, k]
#pragma unroll
    for (int i = 0; i < kMaxEl; ++i) {
      const int e = tid + i * kNT;
      if (e < O * K) {
        const int o = e / K, k = e - o * K;
        double la = 0.0;
        if (info) {
          const double a = (double)As[e];
          la = a >= (double)kMinVal ? lstab(a) + 1.0 : lstab(a);
        }
        int sk[kR];
        for (int j = 0; j < R; ++j) sk[j] = sd[3 * k + j];
        double acc0 = 0.0, acc1 = 0.0;
        int c = 0;
        for (; c + 1 < C; c += 2) {
          double w0 = 1.0, w1 = 1.0;
          for (int j = 0; j < R; ++j) {
            w0 *= q1[(int)cof[3 * c + j] + sk[j]];
            w1 *= q1[(int)cof[3 * c + 3 + j] + sk[j]];
          }
          acc0 = fma(GQ[c * O + o] + SB[c] * la, w0, acc0);
          acc1 = fma(GQ[(c + 1) * O + o] + SB[c + 1] * la, w1, acc1);
        }
        if (c < C) {
          double w0 = 1.0;
          for (int j = 0; j < R; ++j) w0 *= q1[(int)cof[3 * c + j] + sk[j]];
          acc0 = fma(GQ[c * O + o] + SB[c] * la, w0, acc0);
        }
        const double acc = acc0 + acc1;
        p.Abar[m][(long long)b * d.A_stride[m] + e] = (float)acc;
      }
    }
    // dq1_{d_j}[u][s] = sum_{c: u_j(c) = u} sum_{k: s_j(k) = s} wb[c, k] W_{-j}[c, k],
    // wb[c, k] = sum_o A[o, k] GQ[c, o] + SB[c] HA[k]; kGrp lanes per bin
    int nb = 0;
    for (int j = 0; j < R; ++j) nb += d.U[d.mdep[m][j]] * nj[j];
    for (int base = 0; base < nb * kGrp; base += kNT) {
      const int slot = base + tid;
      const int bin = slot / kGrp, part = slot % kGrp;
      double acc = 0.0;
      int j = -1, u = 0, sj = 0;
      if (bin < nb) {
        int t = bin;
        j = 0;
        while (t >= d.U[d.mdep[m][j]] * nj[j]) {
          t -= d.U[d.mdep[m][j]] * nj[j];
          ++j;
        }
        u = t / nj[j];
        sj = t - u * nj[j];
        int stride = 1;
        for (int i = j + 1; i < R; ++i) stride *= nj[i];
        const int others = K / nj[j];
        const int nc = C / d.U[d.mdep[m][j]];
        // combinations with u_j(c) = u: c = (ch * U_j + u) * cstride_j + cl
        const int cstr = d.cstride[m][j];
        for (int w = part; w < nc * others; w += kGrp) {
          const int ci = w / others, r = w - ci * others;
          const int c = ((ci / cstr) * d.U[d.mdep[m][j]] + u) * cstr + (ci % cstr);
          const int k = ((r / stride) * nj[j] + sj) * stride + (r % stride);
          double wb = info ? SB[c] * HA[k] : 0.0;
          for (int o = 0; o < O; ++o) wb = fma((double)As[o * K + k], GQ[c * O + o], wb);
          acc = fma(wb, W(c, k, j), acc);
        }
      }
      acc = group_sum_low(acc, 3);
      if (bin < nb && part == 0) q1b[qo_[j] + u * nj[j] + sj] += acc;
    }
    __syncthreads();
  }
  // dB_f[s2, s, u] = q1b_f[u][s2] qs_f[s];  dqs_f[s] = sum_{u, s2} B_f[s2, s, u] q1b_f[u][s2]
  for (int f = 0; f < d.F; ++f) {
    const float* Bg = p.B[f] + (long long)b * d.B_stride[f];
    float* Bb = p.Bbar[f] + (long long)b * d.B_stride[f];
    const int n = d.S[f], U = d.U[f];
    const float* Bf = Bg;
    __syncthreads();
    if (Bmax) {
      for (int e = tid; e < n * n * U; e += kNT) Bs[e] = Bg[e];
      Bf = Bs;
      __syncthreads();
    }
    for (int e = tid; e < n * n * U; e += kNT) {
      const int u = e % U;
      const int s = (e / U) % n;
      const int s2 = e / (U * n);
      Bb[e] = (float)(q1b[d.q1_off[f] + u * n + s2] * qs[d.S_off[f] + s]);
    }
    // dqs_f[s]: kGrp lanes per s over (u, s2)
    for (int b0 = 0; b0 < n * kGrp; b0 += kNT) {
      const int slot = b0 + tid;
      const int s = slot / kGrp, part = slot % kGrp;
      double acc = 0.0;
      if (s < n)
        for (int w = part; w < U * n; w += kGrp) {
          const int u = w / n, s2 = w - u * n;
          acc = fma((double)Bf[(s2 * n + s) * U + u], q1b[d.q1_off[f] + u * n + s2], acc);
        }
      acc = group_sum_low(acc, 3);
      if (s < n && part == 0) p.qsbar[(long long)b * d.totS + d.S_off[f] + s] = (float)acc;
    }
  }
}

// ---------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------
static int einval(const char* pfx, const std::string& m) {
  set_error(std::string(pfx) + m);
  return PYMDP_HIP_EINVAL;
}

static int fill_common(Desc* d, const int64_t* S, int64_t F, const int64_t* O, int64_t M, const int64_t* Adf,
                       const int64_t* Ado, const char* pfx) {
  std::memset(d, 0, sizeof(Desc));
  if (F < 1 || F > kF) return einval(pfx, "adjoint supports 1..8 factors");
  if (M < 1 || M > kM) return einval(pfx, "adjoint supports 1..8 modalities");
  d->F = (int)F;
  d->M = (int)M;
  for (int f = 0; f < F; ++f) {
    if (S[f] < 1 || S[f] > 64) return einval(pfx, "adjoint supports state counts 1..64");
    d->S[f] = (int)S[f];
    d->S_off[f + 1] = d->S_off[f] + d->S[f];
  }
  d->totS = d->S_off[F];
  if (d->totS > kS) return einval(pfx, "adjoint supports at most 256 states in total");
  std::vector<std::vector<int64_t>> tuples;
  int ll = 0;
  for (int m = 0; m < M; ++m) {
    if (O[m] < 1 || O[m] > 64) return einval(pfx, "adjoint supports outcome counts 1..64");
    d->O[m] = (int)O[m];
    const int64_t a = Ado[m], e = Ado[m + 1];
    if (e - a < 1 || e - a > kR) return einval(pfx, "adjoint supports dependency ranks 1..3");
    std::vector<int64_t> dep(Adf + a, Adf + e);
    int K = 1;
    for (size_t j = 0; j < dep.size(); ++j) {
      if (dep[j] < 0 || dep[j] >= F) return einval(pfx, "A_dependencies factor index out of range");
      for (size_t i = 0; i < j; ++i)
        if (dep[i] == dep[j]) return einval(pfx, "duplicate factor in A_dependencies");
      d->mdep[m][j] = (int)dep[j];
      K *= d->S[dep[j]];
    }
    d->mR[m] = (int)dep.size();
    d->mK[m] = K;
    int g = -1;
    for (size_t k = 0; k < tuples.size(); ++k)
      if (tuples[k] == dep) g = (int)k;
    if (g < 0) {
      tuples.push_back(dep);
      g = (int)tuples.size() - 1;
      d->gR[g] = (int)dep.size();
      for (size_t j = 0; j < dep.size(); ++j) {
        d->gfac[g][j] = (int)dep[j];
        d->gdim[g][j] = d->S[dep[j]];
      }
      d->gK[g] = K;
      d->gll[g] = ll;
      ll += K;
    }
    d->mod_group[m] = g;
  }
  d->G = (int)tuples.size();
  if (ll > kLL) return einval(pfx, "adjoint supports at most 2048 joint states over the dependency groups");
  return 0;
}

}  // namespace adj
}  // namespace pmh

using namespace pmh;
using namespace pmh::adj;

extern "C" int pymdp_hip_fpi_adjoint(const float* const* A_ptrs, const int64_t* A_batch_stride,
                                     const int32_t* obs_idx, const float* const* prior_ptrs,
                                     const int64_t* prior_batch_stride, const float* q_bar,
                                     float* const* A_bar_ptrs, float* prior_bar, float* q_out, int32_t* mode_out,
                                     int64_t batch, const int64_t* S, int64_t F, const int64_t* O, int64_t M,
                                     const int64_t* A_dep_flat, const int64_t* A_dep_offsets, int32_t num_iter,
                                     int32_t mode, float implicit_tol, void* stream) {
  const char* pfx = "fpi_ffi: ";
  set_error("");
  if (batch < 0) return einval(pfx, "batch must be >= 0");
  if (num_iter < 1) return einval(pfx, "num_iter must be >= 1");
  if (!A_ptrs || !A_batch_stride || !obs_idx || !prior_ptrs || !prior_batch_stride || !q_bar || !A_bar_ptrs ||
      !prior_bar)
    return einval(pfx, "null pointer argument");
  Desc d;
  int rc = fill_common(&d, S, F, O, M, A_dep_flat, A_dep_offsets, pfx);
  if (rc) return rc;
  d.iters = (F == 1) ? 1 : num_iter;
  d.mode = mode ? 1 : 0;
  d.tol = implicit_tol;
  Ptrs p;
  std::memset(&p, 0, sizeof(p));
  for (int m = 0; m < M; ++m) {
    if (A_batch_stride[m] == 0 && batch > 1) return einval(pfx, "adjoint needs per-agent A (batch stride > 0)");
    d.A_stride[m] = A_batch_stride[m];
    p.A[m] = A_ptrs[m];
    p.Abar[m] = A_bar_ptrs[m];
  }
  for (int f = 0; f < F; ++f) {
    d.prior_stride[f] = prior_batch_stride[f];
    p.prior[f] = prior_ptrs[f];
  }
  p.obs = obs_idx;
  p.qbar = q_bar;
  p.priorbar = prior_bar;
  p.qout = q_out;
  p.modeout = mode_out;
  if (batch == 0) return 0;
  d.ws_stride = (long long)(d.iters + 1) * d.totS;
  double* ws = nullptr;
  if ((rc = scratch_get(kPoolAdjoint, static_cast<hipStream_t>(stream), sizeof(double) * (size_t)d.ws_stride * (size_t)batch,
                        reinterpret_cast<void**>(&ws), pfx)))
    return rc;
  p.ws = ws;
  if (ensure_log_table()) {
    set_error("fpi_ffi: cannot initialise the log table");
    return PYMDP_HIP_EDEVICE;
  }
  hipLaunchKernelGGL(fpi_adjoint_kernel, dim3((unsigned)batch), dim3(kNT), 0, static_cast<hipStream_t>(stream), d, p);
  PMH_HIP_CHECK(hipGetLastError(), pfx);
  return 0;
}

extern "C" int pymdp_hip_neg_efe_adjoint(const int32_t* policy_dev, int64_t P, int64_t T, const float* qs_flat,
                                         const float* const* A_ptrs, const int64_t* A_batch_stride,
                                         const float* const* B_ptrs, const int64_t* B_batch_stride,
                                         const float* const* C_ptrs, const int64_t* C_batch_stride,
                                         const float* G_bar, float* qs_bar, float* const* A_bar_ptrs,
                                         float* const* B_bar_ptrs, float* const* C_bar_ptrs, int64_t batch,
                                         const int64_t* S, int64_t F, const int64_t* O, int64_t M, const int64_t* U,
                                         const int64_t* A_dep_flat, const int64_t* A_dep_offsets,
                                         const int64_t* B_dep_flat, const int64_t* B_dep_offsets, int32_t flags,
                                         void* stream) {
  const char* pfx = "efe_ffi: ";
  set_error("");
  if (T != 1) return einval(pfx, "the neg-EFE adjoint supports policy horizon 1");
  if (P < 1) return einval(pfx, "P must be positive");
  if (!policy_dev || !qs_flat || !A_ptrs || !B_ptrs || !C_ptrs || !G_bar || !qs_bar || !A_bar_ptrs || !B_bar_ptrs ||
      !C_bar_ptrs)
    return einval(pfx, "null pointer argument");
  if (flags & ~(PYMDP_HIP_FLAG_USE_UTILITY | PYMDP_HIP_FLAG_USE_STATES_INFO_GAIN))
    return einval(pfx, "the neg-EFE adjoint supports the utility and state-information-gain terms only");
  Desc d;
  int rc = fill_common(&d, S, F, O, M, A_dep_flat, A_dep_offsets, pfx);
  if (rc) return rc;
  d.P = (int)P;
  d.flags = flags;
  for (int f = 0; f < F; ++f) {
    if (B_dep_offsets[f + 1] - B_dep_offsets[f] != 1 || B_dep_flat[B_dep_offsets[f]] != f)
      return einval(pfx, "the neg-EFE adjoint supports factor-local B_dependencies ([[0], [1], ...])");
    if (U[f] < 1) return einval(pfx, "U[f] must be positive");
    d.U[f] = (int)U[f];
    d.q1_off[f + 1] = d.q1_off[f] + d.U[f] * d.S[f];
    if (B_batch_stride[f] == 0 && batch > 1) return einval(pfx, "adjoint needs per-agent B (batch stride > 0)");
    d.B_stride[f] = B_batch_stride[f];
  }
  if (d.q1_off[F] > kQ1) return einval(pfx, "adjoint supports sum_f U_f S_f <= 2048");
  for (int m = 0; m < M; ++m) {
    long long C = 1;
    for (int j = d.mR[m] - 1; j >= 0; --j) {
      d.cstride[m][j] = (int)C;
      C *= d.U[d.mdep[m][j]];
    }
    if (C * d.O[m] > kCO) return einval(pfx, "adjoint supports (action combinations x outcomes) <= 4096 per modality");
    d.mC[m] = (int)C;
    if ((A_batch_stride[m] == 0 || C_batch_stride[m] == 0) && batch > 1)
      return einval(pfx, "adjoint needs per-agent A and C (batch stride > 0)");
    d.A_stride[m] = A_batch_stride[m];
    d.C_stride[m] = C_batch_stride[m];
  }
  Ptrs p;
  std::memset(&p, 0, sizeof(p));
  for (int m = 0; m < M; ++m) {
    p.A[m] = A_ptrs[m];
    p.C[m] = C_ptrs[m];
    p.Abar[m] = A_bar_ptrs[m];
    p.Cbar[m] = C_bar_ptrs[m];
  }
  for (int f = 0; f < F; ++f) {
    p.B[f] = B_ptrs[f];
    p.Bbar[f] = B_bar_ptrs[f];
  }
  p.qs = qs_flat;
  p.Gbar = G_bar;
  p.qsbar = qs_bar;
  p.policy = policy_dev;
  if (batch == 0) return 0;
  int Kmax = 1, Cmax = 1, COmax = 1, OKmax = 1;
  for (int m = 0; m < M; ++m) {
    Kmax = std::max(Kmax, d.mK[m]);
    Cmax = std::max(Cmax, d.mC[m]);
    COmax = std::max(COmax, d.mC[m] * d.O[m]);
    OKmax = std::max(OKmax, d.mK[m] * d.O[m]);
  }
  if (OKmax > 16 * kNT) return einval(pfx, "adjoint supports O_m * prod S_deps <= 4096 per modality");
  if (P > 4096) return einval(pfx, "adjoint supports at most 4096 policies");
  for (int f = 0; f < F; ++f)
    if (d.U[f] > 255) return einval(pfx, "adjoint supports at most 255 controls per factor");
  int Bmax = 0;
  for (int f = 0; f < F; ++f) Bmax = std::max(Bmax, d.S[f] * d.S[f] * d.U[f]);
  auto lds_of = [&](int bm) {
    return sizeof(double) * (size_t)(2 * d.q1_off[F] + d.totS + Kmax + Cmax + 2 * COmax + P) +
           sizeof(float) * (size_t)(((OKmax + 1) & ~1) + ((bm + 1) & ~1)) + 6 * (size_t)Cmax + 3 * (size_t)Kmax +
           (size_t)P * F;
  };
  if (lds_of(Bmax) > 64 * 1024) Bmax = 0;  // B read from HBM
  const size_t lds = lds_of(Bmax);
  if (lds > 64 * 1024) return einval(pfx, "adjoint model needs " + std::to_string(lds) + " bytes of LDS (limit 65536)");
  p.ws = nullptr;
  const hipStream_t st = static_cast<hipStream_t>(stream);
  if (OKmax <= 4 * kNT)
    hipLaunchKernelGGL((efe_adjoint_kernel<4>), dim3((unsigned)batch), dim3(kNT), lds, st, d, p, Kmax, Cmax, COmax, OKmax, Bmax);
  else
    hipLaunchKernelGGL((efe_adjoint_kernel<16>), dim3((unsigned)batch), dim3(kNT), lds, st, d, p, Kmax, Cmax, COmax, OKmax, Bmax);
  PMH_HIP_CHECK(hipGetLastError(), pfx);
  return 0;
}
