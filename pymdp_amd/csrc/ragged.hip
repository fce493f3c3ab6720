// Batched ragged agents: one step of infer_states -> infer_policies ->
// sample_action for thousands of agents whose factor, modality, control and
// policy counts all differ, as three stream-ordered launches that a caller can
// capture in one hipGraph.
//
//   rg_fpi_kernel       persistent; a workgroup takes the next agent from an
//                       atomic queue ordered by decreasing cost, runs the
//                       fused log-likelihood + Jacobi FPI of
//                       pymdp/algos.py:88-155 and then the action-history
//                       rollout of pymdp/control.py:351-389 (B . q per
//                       distinct history), writing the posterior (FP32 out,
//                       FP64 internal) and every history vector.
//   rg_efe_kernel       persistent; work items are (agent, modality, level,
//                       chunk of first-axis histories), largest first.  An item
//                       streams its modality's A once through LDS tiles, forms
//                       sum_o A ln A on the fly, contracts [A ; that row] with
//                       every history of each dependency in turn and scores
//                       every history combination: entropy of qo + ambiguity
//                       + qo . C (pymdp/control.py:420-509).
//   rg_assemble_kernel  one workgroup per agent: G[p] = sum_t sum_m of the
//                       combination scores its histories select, q_pi =
//                       softmax(gamma G + ln E) (pymdp/control.py:963), the
//                       first-index argmax and that policy's first actions
//                       (pymdp/control.py:115-150).
//
// The plan (host) lays out one packed arena for all agents' parameters, one
// for observation indices, one for outputs; it allocates every device buffer
// the step needs once, so a step allocates nothing and never synchronises.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/pymdp_hip.h"
#include "common.hpp"
#include "history.hpp"

namespace pmh {
namespace rg {

constexpr int kF = 8;     // factors per agent
constexpr int kM = 8;     // modalities per agent
constexpr int kR = 3;     // dependency rank of a modality
constexpr int kRB = 3;    // dependency rank of a transition
constexpr int kT = 4;     // policy horizon
constexpr int kDim = 64;  // state / outcome count
constexpr int kNT = 256;
constexpr int kNW = kNT / kWave;
constexpr int kHC = 16;                 // max first-axis histories per EFE item
constexpr int kFpiScr = 2048;           // FP64 LL + stage scratch kept in LDS
constexpr int kTileFloats = 4096;       // A tile of an EFE item (floats)

struct Agent {
  int F, M, T, P, G, iters, totS, has_E;
  float gamma;
  int scr_lds;  // LL and stage buffers in LDS (1) or the workgroup's workspace slot (0)
  int ll_total, xa_off, xb_off, scr_total;
  int S[kF], S_off[kF + 1];
  int O[kM], mod_group[kM];
  int gR[kM], gfac[kM][kR], gdim[kM][kR], gll[kM];
  int BR[kF], Bdep[kF][kRB], BK[kF], U[kF];
  int C_time[kM];
  int H[kT][kF], hist_lvl[kT][kF], roll_lvl[kT][kF];
  int score_base[kT][kM], score_stride[kT][kM][kR];
  long long K[kM];
  long long A_off[kM], B_off[kF], C_off[kM], D_off[kF], E_off;
  long long obs_off, qs_off, G_off, qpi_off, act_off;
  long long hist_off, roll_off, p2h_off, pol0_off, score_off, logit_off;
};

struct Item {
  int agent, m, t, c0, c1;
};

struct Bufs {
  const Agent* ag;
  const int* fpi_order;
  const Item* items;
  const int* iarena;
  const float* params;
  const int* obs;
  float* out;
  int* act;
  double* hist;
  double* score;
  double* ws1;
  double* ws2;
  long long ws1_slot, ws2_slot;
  int* cnt;
  long long* trace;  // optional: [agent or item][start, end] wall clock ticks
  int n_agents, n_items;
};

__device__ __forceinline__ float xlogx_f(float x) { return x == 0.0f ? 0.0f : x * logf(fmaxf(x, kMinVal)); }

// ---------------------------------------------------------------------------
// FPI + rollout, one agent per workgroup iteration
// ---------------------------------------------------------------------------
__device__ void rg_softmax(const Agent& a, const double* src, double* dst, int lane, int wave) {
  for (int f = wave; f < a.F; f += kNW) {
    const int off = a.S_off[f], n = a.S[f];
    const double x = lane < n ? src[off + lane] : -INFINITY;
    const double mx = wave_max(x);
    const double e = lane < n ? exp(x - mx) : 0.0;
    const double s = wave_sum(e);
    if (lane < n) dst[off + lane] = e / s;
  }
}

// One stage of the staged leave-one-out contraction over rows [0, rows) of
// X [rows x n] (lane-per-column mapping): acc += w[r] X[r, col] (weights =
// Kronecker product of the leading factors' q), dst[r] = X[r, :] . q_c.
// U rows per lane are loaded before any is used.
template <int U>
__device__ __forceinline__ void rg_stage_rows(const double* src, int rows, int n, int c, int lg, int per, int rbase,
                                              int col, const double* q0p, const double* q1p, int d1, double inv1,
                                              double qcol, double* dst, double& acc) {
  for (int r0 = 0; r0 < rows; r0 += U * per) {
    double x[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int r = r0 + u * per + rbase;
      x[u] = (r < rows && col < n) ? src[(long long)r * n + col] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int r = r0 + u * per + rbase;
      const bool rv = r < rows;
      double w = 0.0;
      if (rv) {
        if (c == 0) {
          w = 1.0;
        } else if (c == 1) {
          w = q0p[r];
        } else {
          const int hi = (int)(r * inv1 + 1e-9);  // r / d1 (exact: r < 2^22)
          w = q0p[hi] * q1p[r - hi * d1];
        }
      }
      acc = fma(w, x[u], acc);
      const double v = group_sum_low(x[u] * qcol, lg);
      if (dst && rv && col == 0) dst[r] = v;
    }
  }
}

__global__ __launch_bounds__(kNT) void rg_fpi_kernel(Bufs b) {
  __shared__ double s_logq[kF * kDim], s_qv[kF * kDim], s_lp[kF * kDim];
  __shared__ double s_red[2][kNW * kWave];
  __shared__ double s_scr[kFpiScr];
  __shared__ double s_ltab[256];
  __shared__ int s_obs[kM];
  __shared__ int s_item;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid >> 6;
  for (int i = tid; i < 256; i += kNT) s_ltab[i] = g_logtab[i];
  double* wsl = b.ws1 + (long long)blockIdx.x * b.ws1_slot;

  for (;;) {
    if (tid == 0) s_item = atomicAdd(&b.cnt[0], 1);
    __syncthreads();
    const int idx = s_item;
    if (idx >= b.n_agents) break;
    const Agent& a = b.ag[b.fpi_order[idx]];
    double* scr = a.scr_lds ? s_scr : wsl;
    if (b.trace && tid == 0) b.trace[2 * b.fpi_order[idx]] = wall_clock64();

    for (int f = 0; f < a.F; ++f)
      for (int i = tid; i < a.S[f]; i += kNT) {
        s_lp[a.S_off[f] + i] = log_stable_f64(b.params[a.D_off[f] + i]);
        s_logq[a.S_off[f] + i] = 0.0;
      }
    if (tid < a.M) s_obs[tid] = b.obs[a.obs_off + tid];
    __syncthreads();

    // log-likelihood per dependency group: sum over its modalities of
    // ln max(A_m[o_m, s], MINVAL), formed as one FP64 log of the product
    // (an out-of-range index is an all-zero one-hot row: no contribution).
    for (int g = 0; g < a.G; ++g) {
      long long Kg = 1;
      for (int j = 0; j < a.gR[g]; ++j) Kg *= a.gdim[g][j];
      double* ll = scr + a.gll[g];
      for (long long e = tid; e < Kg; e += kNT) {
        double prod = 1.0;
        for (int m = 0; m < a.M; ++m) {
          if (a.mod_group[m] != g) continue;
          const int o = s_obs[m];
          if (o < 0 || o >= a.O[m]) continue;
          prod *= (double)fmaxf(b.params[a.A_off[m] + (long long)o * a.K[m] + e], kMinVal);
        }
        ll[e] = log_f64d(prod, s_ltab);
      }
    }
    __threadfence_block();
    __syncthreads();

    for (int it = 0; it < a.iters; ++it) {
      rg_softmax(a, s_logq, s_qv, lane, wave);
      __syncthreads();
      for (int i = tid; i < a.totS; i += kNT) s_logq[i] = s_lp[i];
      __syncthreads();
      for (int g = 0; g < a.G; ++g) {
        const int R = a.gR[g];
        const double* src = scr + a.gll[g];
        for (int k = 0; k < R; ++k) {
          const int c = R - 1 - k;
          const int n = a.gdim[g][c];
          const int f = a.gfac[g][c];
          const int lg = ceil_log2(n);
          const int rpw_lg = 6 - lg;
          int rows = 1;
          for (int j = 0; j < c; ++j) rows *= a.gdim[g][j];
          double* dst = c > 0 ? scr + ((k & 1) ? a.xb_off : a.xa_off) : nullptr;
          double* red = s_red[k & 1];
          const int col = lane & ((1 << lg) - 1);
          const int lr = lane >> lg;
          const double qcol = col < n ? s_qv[a.S_off[f] + col] : 0.0;
          double acc = 0.0;
          const int per = kNW << rpw_lg;
          const int rbase = (wave << rpw_lg) + lr;
          const double* q0p = c >= 1 ? s_qv + a.S_off[a.gfac[g][0]] : nullptr;
          const double* q1p = c == 2 ? s_qv + a.S_off[a.gfac[g][1]] : nullptr;
          const int d1 = c == 2 ? a.gdim[g][1] : 1;
          const double inv1 = 1.0 / d1;
          // streamed stage 0 of a large agent: 16 row loads in flight per lane
          if (!a.scr_lds && k == 0 && rows >= 16 * per)
            rg_stage_rows<16>(src, rows, n, c, lg, per, rbase, col, q0p, q1p, d1, inv1, qcol, dst, acc);
          else
            rg_stage_rows<4>(src, rows, n, c, lg, per, rbase, col, q0p, q1p, d1, inv1, qcol, dst, acc);
          acc = group_sum_high(acc, lg);
          if ((lane >> lg) == 0 && col < n) red[wave * kWave + col] = acc;
          __threadfence_block();
          __syncthreads();
          if (tid < n) {
            double s = 0.0;
#pragma unroll
            for (int w = 0; w < kNW; ++w) s += red[w * kWave + tid];
            s_logq[a.S_off[f] + tid] += s;
          }
          src = dst;
        }
        __syncthreads();
      }
    }
    rg_softmax(a, s_logq, s_qv, lane, wave);
    __syncthreads();

    // posterior out (FP32) and as the level -1 history vector (FP64)
    double* hist = b.hist + a.hist_off;
    for (int i = tid; i < a.totS; i += kNT) {
      b.out[a.qs_off + i] = (float)s_qv[i];
      hist[i] = s_qv[i];
    }
    __syncthreads();

    // rollout: q_f[t][h] = B_f[:, k, u_h] . (x_j q_{dep_j}[t-1][parent_j(h)])[k]
    const int* roll = b.iarena + a.roll_off;
    for (int t = 0; t < a.T; ++t) {
      for (int f = 0; f < a.F; ++f) {
        const int Sf = a.S[f], rb = a.BR[f], KB = a.BK[f], Uf = a.U[f];
        const int nh = a.H[t][f];
        const float* Bf = b.params + a.B_off[f];
        for (int e = tid; e < nh * Sf; e += kNT) {
          const int h = e / Sf, s = e - h * Sf;
          const int* rec = roll + a.roll_lvl[t][f] + h * (1 + rb);
          const int u = rec[0];
          const double* qp[kRB];
          int dd[kRB];
          for (int j = 0; j < rb; ++j) {
            const int d = a.Bdep[f][j];
            dd[j] = a.S[d];
            qp[j] = hist + (t == 0 ? a.S_off[d] : a.hist_lvl[t - 1][d] + rec[1 + j] * a.S[d]);
          }
          double acc = 0.0;
          for (int k = 0; k < KB; ++k) {
            double w;
            if (rb == 1) {
              w = qp[0][k];
            } else if (rb == 2) {
              const int k0 = k / dd[1];
              w = qp[0][k0] * qp[1][k - k0 * dd[1]];
            } else {
              const int k01 = k / dd[2], k2 = k - k01 * dd[2];
              const int k0 = k01 / dd[1];
              w = qp[0][k0] * qp[1][k01 - k0 * dd[1]] * qp[2][k2];
            }
            acc = fma((double)Bf[((long long)s * KB + k) * Uf + u], w, acc);
          }
          hist[a.hist_lvl[t][f] + h * Sf + s] = acc;
        }
      }
      __syncthreads();
    }
    if (b.trace && tid == 0) b.trace[2 * b.fpi_order[idx] + 1] = wall_clock64();
  }
}

// ---------------------------------------------------------------------------
// EFE items
// ---------------------------------------------------------------------------
// Every contraction stage is a batched small GEMM Y[b*H + h][row] =
// sum_i X[b][row][i] q_h[i] over a row-major X whose last axis is the one
// being contracted; X tiles are staged through LDS with coalesced loads and
// each thread computes four histories of one row from LDS.
constexpr int kTileBytes = (kTileFloats + 2 * kDim) * 4;
constexpr int kHaPart = 512;

// X (FP64, global) [rows x n]; row R belongs to slab b = R / slab, and its
// outputs for histories [h0, h0 + hn) go to Y[(b * H + h) * slab + R % slab].
__device__ void rg_stage_tile(const double* X, long long rows, long long slab, int n, const double* Q, int h0,
                              int hn, int H, double* Y, double* tile, double* qh, int tid) {
  constexpr int kTD = kTileBytes / 8;
  constexpr int kPF = (kTD + kNT - 1) / kNT;
  const int np1 = n + 1;
  int rt = kTD / np1;
  if (rt > rows) rt = (int)rows;
  for (int i = tid; i < hn * n; i += kNT) qh[i] = Q[(long long)h0 * n + i];
  const int hb_n = (hn + 3) / 4;
  // the next tile's values are loaded into registers while this one is used
  double pf[kPF];
  auto fetch = [&](long long r0, int nr) {
#pragma unroll
    for (int j = 0; j < kPF; ++j) {
      const int e = tid + j * kNT;
      pf[j] = e < nr * n ? X[r0 * n + e] : 0.0;
    }
  };
  fetch(0, (int)min((long long)rt, rows));
  for (long long r0 = 0; r0 < rows; r0 += rt) {
    const int nr = (int)min((long long)rt, rows - r0);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kPF; ++j) {
      const int e = tid + j * kNT;
      if (e < nr * n) {
        const int r = e / n, i = e - r * n;
        tile[r * np1 + i] = pf[j];
      }
    }
    __syncthreads();
    if (r0 + rt < rows) fetch(r0 + rt, (int)min((long long)rt, rows - r0 - rt));
    for (int e = tid; e < hb_n * nr; e += kNT) {
      const int hb = e / nr, r = e - hb * nr;
      const int hh0 = hb * 4;
      double acc[4] = {0.0, 0.0, 0.0, 0.0};
      for (int i = 0; i < n; ++i) {
        const double x = tile[r * np1 + i];
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (hh0 + u < hn) acc[u] = fma(x, qh[(hh0 + u) * n + i], acc[u]);
      }
      const long long R = r0 + r;
      const long long bb = R / slab;
      double* y = Y + (bb * H + h0 + hh0) * slab + (R - bb * slab);
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (hh0 + u < hn) y[(long long)u * slab] = acc[u];
    }
  }
  __syncthreads();
}

__global__ __launch_bounds__(kNT) void rg_efe_kernel(Bufs b) {
  __shared__ __align__(16) unsigned char s_tile_raw[kTileBytes];
  __shared__ double s_hap[kHaPart];
  __shared__ double s_ha[kHaPart];
  __shared__ double s_qh[kHC * kDim];
  __shared__ int s_item;
  float* s_tile = reinterpret_cast<float*>(s_tile_raw);
  double* s_tiled = reinterpret_cast<double*>(s_tile_raw);
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid >> 6;
  double* Z = b.ws2 + (long long)blockIdx.x * b.ws2_slot;

  for (;;) {
    if (tid == 0) s_item = atomicAdd(&b.cnt[1], 1);
    __syncthreads();
    const int idx = s_item;
    if (idx >= b.n_items) break;
    const Item itm = b.items[idx];
    const Agent& a = b.ag[itm.agent];
    const int m = itm.m, t = itm.t;
    long long* tr = b.trace ? b.trace + 2 * ((long long)b.n_agents + idx) : nullptr;
    if (tr && tid == 0) tr[0] = wall_clock64();
    const int g = a.mod_group[m];
    const int R = a.gR[g];
    const int O = a.O[m], O1 = O + 1;
    const long long K = a.K[m];
    const int n = a.gdim[g][R - 1];
    const int pre = (int)(K / n);
    const int hc = itm.c1 - itm.c0;
    const double* hist = b.hist + a.hist_off;
    const float* Am = b.params + a.A_off[m];

    {
      const double* q = hist + a.hist_lvl[t][a.gfac[g][R - 1]] + (long long)itm.c0 * n;
      for (int i = tid; i < hc * n; i += kNT) s_qh[i] = q[i];
    }
    // stage 1: Z1[h][o'][p] = sum_i X[o'][p][i] q_h[i], X = [A ; sum_o A ln A]
    const int np1 = n + 1;
    int pt = kTileFloats / (O * np1);
    if (pt > kHaPart / np1) pt = kHaPart / np1;
    if (pt < 1) pt = 1;
    if (pt > pre) pt = pre;
    const int hb_n = (hc + 3) / 4;
    constexpr int kPF1 = (kTileFloats + 2 * kDim + kNT - 1) / kNT;
    float pf[kPF1];
    auto fetch1 = [&](int p0, int np) {
      const int ncol = np * n;
#pragma unroll
      for (int j = 0; j < kPF1; ++j) {
        const int e = tid + j * kNT;
        if (e < O * ncol) {
          const int o = e / ncol;
          const int rem = e - o * ncol;
          pf[j] = __builtin_nontemporal_load(Am + (long long)o * K + (long long)p0 * n + rem);
        }
      }
    };
    fetch1(0, min(pt, pre));
    for (int p0 = 0; p0 < pre; p0 += pt) {
      const int np = min(pt, pre - p0);
      const int ncol = np * n;
      __syncthreads();
#pragma unroll
      for (int j = 0; j < kPF1; ++j) {
        const int e = tid + j * kNT;
        if (e < O * ncol) {
          const int o = e / ncol;
          const int rem = e - o * ncol;
          const int p = rem / n, i = rem - p * n;
          s_tile[(o * np + p) * np1 + i] = pf[j];
        }
      }
      __syncthreads();
      if (p0 + pt < pre) fetch1(p0 + pt, min(pt, pre - p0 - pt));
      // sum_o x ln x per column, the outcome axis split over oc chunks
      {
        int oc = kNT / ncol;
        if (oc < 1) oc = 1;
        if (oc > O) oc = O;
        const int ochunk = (O + oc - 1) / oc;
        for (int e = tid; e < oc * ncol; e += kNT) {
          const int ch = e / ncol, col = e - ch * ncol;
          const int p = col / n, i = col - p * n;
          double sacc = 0.0;
          const int o1 = min(O, (ch + 1) * ochunk);
          for (int o = ch * ochunk; o < o1; ++o) sacc += (double)xlogx_f(s_tile[(o * np + p) * np1 + i]);
          s_hap[e] = sacc;
        }
        __syncthreads();
        for (int col = tid; col < ncol; col += kNT) {
          double sacc = 0.0;
          for (int ch = 0; ch < oc; ++ch) sacc += s_hap[ch * ncol + col];
          const int p = col / n, i = col - p * n;
          s_ha[p * np1 + i] = sacc;
        }
      }
      __syncthreads();
      const int outs = hb_n * O1 * np;
      for (int e = tid; e < outs; e += kNT) {
        const int hb = e / (O1 * np);
        const int rem = e - hb * (O1 * np);
        const int o = rem / np, p = rem - o * np;
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        const int h0 = hb * 4;
        for (int i = 0; i < n; ++i) {
          const double x = o < O ? (double)s_tile[(o * np + p) * np1 + i] : s_ha[p * np1 + i];
#pragma unroll
          for (int hh = 0; hh < 4; ++hh)
            if (h0 + hh < hc) acc[hh] = fma(x, s_qh[(h0 + hh) * n + i], acc[hh]);
        }
#pragma unroll
        for (int hh = 0; hh < 4; ++hh)
          if (h0 + hh < hc) Z[((long long)(h0 + hh) * O1 + o) * pre + p0 + p] = acc[hh];
      }
    }
    __threadfence_block();
    __syncthreads();

    // stages 2..R: contract the next axis with every history of its factor
    double* zin = Z;
    long long Bin = hc, rows_in = pre;  // zin: [Bin][O1 * rows_in], last axis contracted next
    for (int j = R - 2; j >= 0; --j) {
      const int nj = a.gdim[g][j];
      const int fj = a.gfac[g][j];
      const int Hj = a.H[t][fj];
      const double* qj = hist + a.hist_lvl[t][fj];
      const long long rows_out = rows_in / nj;
      const long long slab = O1 * rows_out;  // output values per (b, h)
      double* zout = zin + Bin * O1 * rows_in;
      for (int h0 = 0; h0 < Hj; h0 += kHC)
        rg_stage_tile(zin, Bin * slab, slab, nj, qj, h0, min(kHC, Hj - h0), Hj, zout, s_tiled, s_qh, tid);
      __threadfence_block();
      __syncthreads();
      zin = zout;
      Bin *= Hj;
      rows_in = rows_out;
    }

    // score every combination: H(qo) + sum_s (sum_o A ln A) x q + qo . C
    const float* Cm = b.params + a.C_off[m] + (a.C_time[m] ? (long long)t * O : 0);
    double* sc = b.score + a.score_off + a.score_base[t][m] + (long long)itm.c0 * a.score_stride[t][m][R - 1];
    const float cl = lane < O ? Cm[lane] : 0.0f;
    for (long long c = wave; c < Bin; c += kNW) {
      const double* z = zin + c * O1;
      double v = 0.0;
      if (lane < O) {
        const double q = z[lane];
        const float qf = (float)q;
        v = q * (double)cl;
        if (qf != 0.0f) v -= q * (double)logf(fmaxf(qf, kMinVal));
      }
      v = wave_sum(v);
      if (lane == 0) sc[c] = v + z[O];
    }
    if (tr && tid == 0) tr[1] = wall_clock64();
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Per-agent G, q_pi, action
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kNT) void rg_assemble_kernel(Bufs b) {
  __shared__ double s_red[kNW];
  __shared__ float s_redf[kNW];
  __shared__ int s_redi[kNW];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid >> 6;
  const Agent& a = b.ag[blockIdx.x];
  const int P = a.P, T = a.T, F = a.F;
  const int* p2h = b.iarena + a.p2h_off;
  const double* sc = b.score + a.score_off;
  double* logit = b.score + a.logit_off;
  double lmax = -INFINITY;
  for (int p = tid; p < P; p += kNT) {
    double G = 0.0;
    for (int t = 0; t < T; ++t)
      for (int m = 0; m < a.M; ++m) {
        const int g = a.mod_group[m];
        long long ix = a.score_base[t][m];
        for (int j = 0; j < a.gR[g]; ++j)
          ix += (long long)p2h[((long long)p * T + t) * F + a.gfac[g][j]] * a.score_stride[t][m][j];
        G += sc[ix];
      }
    b.out[a.G_off + p] = (float)G;
    double l = (double)a.gamma * G;
    if (a.has_E) l += log_stable_f64(b.params[a.E_off + p]);
    logit[p] = l;
    lmax = fmax(lmax, l);
  }
  lmax = wave_max(lmax);
  if (lane == 0) s_red[wave] = lmax;
  __syncthreads();
  double mx = s_red[0];
  for (int w = 1; w < kNW; ++w) mx = fmax(mx, s_red[w]);
  __syncthreads();
  double se = 0.0;
  for (int p = tid; p < P; p += kNT) se += exp(logit[p] - mx);
  se = wave_sum(se);
  if (lane == 0) s_red[wave] = se;
  __syncthreads();
  double tot = 0.0;
  for (int w = 0; w < kNW; ++w) tot += s_red[w];
  // q_pi in FP32 and the first index of its maximum (argmax of the returned q_pi)
  float best = -1.0f;
  int bi = 0x7fffffff;
  for (int p = tid; p < P; p += kNT) {
    const float q = (float)(exp(logit[p] - mx) / tot);
    b.out[a.qpi_off + p] = q;
    if (q > best) {
      best = q;
      bi = p;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    const float ob = __shfl_xor(best, o, kWave);
    const int oi = __shfl_xor(bi, o, kWave);
    if (ob > best || (ob == best && oi < bi)) {
      best = ob;
      bi = oi;
    }
  }
  if (lane == 0) {
    s_redf[wave] = best;
    s_redi[wave] = bi;
  }
  __syncthreads();
  if (tid == 0) {
    float bv = s_redf[0];
    int bx = s_redi[0];
    for (int w = 1; w < kNW; ++w)
      if (s_redf[w] > bv || (s_redf[w] == bv && s_redi[w] < bx)) {
        bv = s_redf[w];
        bx = s_redi[w];
      }
    const int* pol0 = b.iarena + a.pol0_off;
    for (int f = 0; f < F; ++f) b.act[a.act_off + f] = pol0[(long long)bx * F + f];
  }
  // the queues of the two persistent kernels start from zero at the next step
  if (blockIdx.x == 0 && tid == 0) {
    b.cnt[0] = 0;
    b.cnt[1] = 0;
  }
}

// ---------------------------------------------------------------------------
// Host plan
// ---------------------------------------------------------------------------
struct Plan {
  int device = 0;
  int64_t n_agents = 0, sumF = 0, sumM = 0;
  std::vector<Agent> agents;
  std::vector<int> fpi_order;
  std::vector<Item> items;
  std::vector<int32_t> iarena;
  std::vector<int64_t> layout;  // see pymdp_hip_ragged_layout
  long long hist_total = 0, score_total = 0, ws1_need = 0, ws2_need = 0;
  int grid1 = 0, grid2 = 0, grid_max = 0;
  bool tuned = false;
  long long* trace = nullptr;  // caller-owned device buffer, see pymdp_hip_ragged_set_trace
  // device
  Agent* d_ag = nullptr;
  int* d_order = nullptr;
  Item* d_items = nullptr;
  int32_t* d_iarena = nullptr;
  double *d_hist = nullptr, *d_score = nullptr, *d_ws1 = nullptr, *d_ws2 = nullptr;
  int* d_cnt = nullptr;
};

static int einval(const std::string& m) {
  set_error("ragged: " + m);
  return PYMDP_HIP_EINVAL;
}

static void free_plan(Plan* p) {
  if (!p) return;
  int prev = 0;
  (void)hipGetDevice(&prev);
  (void)hipSetDevice(p->device);
  (void)hipFree(p->d_ag);
  (void)hipFree(p->d_order);
  (void)hipFree(p->d_items);
  (void)hipFree(p->d_iarena);
  (void)hipFree(p->d_hist);
  (void)hipFree(p->d_score);
  (void)hipFree(p->d_ws1);
  (void)hipFree(p->d_ws2);
  (void)hipFree(p->d_cnt);
  (void)hipSetDevice(prev);
  delete p;
}

// Stage costs of contracting axes last-to-first in storage order `ax`.
static double efe_cost(const std::vector<int>& dims, const std::vector<int>& Hs, int O) {
  double rows = O + 1.0, c = 0.0, mult = 1.0;
  for (int d : dims) rows *= d;
  for (int j = (int)dims.size() - 1; j >= 0; --j) {
    mult *= Hs[j];
    c += mult * rows;
    rows /= dims[j];
  }
  return c + mult * O;
}

}  // namespace rg
}  // namespace pmh

using namespace pmh;
using namespace pmh::rg;

extern "C" int pymdp_hip_ragged_plan_create(int64_t n_agents, const int64_t* hdr, const int64_t* S_flat,
                                             const int64_t* U_flat, const int64_t* O_flat,
                                             const int64_t* A_dep_rank, const int64_t* A_dep_flat,
                                             const int64_t* B_dep_rank, const int64_t* B_dep_flat,
                                             const int64_t* C_time, const int32_t* E_present,
                                             const float* gamma, const int32_t* policy_flat, int32_t num_iter,
                                             void** plan_out) {
  if (!plan_out) return einval("plan_out is NULL");
  *plan_out = nullptr;
  if (n_agents <= 0) return einval("n_agents must be positive");
  if (num_iter < 0) return einval("num_iter must be >= 0");
  Plan* pl = new Plan();
  pl->n_agents = n_agents;
  long long pA = 0, pO = 0, pS = 0;  // running offsets: A_dep_flat, modality, factor
  long long pB = 0, pPol = 0;
  long long params = 0, obs = 0, outf = 0, acts = 0, hist = 0, score = 0;
  std::vector<int64_t> A_off, A_perm, B_off, C_off, D_off, E_off, obs_off, qs_off, G_off, qpi_off, act_off;
  std::vector<double> fpi_cost(n_agents);
  struct ItemC {
    Item it;
    double cost;
  };
  std::vector<ItemC> items;
  double total_cost = 0.0;
  std::vector<std::vector<ItemC>> per_agent_items(n_agents);
  for (int64_t ai = 0; ai < n_agents; ++ai) {
    const int64_t F = hdr[4 * ai], M = hdr[4 * ai + 1], T = hdr[4 * ai + 2], P = hdr[4 * ai + 3];
    auto fail = [&](const std::string& m) {
      delete pl;
      return einval("agent " + std::to_string(ai) + ": " + m);
    };
    if (F < 1 || F > kF) return fail("number of factors must be in [1, 8]");
    if (M < 1 || M > kM) return fail("number of modalities must be in [1, 8]");
    if (T < 1 || T > kT) return fail("policy horizon must be in [1, 4]");
    if (P < 1 || P > (1 << 24)) return fail("number of policies must be in [1, 2^24]");
    Agent a;
    std::memset(&a, 0, sizeof(a));
    a.F = (int)F;
    a.M = (int)M;
    a.T = (int)T;
    a.P = (int)P;
    a.has_E = E_present ? E_present[ai] : 0;
    a.gamma = gamma ? gamma[ai] : 1.0f;
    std::vector<int64_t> S(S_flat + pS, S_flat + pS + F), U(U_flat + pS, U_flat + pS + F), O(O_flat + pO, O_flat + pO + M);
    for (int f = 0; f < F; ++f) {
      if (S[f] < 1 || S[f] > kDim) return fail("state counts must be in [1, 64]");
      if (U[f] < 1) return fail("control counts must be positive");
      a.S[f] = (int)S[f];
      a.U[f] = (int)U[f];
      a.S_off[f + 1] = a.S_off[f] + a.S[f];
    }
    a.totS = a.S_off[F];
    for (int m = 0; m < M; ++m) {
      if (O[m] < 1 || O[m] > kDim) return fail("outcome counts must be in [1, 64]");
      a.O[m] = (int)O[m];
    }
    // dependency lists (as given; sorted into sets below)
    std::vector<int64_t> Adf, Ado{0}, Bdf, Bdo{0};
    for (int m = 0; m < M; ++m) {
      const int64_t r = A_dep_rank[pO + m];
      if (r < 1 || r > kR) return fail("A_dependencies rank must be in [1, 3]");
      for (int j = 0; j < r; ++j) Adf.push_back(A_dep_flat[pA++]);
      Ado.push_back((int64_t)Adf.size());
    }
    for (int f = 0; f < F; ++f) {
      const int64_t r = B_dep_rank[pS + f];
      if (r < 1 || r > kRB) return fail("B_dependencies rank must be in [1, 3]");
      for (int j = 0; j < r; ++j) Bdf.push_back(B_dep_flat[pB++]);
      Bdo.push_back((int64_t)Bdf.size());
    }
    EfeHostPlan hp;
    const std::string err = build_efe_host_plan(policy_flat + pPol, P, T, F, S.data(), U.data(), O.data(), M,
                                                Adf.data(), Ado.data(), Bdf.data(), Bdo.data(), &hp);
    if (!err.empty()) return fail(err);
    for (int f = 0; f < F; ++f) {
      a.BR[f] = (int)(Bdo[f + 1] - Bdo[f]);
      a.BK[f] = 1;
      for (int j = 0; j < a.BR[f]; ++j) {
        a.Bdep[f][j] = (int)Bdf[Bdo[f] + j];
        a.BK[f] *= a.S[a.Bdep[f][j]];
      }
    }
    // dependency sets -> storage order minimising the EFE contraction cost
    std::vector<std::vector<int>> sets;
    std::vector<std::vector<int>> set_order;
    for (int m = 0; m < M; ++m) {
      std::vector<int> d(Adf.begin() + Ado[m], Adf.begin() + Ado[m + 1]);
      for (int x : d)
        if (x < 0 || x >= F) return fail("A_dependencies factor index out of range");
      std::vector<int> sd = d;
      std::sort(sd.begin(), sd.end());
      for (size_t i = 1; i < sd.size(); ++i)
        if (sd[i] == sd[i - 1]) return fail("duplicate factor in A_dependencies");
      int g = -1;
      for (size_t k = 0; k < sets.size(); ++k)
        if (sets[k] == sd) g = (int)k;
      if (g < 0) {
        sets.push_back(sd);
        g = (int)sets.size() - 1;
        std::vector<int> best = sd, perm = sd;
        double bc = 1e300;
        do {
          std::vector<int> dims;
          for (int x : perm) dims.push_back(a.S[x]);
          double c = 0;
          for (int t = 0; t < T; ++t) {
            std::vector<int> Hs;
            for (int x : perm) Hs.push_back(hp.fac[t][x].n);
            c += efe_cost(dims, Hs, 1);
          }
          if (c < bc) {
            bc = c;
            best = perm;
          }
        } while (std::next_permutation(perm.begin(), perm.end()));
        set_order.push_back(best);
      }
      a.mod_group[m] = g;
      // storage permutation of modality m: positions of the stored axes in its given list
      for (int j = 0; j < kR; ++j) {
        int pos = -1;
        if (j < (int)d.size())
          for (size_t i = 0; i < d.size(); ++i)
            if (d[i] == set_order[g][j]) pos = (int)i;
        A_perm.push_back(pos);
      }
    }
    a.G = (int)sets.size();
    int ll = 0;
    int xa = 0, xb = 0;
    for (int g = 0; g < a.G; ++g) {
      a.gR[g] = (int)set_order[g].size();
      int K = 1;
      for (int j = 0; j < a.gR[g]; ++j) {
        a.gfac[g][j] = set_order[g][j];
        a.gdim[g][j] = a.S[set_order[g][j]];
        K *= a.gdim[g][j];
      }
      a.gll[g] = ll;
      ll += K;
      if (a.gR[g] >= 2) xa = std::max(xa, K / a.gdim[g][a.gR[g] - 1]);
      if (a.gR[g] >= 3) xb = std::max(xb, a.gdim[g][0]);
    }
    a.ll_total = ll;
    a.xa_off = ll;
    a.xb_off = ll + xa;
    a.scr_total = ll + xa + xb;
    a.scr_lds = a.scr_total <= kFpiScr ? 1 : 0;
    if (!a.scr_lds) pl->ws1_need = std::max(pl->ws1_need, (long long)a.scr_total);
    a.iters = (F == 1 && num_iter > 0) ? 1 : num_iter;
    // params arena
    for (int m = 0; m < M; ++m) {
      long long K = 1;
      for (int j = 0; j < a.gR[a.mod_group[m]]; ++j) K *= a.gdim[a.mod_group[m]][j];
      a.K[m] = K;
      a.A_off[m] = params;
      A_off.push_back(params);
      params += (long long)a.O[m] * K;
    }
    for (int f = 0; f < F; ++f) {
      a.B_off[f] = params;
      B_off.push_back(params);
      params += (long long)a.S[f] * a.BK[f] * a.U[f];
    }
    for (int m = 0; m < M; ++m) {
      a.C_time[m] = C_time ? (int)C_time[pO + m] : 0;
      a.C_off[m] = params;
      C_off.push_back(params);
      params += (long long)a.O[m] * (a.C_time[m] ? T : 1);
    }
    for (int f = 0; f < F; ++f) {
      a.D_off[f] = params;
      D_off.push_back(params);
      params += a.S[f];
    }
    a.E_off = a.has_E ? params : -1;
    E_off.push_back(a.E_off);
    if (a.has_E) params += P;
    params = (params + 3) & ~3LL;
    a.obs_off = obs;
    obs_off.push_back(obs);
    obs += M;
    a.qs_off = outf;
    qs_off.push_back(outf);
    outf += a.totS;
    a.G_off = outf;
    G_off.push_back(outf);
    outf += P;
    a.qpi_off = outf;
    qpi_off.push_back(outf);
    outf += P;
    a.act_off = acts;
    act_off.push_back(acts);
    acts += F;
    // histories, rollout records, policy tables
    a.hist_off = hist;
    long long hl = a.totS;
    a.roll_off = (long long)pl->iarena.size();
    int rl = 0;
    for (int t = 0; t < T; ++t)
      for (int f = 0; f < F; ++f) {
        const FactorLevelPlan& lv = hp.fac[t][f];
        a.H[t][f] = lv.n;
        a.hist_lvl[t][f] = (int)hl;
        hl += (long long)lv.n * a.S[f];
        a.roll_lvl[t][f] = rl;
        for (int h = 0; h < lv.n; ++h) {
          pl->iarena.push_back(lv.action[h]);
          for (int j = 0; j < lv.rank; ++j) pl->iarena.push_back(lv.parents[(size_t)h * lv.rank + j]);
        }
        rl += lv.n * (1 + lv.rank);
      }
    if (hl > INT32_MAX) return fail("too many history vectors");
    hist += hl;
    a.p2h_off = (long long)pl->iarena.size();
    for (int64_t p = 0; p < P; ++p)
      for (int t = 0; t < T; ++t)
        for (int f = 0; f < F; ++f) pl->iarena.push_back(hp.fac[t][f].p2h[p]);
    a.pol0_off = (long long)pl->iarena.size();
    for (int64_t p = 0; p < P; ++p)
      for (int f = 0; f < F; ++f) pl->iarena.push_back(policy_flat[pPol + (p * T + 0) * F + f]);
    // score layout and EFE items
    a.score_off = score;
    long long sl = 0;
    double acost = 0.0;
    for (int t = 0; t < T; ++t)
      for (int m = 0; m < M; ++m) {
        const int g = a.mod_group[m], R = a.gR[g];
        long long stride = 1;
        for (int j = 0; j < R; ++j) {
          a.score_stride[t][m][j] = (int)stride;
          stride *= a.H[t][a.gfac[g][j]];
        }
        if (sl + stride > INT32_MAX) return fail("too many history combinations");
        a.score_base[t][m] = (int)sl;
        sl += stride;
        // items: chunks of the first contracted axis' histories
        std::vector<int> dims, Hs;
        for (int j = 0; j < R; ++j) {
          dims.push_back(a.gdim[g][j]);
          Hs.push_back(a.H[t][a.gfac[g][j]]);
        }
        const double c = efe_cost(dims, Hs, a.O[m]);
        const int HR = Hs[R - 1];
        // workspace (doubles) per first-axis history
        double per_h = (a.O[m] + 1.0) * (double)(a.K[m] / dims[R - 1]);
        {
          double rows = (double)(a.K[m] / dims[R - 1]), mult = 1.0, acc = per_h;
          for (int j = R - 2; j >= 0; --j) {
            mult *= Hs[j];
            rows /= dims[j];
            acc += mult * (a.O[m] + 1.0) * rows;
          }
          per_h = acc;
        }
        const int hc_max = std::min(HR, kHC);
        for (int c0 = 0; c0 < HR; c0 += hc_max) {
          const int c1 = std::min(HR, c0 + hc_max);
          ItemC ic{Item{(int)ai, m, t, c0, c1}, c * (c1 - c0) / HR};
          per_agent_items[ai].push_back(ic);
          pl->ws2_need = std::max(pl->ws2_need, (long long)std::ceil(per_h * (c1 - c0)));
        }
        acost += c;
      }
    a.logit_off = score + sl;
    score += sl + P;
    total_cost += acost;
    {
      double fc = 0.0;
      for (int g = 0; g < a.G; ++g) {
        double K = 1;
        for (int j = 0; j < a.gR[g]; ++j) K *= a.gdim[g][j];
        fc += K * (2.0 + 3.0 * a.iters * a.gR[g]);
      }
      fpi_cost[ai] = fc;
    }
    pl->agents.push_back(a);
    pS += F;
    pO += M;
    pPol += P * T * F;
  }
  for (auto& v : per_agent_items)
    for (auto& ic : v) items.push_back(ic);
  std::stable_sort(items.begin(), items.end(), [](const ItemC& x, const ItemC& y) { return x.cost > y.cost; });
  for (auto& ic : items) pl->items.push_back(ic.it);
  pl->fpi_order.resize(n_agents);
  std::iota(pl->fpi_order.begin(), pl->fpi_order.end(), 0);
  std::stable_sort(pl->fpi_order.begin(), pl->fpi_order.end(),
                   [&](int x, int y) { return fpi_cost[x] > fpi_cost[y]; });
  pl->sumF = pS;
  pl->sumM = pO;
  pl->hist_total = hist;
  pl->score_total = score;
  // layout vector (host)
  auto& L = pl->layout;
  L = {params, obs, outf, acts};
  L.insert(L.end(), A_off.begin(), A_off.end());
  L.insert(L.end(), A_perm.begin(), A_perm.end());
  L.insert(L.end(), B_off.begin(), B_off.end());
  L.insert(L.end(), C_off.begin(), C_off.end());
  L.insert(L.end(), D_off.begin(), D_off.end());
  L.insert(L.end(), E_off.begin(), E_off.end());
  L.insert(L.end(), obs_off.begin(), obs_off.end());
  L.insert(L.end(), qs_off.begin(), qs_off.end());
  L.insert(L.end(), G_off.begin(), G_off.end());
  L.insert(L.end(), qpi_off.begin(), qpi_off.end());
  L.insert(L.end(), act_off.begin(), act_off.end());

  // launch grids (workgroups per CU of the 256-CU MI355X; see pymdp_hip_ragged_autotune)
  const int cus = 256;
  pl->grid_max = cus * 8;
  pl->grid1 = (int)std::min<long long>(n_agents, cus * 4);
  pl->grid2 = (int)std::min<long long>((long long)pl->items.size(), cus * 4);
  // keep the EFE workspace within 16 GiB
  const long long ws2_slot = std::max<long long>(pl->ws2_need, 1);
  const long long cap = (16LL << 30) / 8 / ws2_slot;
  if (cap < 1) {
    free_plan(pl);
    return einval("an EFE work item needs more than 16 GiB of workspace");
  }
  pl->grid_max = (int)std::min<long long>(pl->grid_max, std::max<long long>(cap, 1));
  pl->grid1 = std::min(pl->grid1, pl->grid_max);
  pl->grid2 = std::max(1, std::min(pl->grid2, pl->grid_max));
  pl->ws2_need = ws2_slot;
  pl->ws1_need = std::max<long long>(pl->ws1_need, 1);
  *plan_out = pl;
  return PYMDP_HIP_OK;
}

static int upload(Plan* pl) {
  if (pl->d_ag) return 0;
  auto fail_dev = [&](hipError_t e, const char* what) {
    set_error(std::string("ragged: ") + what + ": " + hipGetErrorString(e));
    return PYMDP_HIP_EDEVICE;
  };
  hipError_t e;
  if ((e = hipMalloc(&pl->d_ag, sizeof(Agent) * pl->agents.size())) != hipSuccess) return fail_dev(e, "hipMalloc");
  if ((e = hipMalloc(&pl->d_order, sizeof(int) * pl->fpi_order.size())) != hipSuccess) return fail_dev(e, "hipMalloc");
  if ((e = hipMalloc(&pl->d_items, sizeof(Item) * std::max<size_t>(1, pl->items.size()))) != hipSuccess)
    return fail_dev(e, "hipMalloc");
  if ((e = hipMalloc(&pl->d_iarena, sizeof(int32_t) * std::max<size_t>(1, pl->iarena.size()))) != hipSuccess)
    return fail_dev(e, "hipMalloc");
  if ((e = hipMalloc(&pl->d_hist, sizeof(double) * std::max<long long>(1, pl->hist_total))) != hipSuccess)
    return fail_dev(e, "hipMalloc");
  if ((e = hipMalloc(&pl->d_score, sizeof(double) * std::max<long long>(1, pl->score_total))) != hipSuccess)
    return fail_dev(e, "hipMalloc");
  if ((e = hipMalloc(&pl->d_ws1, sizeof(double) * (size_t)pl->ws1_need * pl->grid_max)) != hipSuccess)
    return fail_dev(e, "hipMalloc");
  if ((e = hipMalloc(&pl->d_ws2, sizeof(double) * (size_t)pl->ws2_need * pl->grid_max)) != hipSuccess)
    return fail_dev(e, "hipMalloc");
  if ((e = hipMalloc(&pl->d_cnt, sizeof(int) * 4)) != hipSuccess) return fail_dev(e, "hipMalloc");
  if ((e = hipMemcpy(pl->d_ag, pl->agents.data(), sizeof(Agent) * pl->agents.size(), hipMemcpyHostToDevice)) != hipSuccess)
    return fail_dev(e, "hipMemcpy");
  if ((e = hipMemcpy(pl->d_order, pl->fpi_order.data(), sizeof(int) * pl->fpi_order.size(), hipMemcpyHostToDevice)) !=
      hipSuccess)
    return fail_dev(e, "hipMemcpy");
  if (!pl->items.empty() &&
      (e = hipMemcpy(pl->d_items, pl->items.data(), sizeof(Item) * pl->items.size(), hipMemcpyHostToDevice)) !=
          hipSuccess)
    return fail_dev(e, "hipMemcpy");
  if (!pl->iarena.empty() && (e = hipMemcpy(pl->d_iarena, pl->iarena.data(), sizeof(int32_t) * pl->iarena.size(),
                                            hipMemcpyHostToDevice)) != hipSuccess)
    return fail_dev(e, "hipMemcpy");
  if ((e = hipMemset(pl->d_cnt, 0, sizeof(int) * 4)) != hipSuccess) return fail_dev(e, "hipMemset");
  if ((e = hipDeviceSynchronize()) != hipSuccess) return fail_dev(e, "hipDeviceSynchronize");
  return PYMDP_HIP_OK;
}

extern "C" int pymdp_hip_ragged_plan_upload(void* plan) {
  Plan* pl = static_cast<Plan*>(plan);
  if (!pl) return einval("plan is NULL");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    set_error("ragged: no HIP device");
    return PYMDP_HIP_EDEVICE;
  }
  pl->device = dev;
  return upload(pl);
}

extern "C" int pymdp_hip_ragged_tables(void* plan, int64_t* sizes8, void* agents_out, int32_t* order_out,
                                       int32_t* items_out, int32_t* iarena_out) {
  Plan* pl = static_cast<Plan*>(plan);
  if (!pl) return einval("plan is NULL");
  if (sizes8) {
    sizes8[0] = (int64_t)sizeof(Agent);
    sizes8[1] = (int64_t)sizeof(Item);
    sizes8[2] = (int64_t)pl->items.size();
    sizes8[3] = (int64_t)pl->iarena.size();
    sizes8[4] = pl->grid_max;
    sizes8[5] = 0;
    sizes8[6] = 0;
    sizes8[7] = 0;
  }
  if (agents_out) std::memcpy(agents_out, pl->agents.data(), sizeof(Agent) * pl->agents.size());
  if (order_out) std::copy(pl->fpi_order.begin(), pl->fpi_order.end(), order_out);
  if (items_out) std::memcpy(items_out, pl->items.data(), sizeof(Item) * pl->items.size());
  if (iarena_out) std::copy(pl->iarena.begin(), pl->iarena.end(), iarena_out);
  return PYMDP_HIP_OK;
}

extern "C" int pymdp_hip_ragged_layout(void* plan, int64_t* out, int64_t n_out) {
  Plan* pl = static_cast<Plan*>(plan);
  if (!pl) return einval("plan is NULL");
  if (!out) return (int)std::min<size_t>(INT32_MAX, pl->layout.size());
  if (n_out < (int64_t)pl->layout.size())
    return einval("layout buffer holds " + std::to_string(n_out) + " values, " + std::to_string(pl->layout.size()) +
                  " needed");
  std::copy(pl->layout.begin(), pl->layout.end(), out);
  return PYMDP_HIP_OK;
}

extern "C" int pymdp_hip_ragged_info(void* plan, int64_t* out8) {
  Plan* pl = static_cast<Plan*>(plan);
  if (!pl || !out8) return einval("plan or out is NULL");
  out8[0] = pl->n_agents;
  out8[1] = (int64_t)pl->items.size();
  out8[2] = pl->grid1;
  out8[3] = pl->grid2;
  out8[4] = pl->hist_total;
  out8[5] = pl->score_total;
  out8[6] = pl->ws1_need;
  out8[7] = pl->ws2_need;
  return PYMDP_HIP_OK;
}

static int ragged_launch(Plan* pl, const float* params, const int32_t* obs, float* out, int32_t* act, hipStream_t s,
                         int grid1, int grid2) {
  Bufs b;
  b.ag = pl->d_ag;
  b.fpi_order = pl->d_order;
  b.items = pl->d_items;
  b.iarena = pl->d_iarena;
  b.params = params;
  b.obs = obs;
  b.out = out;
  b.act = act;
  b.hist = pl->d_hist;
  b.score = pl->d_score;
  b.ws1 = pl->d_ws1;
  b.ws2 = pl->d_ws2;
  b.ws1_slot = pl->ws1_need;
  b.ws2_slot = pl->ws2_need;
  b.cnt = pl->d_cnt;
  b.n_agents = (int)pl->n_agents;
  b.n_items = (int)pl->items.size();
  b.trace = pl->trace;
  hipLaunchKernelGGL(rg_fpi_kernel, dim3(grid1), dim3(kNT), 0, s, b);
  hipLaunchKernelGGL(rg_efe_kernel, dim3(grid2), dim3(kNT), 0, s, b);
  hipLaunchKernelGGL(rg_assemble_kernel, dim3((unsigned)pl->n_agents), dim3(kNT), 0, s, b);
  PMH_HIP_CHECK(hipGetLastError(), "ragged: ");
  return 0;
}

extern "C" int pymdp_hip_ragged_step(void* plan, const float* params, const int32_t* obs, float* out, int32_t* act,
                                     void* stream) {
  Plan* pl = static_cast<Plan*>(plan);
  if (!pl) return einval("plan is NULL");
  if (!params || !obs || !out || !act) return einval("arena pointer is NULL");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev != pl->device) return einval("plan was created on another device");
  if (ensure_log_table()) {
    set_error("ragged: cannot initialise the log table");
    return PYMDP_HIP_EDEVICE;
  }
  return ragged_launch(pl, params, obs, out, act, static_cast<hipStream_t>(stream), pl->grid1, pl->grid2);
}

// Times every (grid1, grid2) pair in {2, 4, 8} workgroups per CU (3 steps
// each after one warm-up, hipEvents on `stream`) and keeps the fastest for
// later steps.  Must not be called while `stream` is being captured.
extern "C" int pymdp_hip_ragged_autotune(void* plan, const float* params, const int32_t* obs, float* out, int32_t* act,
                                         void* stream, float* best_ms_out) {
  Plan* pl = static_cast<Plan*>(plan);
  if (!pl) return einval("plan is NULL");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)
    return einval("autotune cannot run while the stream is being captured");
  if (ensure_log_table()) {
    set_error("ragged: cannot initialise the log table");
    return PYMDP_HIP_EDEVICE;
  }
  int cus = pl->grid_max / 8;
  if (cus < 1) cus = 1;
  hipEvent_t e0, e1;
  PMH_HIP_CHECK(hipEventCreate(&e0), "ragged: ");
  PMH_HIP_CHECK(hipEventCreate(&e1), "ragged: ");
  float best = 1e30f;
  int b1 = pl->grid1, b2 = pl->grid2;
  for (int w1 : {2, 4, 8})
    for (int w2 : {2, 4, 8}) {
      const int g1 = (int)std::max<long long>(1, std::min<long long>({(long long)cus * w1, pl->n_agents, (long long)pl->grid_max}));
      const int g2 = (int)std::max<long long>(
          1, std::min<long long>({(long long)cus * w2, (long long)pl->items.size(), (long long)pl->grid_max}));
      int rc = ragged_launch(pl, params, obs, out, act, s, g1, g2);
      if (rc) return rc;
      PMH_HIP_CHECK(hipEventRecord(e0, s), "ragged: ");
      for (int r = 0; r < 3; ++r)
        if ((rc = ragged_launch(pl, params, obs, out, act, s, g1, g2))) return rc;
      PMH_HIP_CHECK(hipEventRecord(e1, s), "ragged: ");
      PMH_HIP_CHECK(hipEventSynchronize(e1), "ragged: ");
      float ms = 0.f;
      PMH_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1), "ragged: ");
      if (ms < best) {
        best = ms;
        b1 = g1;
        b2 = g2;
      }
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  pl->grid1 = b1;
  pl->grid2 = b2;
  pl->tuned = true;
  if (best_ms_out) *best_ms_out = best / 3.0f;
  return PYMDP_HIP_OK;
}

extern "C" int pymdp_hip_ragged_set_trace(void* plan, int64_t* trace) {
  Plan* pl = static_cast<Plan*>(plan);
  if (!pl) return einval("plan is NULL");
  pl->trace = reinterpret_cast<long long*>(trace);
  return PYMDP_HIP_OK;
}

extern "C" int pymdp_hip_ragged_plan_destroy(void* plan) {
  free_plan(static_cast<Plan*>(plan));
  return PYMDP_HIP_OK;
}
