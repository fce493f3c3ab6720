// Batched ragged agents: one step of infer_states -> infer_policies ->
// sample_action for thousands of agents whose factor, modality, control and
// policy counts all differ, as stream-ordered launches that a caller can
// capture in one hipGraph.
//
//   rg_ha_kernel        sum_o A ln A of every modality (once under static
//                       parameters).
//   rg_fpi_kernel       persistent; a workgroup takes the next FPI unit from
//                       an atomic queue.  A unit holds (a row range of) an
//                       agent's log-likelihood tensors in LDS for all sweeps
//                       of the fused log-likelihood + Jacobi FPI of
//                       pymdp/algos.py:88-155.  An agent too large for one
//                       unit is split over a team of units whose workgroups
//                       exchange partial marginals once per sweep.
//   rg_roll_kernel      one launch per policy level: the action-history
//                       rollout of pymdp/control.py:351-389 (B . q per
//                       distinct history).
//   rg_rows_kernel<1>   stage 1, one row per thread: rows (o, i, j) of
//                       [A ; sum_o A ln A] of one modality contracted over
//                       its last storage axis with the history vectors of
//                       every level at once (A is read once per step).
//   rg_rows_kernel<2>   stage 2, the same row streaming over the stage-1
//                       output: the middle axis contracted with the middle
//                       factor's histories of one level.
//   rg_rows_kernel<3>   stage 3, the first axis contracted with the first
//                       factor's histories and every history combination scored:
//                       entropy of qo + ambiguity + qo . C
//                       (pymdp/control.py:420-509).
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
constexpr int kCap = 6144;        // log-likelihood doubles one FPI unit keeps in LDS
constexpr int kRowsCap = 1024;    // rows of one FPI segment
constexpr int kTeamMax = 128;     // FPI units of one agent
constexpr long long kSpinMax = 1LL << 22;
// Team exchange tags: the low kTagBits mantissa bits of every published
// partial carry the sweep's sequence number (step * iters + sweep), cleared
// again on read (a relative change below 2^-42); a fresh partial buffer holds
// kTagFresh, which no sweep of the first step carries (iters <= kTagMaxIters).
constexpr int kTagBits = 10;
constexpr unsigned long long kTagMask = (1ull << kTagBits) - 1;
constexpr unsigned long long kTagFresh = kTagMask;
constexpr int kTagMaxIters = (int)kTagMask - 1;

struct Agent {
  int F, M, T, P, G, iters, totS, has_E, team;
  float gamma;
  int S[kF], S_off[kF + 1];
  int O[kM], mod_group[kM];
  int gR[kM], gfac[kM][kR], gdim[kM][kR];
  int BR[kF], Bdep[kF][kRB], BK[kF], U[kF];
  int C_time[kM];
  int H[kT][kF], hist_lvl[kT][kF];
  int score_base[kT][kM], score_stride[kT][kM][kR];
  // EFE view of modality m: padded dims (ed0, ed1, ed2), ed2 contracted first;
  // ef = factor of each padded dim (-1: padding, one state, one history, q = 1)
  int ed[kM][3], ef[kM][3];
  int tstart[kM][kT + 1];  // first concatenated history column of each level
  long long K[kM];
  long long A_off[kM], B_off[kF], C_off[kM], D_off[kF], E_off;
  long long obs_off, qs_off, G_off, qpi_off, act_off;
  long long hist_off, p2h_off, pol0_off, score_off, logit_off, part_off;
  long long z1_off[kM], z2_off[kM][kT];
  long long ha_off[kM];  // sum_o A ln A table of modality m ([pos][d2] doubles)
  long long pred_off;  // predicted prior B . q for the chosen action, [totS] in `out`
};

// One FPI unit: up to kM segments, each a row range [r0, r1) of the stage-0
// matrix [rows x n] (n = last storage dim) of dependency group g, held in
// LDS at offset lds with row stride n | 1.  x32 = 0: the log-likelihood is
// held as doubles (kCap of them); x32 = 1: as floats (2 kCap), for agents
// whose log-likelihood does not fit one unit as doubles (offsets then count
// floats).
struct Unit {
  int agent, slice, nseg, x32;
  int g[kM], r0[kM], r1[kM], lds[kM];
};

// One EFE tile: rows [r0, r1) and columns [c0, c1) of one stage of modality
// m of `agent` at level t (E1 tiles span every level).
struct Tile {
  int agent, m, t, r0, r1, c0, c1, pad;
};

// Stage-1 / stage-2 work item: rows [e0, e1) of one row space (see
// rg_rows); every address the item needs comes from this record (history
// column n of level t, ts[t] <= n < ts[t + 1], is the vector at
// hb[t] + (n - ts[t]) * d2 of the history buffer).  Stage 1: a_off / h_off
// locate A_m and its sum_o A ln A table, blk = O1 d01 (one output block);
// stage 2: a_off is the z1 offset of the level's first row (rows of d2
// floats, contiguous), blk = O1 d0 rows per output block of N1 columns.
struct E1Rec {
  long long a_off;   // params offset of A_m
  long long h_off;   // ha offset of the modality's sum_o A ln A table ([pos][d2])
  long long z_off;   // z1 offset of the modality
  long long hb[kT];  // hist offset of level t's first last-axis history vector
  int ts[kT + 1];    // first concatenated column of each level (ts[T..kT] = N1)
  int O, d2, d01, e0, e1, N1, blk;
};
static_assert(sizeof(E1Rec) == 104, "stage-1 record layout");

// One block of the sum_o A ln A table (FP32): elements [e0, e0 + kNT) of
// modality m's [pos][k] table (n elements).
struct HaTask {
  long long a_off, ha_off;
  int O, d2, n, e0;
};

// One block of the rollout at level t of factor f of one agent: entries
// [0, nh) of a list of that level's histories sorted by action, each entry
// (h, u, parent_0 .. parent_{rb-1}) in the int arena, every address resolved
// on the host.  q_f[t][h] = B_f[u] . (x_j q_{dep_j}[t-1][parent_j]) with B_f
// stored [U_f][K_B][S_f] (state index fastest) in the params arena.
struct RollTask {
  long long b_off;          // params offset of B_f
  long long out;            // hist index of the level's first vector (h = 0)
  long long par[kRB];       // hist index of each dependency's parent level (t = 0: the posterior)
  long long ent;            // iarena index of the task's first (h, u, parents) entry
  int Sf, KB, rb, nh, t;
  int dd[kRB];              // state count of each dependency
};

struct Bufs {
  const Agent* ag;
  const Unit* units;
  const Tile* t1;
  const E1Rec* e1;
  const E1Rec* e2;
  const E1Rec* e3;
  const HaTask* hat;
  const RollTask* roll;
  const Tile* t2;
  const Tile* t3;
  const int* iarena;
  const float* params;
  const int* obs;
  float* out;
  int* act;
  double* hist;
  double* score;
  float* ha;  // sum_o A ln A rows of every modality, [pos][d2] (FP32)
  float* z1;
  float* z2;
  double* part;
  int* cnt;          // [0] FPI queue, [1] step sequence number, [3] team-exchange timeout flag
  long long* trace;  // optional: [unit or tile][start, end] wall clock ticks
  int n_agents, n_units, n_t1, n_t2, n_t3;
  int carry;  // 1: the FPI prior is the previous step's predicted prior (out + pred_off), 0: D
};

// x ln x of a float with the reference's clip (0 for x == 0) in FP64, the
// log from ln_fast (exact exponent, hardware log2 of the mantissa: an
// absolute error below ~1.5e-7 at any magnitude); the outcome entropies and
// the sum_o A ln A it feeds are summed in FP64.
__device__ __forceinline__ double xlogx_fast(float x) {
  return x == 0.0f ? 0.0 : (double)x * ln_fast((double)fmaxf(x, kMinVal));
}

// *dst = *src for a plain record in global memory, copied into LDS with every
// load issued before the first store.  The kernels read their agent / unit
// records from LDS: a field read is then an LDS access instead of a global
// round trip that the compiler repeats after every global store it cannot
// prove unaliased (and chains when one field indexes another).  The caller
// places a barrier before the copy is read.
template <typename T, int NT = kNT>
__device__ __forceinline__ void rg_copy_lds(const T* src, T* dst, int tid) {
  static_assert(sizeof(T) % 4 == 0, "record of whole dwords");
  constexpr int n = (int)(sizeof(T) / 4), NB = (n + NT - 1) / NT;
  const int* sp = reinterpret_cast<const int*>(src);
  int* dp = reinterpret_cast<int*>(dst);
  int v[NB];
#pragma unroll
  for (int q = 0; q < NB; ++q)
    if (tid + q * NT < n) v[q] = sp[tid + q * NT];
#pragma unroll
  for (int q = 0; q < NB; ++q)
    if (tid + q * NT < n) dp[tid + q * NT] = v[q];
}

// ---------------------------------------------------------------------------
// FPI units
// ---------------------------------------------------------------------------
// dst_f = softmax(x_f) per factor, one wave per factor.  x = src, or with
// part != nullptr x = src + part (the log prior plus this sweep's partial
// marginals), part then being cleared for the next sweep: that folds the
// log q update of a single-unit agent into its softmax (no separate pass and
// barrier).
__device__ void rg_softmax(const Agent& a, const double* src, double* part, double* dst, int lane, int wave) {
  for (int f = wave; f < a.F; f += kNW) {
    const int off = a.S_off[f], n = a.S[f];
    double x = -INFINITY;
    if (lane < n) {
      x = src[off + lane];
      if (part) {
        x += part[off + lane];
        part[off + lane] = 0.0;
      }
    }
    const double mx = wave_max_dpp(x);
    const double e = lane < n ? exp(x - mx) : 0.0;
    const double s = wave_sum_dpp(e);
    if (lane < n) dst[off + lane] = e / s;
  }
}

// log2 of the threads that share one row's dot product (about 8 columns each)
__device__ __forceinline__ int rows_lg(int n) { return n <= 8 ? 0 : ceil_log2(n) - 3; }

// Adds one segment's contribution to the partial marginals s_part of every
// factor of group g (one pass over the rows; pymdp/algos.py:26-51 for all
// factors of the group at once):
//   R = 1: m_0[k]  += X[0, k]
//   R = 2: rows i: m_1[k] += sum_i q_0[i] X[i, k];  m_0[i] += X[i, :] . q_1
//   R = 3: rows r = i*d1 + j: m_2[k] += sum_r q_0[i] q_1[j] X[r, k];
//          y_r = X[r, :] . q_2;  m_1[j] += sum_i q_0[i] y_ij;  m_0[i] += sum_j q_1[j] y_ij
// Two barriers; every sum has a fixed order (bitwise reproducible).  X is
// double or float (the unit's x32); products and sums are FP64 either way.
template <typename TX>
__device__ void rg_fpi_segment(const Agent& a, int g, int r0, int nr, const TX* X, double* s_y,
                               double* s_red, const double* s_qv, double* s_part, int tid) {
  const int R = a.gR[g];
  const int n = a.gdim[g][R - 1], ld = n | 1;
  const int fl = a.gfac[g][R - 1];
  const double* qc = s_qv + a.S_off[fl];
  const double* q0 = R >= 2 ? s_qv + a.S_off[a.gfac[g][0]] : nullptr;
  const double* q1 = R == 3 ? s_qv + a.S_off[a.gfac[g][1]] : nullptr;
  const int d1 = R == 3 ? a.gdim[g][1] : 1;
  // row dot products (R >= 2): 2^lt lanes per row, four accumulators each
  // (eight LDS reads in flight per wait)
  if (R >= 2) {
    // lanes per row: about 8 columns each, fewer when that would take more
    // than one round of the workgroup's threads
    int lt = rows_lg(n);
    while (lt > 0 && (nr << lt) > kNT) --lt;
    const int tpr = 1 << lt;
    for (int e = tid; e < nr * tpr; e += kNT) {
      const int rl = e >> lt, sub = e & (tpr - 1);
      const TX* x = X + rl * ld;
      double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
      int k = sub;
      for (; k + 3 * tpr < n; k += 4 * tpr) {
        v0 = fma((double)x[k], qc[k], v0);
        v1 = fma((double)x[k + tpr], qc[k + tpr], v1);
        v2 = fma((double)x[k + 2 * tpr], qc[k + 2 * tpr], v2);
        v3 = fma((double)x[k + 3 * tpr], qc[k + 3 * tpr], v3);
      }
      for (; k < n; k += tpr) v0 = fma((double)x[k], qc[k], v0);
      double v = group_sum_low_dpp((v0 + v1) + (v2 + v3), lt);
      if (sub == 0) {
        if (R == 2)
          s_part[a.S_off[a.gfac[g][0]] + r0 + rl] += v;
        else
          s_y[rl] = v;
      }
    }
  }
  // weighted column sums: thread (col, rg) over rows rg, rg + nrg, ...;
  // the row's (i, j) advances incrementally (no division per row).  The nrg
  // row groups are summed serially below, so nrg is capped near sqrt(nr):
  // rows per thread + partials per column stays small (a narrow last axis
  // would otherwise give up to 128 groups of one row each).
  const int nrg = std::min(kNT / n, 1 + (int)sqrtf((float)nr));
  const int col = tid % n, rg = tid / n;
  if (rg < nrg) {
    double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0, acc3 = 0.0;
    if (R == 1) {
      if (rg == 0) acc0 = (double)X[col];
    } else if (R == 2) {
      int rl = rg;
      for (; rl + 3 * nrg < nr; rl += 4 * nrg) {
        acc0 = fma(q0[r0 + rl], (double)X[rl * ld + col], acc0);
        acc1 = fma(q0[r0 + rl + nrg], (double)X[(rl + nrg) * ld + col], acc1);
        acc2 = fma(q0[r0 + rl + 2 * nrg], (double)X[(rl + 2 * nrg) * ld + col], acc2);
        acc3 = fma(q0[r0 + rl + 3 * nrg], (double)X[(rl + 3 * nrg) * ld + col], acc3);
      }
      for (; rl < nr; rl += nrg) acc0 = fma(q0[r0 + rl], (double)X[rl * ld + col], acc0);
    } else {
      // two rows per step: (i, j) of rows rl and rl + nrg advance by 2 nrg
      const int di = nrg / d1, dj = nrg - di * d1;
      int i = (r0 + rg) / d1, j = (r0 + rg) - i * d1;
      int i2 = i + di, j2 = j + dj;
      if (j2 >= d1) {
        j2 -= d1;
        ++i2;
      }
      const int ei = (2 * nrg) / d1, ej = 2 * nrg - ei * d1;
      int rl = rg;
      for (; rl + nrg < nr; rl += 2 * nrg) {
        acc0 = fma(q0[i] * q1[j], (double)X[rl * ld + col], acc0);
        acc1 = fma(q0[i2] * q1[j2], (double)X[(rl + nrg) * ld + col], acc1);
        i += ei;
        j += ej;
        if (j >= d1) {
          j -= d1;
          ++i;
        }
        i2 += ei;
        j2 += ej;
        if (j2 >= d1) {
          j2 -= d1;
          ++i2;
        }
      }
      if (rl < nr) acc0 = fma(q0[i] * q1[j], (double)X[rl * ld + col], acc0);
    }
    s_red[tid] = (acc0 + acc1) + (acc2 + acc3);
  }
  __syncthreads();
  if (tid < n) {
    double sacc = 0.0;
    for (int k = 0; k < nrg; ++k) sacc += s_red[k * n + tid];
    s_part[a.S_off[fl] + tid] += sacc;
  }
  if (R == 3) {
    // m_0[i] = sum_j q_1[j] y_ij (thread per i) and m_1[j] = sum_i q_0[i] y_ij
    // (thread per j) in one pass: serial FMA chains over LDS, two / four
    // partial sums each, no cross-lane reductions.  Rows r = i d1 + j outside
    // [r0, r0 + nr) belong to other segments and count 0.
    const int i_lo = r0 / d1, i_hi = (r0 + nr - 1) / d1, ni = i_hi - i_lo + 1;
    for (int e = tid; e < ni + d1; e += kNT) {
      if (e < ni) {
        const int i = i_lo + e, rb = i * d1 - r0;  // y index of (i, j) is rb + j
        const int j0 = std::max(0, -rb), j1 = std::min(d1, nr - rb);
        double v0 = 0.0, v1 = 0.0;
        int j = j0;
        for (; j + 1 < j1; j += 2) {
          v0 = fma(q1[j], s_y[rb + j], v0);
          v1 = fma(q1[j + 1], s_y[rb + j + 1], v1);
        }
        if (j < j1) v0 = fma(q1[j], s_y[rb + j], v0);
        s_part[a.S_off[a.gfac[g][0]] + i] += v0 + v1;
      } else {
        const int j = e - ni;
        auto term = [&](int ii, double acc) {
          const int r = (i_lo + ii) * d1 + j - r0;
          const bool ok = r >= 0 && r < nr;
          return fma(ok ? q0[i_lo + ii] : 0.0, s_y[ok ? r : 0], acc);
        };
        double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
        int ii = 0;
        for (; ii + 3 < ni; ii += 4) {
          v0 = term(ii, v0);
          v1 = term(ii + 1, v1);
          v2 = term(ii + 2, v2);
          v3 = term(ii + 3, v3);
        }
        for (; ii < ni; ++ii) v0 = term(ii, v0);
        s_part[a.S_off[a.gfac[g][1]] + j] += (v0 + v1) + (v2 + v3);
      }
    }
  }
  __syncthreads();
}

// kV: capacity of the per-factor vectors (sum of an agent's state counts)
template <int kV>
__global__ __launch_bounds__(kNT) void rg_fpi_kernel(Bufs b) {
  __shared__ double s_x[kCap];
  __shared__ double s_y[kRowsCap];
  __shared__ double s_logq[kV], s_qv[kV], s_lp[kV], s_part[kV];
  __shared__ double s_red[kNT];
  __shared__ double s_ltab[256];
  __shared__ int s_obs[kM];
  __shared__ int s_item, s_broken, s_seq;
  __shared__ Agent s_ag;
  __shared__ Unit s_un;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid >> 6;
  for (int i = tid; i < 256; i += kNT) s_ltab[i] = g_logtab[i];
  if (tid == 0) s_seq = b.cnt[1];  // this step's sequence number (team-exchange tags)

  for (;;) {
    if (tid == 0) {
      s_item = atomicAdd(&b.cnt[0], 1);
      s_broken = 0;
    }
    __syncthreads();
    const int idx = s_item;
    if (idx >= b.n_units) break;
    rg_copy_lds(b.units + idx, &s_un, tid);
    __syncthreads();
    rg_copy_lds(b.ag + s_un.agent, &s_ag, tid);
    __syncthreads();
    const Unit& u = s_un;
    const Agent& a = s_ag;
    const int totS = a.totS;
    if (b.trace && tid == 0) b.trace[2 * idx] = wall_clock64();

    for (int f = 0; f < a.F; ++f) {
      const float* pr = b.carry ? b.out + a.pred_off + a.S_off[f] : b.params + a.D_off[f];
      for (int i = tid; i < a.S[f]; i += kNT) s_lp[a.S_off[f] + i] = log_stable_f64(pr[i]);
    }
    for (int i = tid; i < totS; i += kNT) {
      s_logq[i] = 0.0;
      s_part[i] = 0.0;
    }
    if (tid < a.M) s_obs[tid] = b.obs[a.obs_off + tid];
    __syncthreads();

    // log-likelihood rows of every segment: sum over the group's modalities
    // of ln max(A_m[o_m, s], MINVAL), formed as one FP64 log of the product
    // (an out-of-range index is an all-zero one-hot row: no contribution)
    for (int sg = 0; sg < u.nseg; ++sg) {
      const int g = u.g[sg], R = a.gR[g];
      const int n = a.gdim[g][R - 1], ld = n | 1;
      const int r0 = u.r0[sg], nr = u.r1[sg] - r0;
      double* X = s_x + u.lds[sg];
      float* X32 = reinterpret_cast<float*>(s_x) + u.lds[sg];
      // four elements per thread and every modality of the group in flight:
      // all loads of a batch are issued (unconditionally, at clamped
      // addresses; the modality guards are uniform) before any product is
      // formed
      const int ne = nr * n;
      for (int e0 = tid; e0 < ne; e0 += 4 * kNT) {
        int rl[4], kk[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int e = min(e0 + q * kNT, ne - 1);
          rl[q] = e / n;
          kk[q] = e - rl[q] * n;
        }
        float v[kM][4];
#pragma unroll
        for (int m = 0; m < kM; ++m) {
          const int o = m < a.M ? s_obs[m] : -1;
          if (m < a.M && a.mod_group[m] == g && o >= 0 && o < a.O[m]) {
            // A_m is outcome-major: row (outcome o, position r) at (o P + r) n
            // with P positions (K_m / n), so the observed rows are contiguous
            const float* Ar = b.params + a.A_off[m] + ((long long)o * (a.K[m] / n) + r0) * n;
            const long long rs = n;
#pragma unroll
            for (int q = 0; q < 4; ++q) v[m][q] = Ar[rl[q] * rs + kk[q]];
          } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) v[m][q] = 1.0f;
          }
        }
        double prod[4] = {1.0, 1.0, 1.0, 1.0};
#pragma unroll
        for (int m = 0; m < kM; ++m)
#pragma unroll
          for (int q = 0; q < 4; ++q) prod[q] *= (double)fmaxf(v[m][q], kMinVal);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int e = e0 + q * kNT;
          if (e < nr * n) {
            const int rl = e / n, k = e - rl * n;
            const double l = log_f64d(prod[q], s_ltab);
            if (u.x32)
              X32[rl * ld + k] = (float)l;
            else
              X[rl * ld + k] = l;
          }
        }
      }
    }
    __syncthreads();

    // a single unit's log q of sweep it > 0 is s_lp + s_part, formed inside
    // the softmax; a team's is summed over its units into s_logq
    const bool single = a.team == 1;
    for (int it = 0; it < a.iters; ++it) {
      if (single && it > 0)
        rg_softmax(a, s_lp, s_part, s_qv, lane, wave);
      else
        rg_softmax(a, s_logq, nullptr, s_qv, lane, wave);
      __syncthreads();
      for (int sg = 0; sg < u.nseg; ++sg) {
        if (u.x32)
          rg_fpi_segment(a, u.g[sg], u.r0[sg], u.r1[sg] - u.r0[sg], reinterpret_cast<const float*>(s_x) + u.lds[sg],
                         s_y, s_red, s_qv, s_part, tid);
        else
          rg_fpi_segment(a, u.g[sg], u.r0[sg], u.r1[sg] - u.r0[sg], s_x + u.lds[sg], s_y, s_red, s_qv, s_part, tid);
      }
      if (a.team > 1) {
        // exchange partial marginals with the other units of this agent
        // (double-buffered by sweep parity: a unit overwrites its slot of
        // buffer it & 1 only after it has read every unit's sweep it - 1
        // partials, which each unit publishes only after reading sweep it - 2
        // from buffer it & 1).  Each partial is published as one agent-scope
        // relaxed atomic 64-bit store (device-coherence scope bits, no
        // L2-flushing fence) with the sweep's sequence number in its low
        // mantissa bits; readers poll the partials themselves with
        // agent-scope relaxed atomic loads until every tag matches, so a
        // sweep costs one round trip instead of a store acknowledgement, an
        // arrival count and a load.  A slot holds the sweep's tag or that of
        // two sweeps earlier (or the fresh-buffer tag), never a match by
        // accident; the poll is bounded (on timeout the agent's qs are NaN
        // and the plan's status flag is set).
        const unsigned long long tag =
            ((unsigned long long)(unsigned)s_seq * (unsigned)a.iters + (unsigned)it) & kTagMask;
        unsigned long long* part = reinterpret_cast<unsigned long long*>(b.part + a.part_off) +
                                   (long long)(it & 1) * a.team * totS;
        if (wave == 0)
          for (int i = lane; i < totS; i += kWave) {
            const unsigned long long bits = ((unsigned long long)__double_as_longlong(s_part[i]) & ~kTagMask) | tag;
            __hip_atomic_store(part + u.slice * totS + i, bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          }
        for (int i = tid; i < totS; i += kNT) {
          const unsigned long long* pp = part + i;
          double s = 0.0;
          for (int s0 = 0; s0 < a.team; s0 += 16) {
            const int nq = min(16, a.team - s0);
            unsigned long long v[16];
#pragma unroll
            for (int q = 0; q < 16; ++q)
              v[q] = q < nq ? __hip_atomic_load(pp + (long long)(s0 + q) * totS, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                            : tag;
            long long spins = 0;
            for (;;) {
              bool ready = true;
#pragma unroll
              for (int q = 0; q < 16; ++q) ready = ready && (v[q] & kTagMask) == tag;
              // after a timeout no thread of the unit waits again: the flag is
              // read with an LDS atomic so the poll cannot be hoisted
              if (ready || __hip_atomic_load(&s_broken, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) break;
              if (++spins > kSpinMax) {
                __hip_atomic_store(&s_broken, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                atomicOr(&b.cnt[3], 1);
                break;
              }
              __builtin_amdgcn_s_sleep(2);
#pragma unroll
              for (int q = 0; q < 16; ++q)
                if ((v[q] & kTagMask) != tag)
                  v[q] = __hip_atomic_load(pp + (long long)(s0 + q) * totS, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
#pragma unroll
            for (int q = 0; q < 16; ++q)
              if (q < nq) s += __longlong_as_double((long long)(v[q] & ~kTagMask));
          }
          s_logq[i] = s_lp[i] + s;
          s_part[i] = 0.0;
        }
        __syncthreads();
      }
    }
    if (single && a.iters > 0)
      rg_softmax(a, s_lp, s_part, s_qv, lane, wave);
    else
      rg_softmax(a, s_logq, nullptr, s_qv, lane, wave);
    __syncthreads();

    if (u.slice == 0) {
      // posterior out (FP32) and as the level -1 history vector (FP64)
      double* hist = b.hist + a.hist_off;
      const bool bad = s_broken != 0;
      for (int i = tid; i < totS; i += kNT) {
        b.out[a.qs_off + i] = bad ? NAN : (float)s_qv[i];
        hist[i] = s_qv[i];
      }
    }
    if (b.trace && tid == 0) b.trace[2 * idx + 1] = wall_clock64();
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Rollout (pymdp/control.py:351-389), one launch per level t after the FPI
// kernel, one workgroup per task (at most kNT outputs).  Output o of a task is
// state s of entry i = o / S_f, entry i being history h with action u:
//   q_f[t][h][s] = sum_k B_f[u][k][s] w_i[k],  w_i = x_j q_{dep_j}[t-1][parent_j]
// (t = 0: the parents are the posterior), one thread per output, loads in
// flight sixteen k steps at a time (one-dependency B).  B_f is stored [U_f][K_B][S_f], so the lanes
// of one history read consecutive states (coalesced) and the same w_i[k] (one
// broadcast).  Measured on the 4096-agent workload: 74 us per level; staging
// the B blocks and parents in LDS (three load rounds per task instead of ten)
// and a fixed grid striding over the tasks were both slower (133-153 us).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kNT) void rg_roll_kernel(Bufs b, int task0) {
  const RollTask tk = b.roll[task0 + blockIdx.x];
  const int Sf = tk.Sf, KB = tk.KB, rb = tk.rb, ne = 2 + rb;
  const int o = threadIdx.x;
  if (o >= tk.nh * Sf) return;
  const int i = o / Sf, s = o - i * Sf;
  const int* en = b.iarena + tk.ent + (long long)i * ne;
  const int h = en[0], u = en[1];
  const float* Bu = b.params + tk.b_off + (long long)u * KB * Sf + s;
  double acc0 = 0.0, acc1 = 0.0;
  if (rb == 1) {
    const double* w = b.hist + tk.par[0] + (tk.t == 0 ? 0 : (long long)en[2] * tk.dd[0]);
    int k = 0;
    for (; k + 16 <= KB; k += 16) {
      float bv[16];
      double wv[16];
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        bv[j] = Bu[(long long)(k + j) * Sf];
        wv[j] = w[k + j];
      }
#pragma unroll
      for (int j = 0; j < 16; j += 2) {
        acc0 = fma((double)bv[j], wv[j], acc0);
        acc1 = fma((double)bv[j + 1], wv[j + 1], acc1);
      }
    }
    for (; k < KB; ++k) acc0 = fma((double)Bu[(long long)k * Sf], w[k], acc0);
  } else {
    const double* qp[kRB];
    for (int d = 0; d < rb; ++d)
      qp[d] = b.hist + tk.par[d] + (tk.t == 0 ? 0 : (long long)en[2 + d] * tk.dd[d]);
    for (int k = 0; k < KB; ++k) {
      double w;
      if (rb == 2) {
        const int k1 = k / tk.dd[1];
        w = qp[0][k1] * qp[1][k - k1 * tk.dd[1]];
      } else {
        const int k01 = k / tk.dd[2], k2 = k - k01 * tk.dd[2];
        const int k1 = k01 / tk.dd[1];
        w = qp[0][k1] * qp[1][k01 - k1 * tk.dd[1]] * qp[2][k2];
      }
      acc0 = fma((double)Bu[(long long)k * Sf], w, acc0);
    }
  }
  b.hist[tk.out + (long long)h * Sf + s] = acc0 + acc1;
}

// ---------------------------------------------------------------------------
// EFE stages
// ---------------------------------------------------------------------------
// Stage 1 on FP64 vector FMAs, one row per thread.  Modality m's row space
// is [O1][d01] (row e = o * d01 + pos, pos = i * d1 + j): row e of
// [A ; sum_o A ln A] is A_m[o, pos, :] (FP32, outcome-major in the params
// arena, so row e sits at A_m + e d2) for o < O and row pos of the
// modality's sum_o A ln A table (FP32)
// for o = O.  For every concatenated last-axis history n of the modality:
//   Z1[(n * O1 + o) * d01 + pos] = row_e . q_n   (summed in FP64, stored FP32)
// A thread loads its row (at most 64 floats) into registers with 16-byte
// loads of 4-byte alignment, all in flight together; the history vectors are
// staged in LDS 16 columns at a time and read by every lane at the same
// address (a broadcast); the stores of one column are one contiguous run of
// the item's rows.  No MFMA: the FP64 vector rate of gfx950 equals its FP64
// matrix rate, most modalities have few history columns (a 16-column MFMA
// block would be mostly padding), and the row-per-thread form keeps every
// operand load in flight at once instead of staging A through LDS.
constexpr int kE1Cols = 16;  // history columns staged in LDS at a time (stages 1, 2)
constexpr int kE3Cols = 8;   // the same for stage 3
// threads per stage-1 row: two for rows of more than 8 vectors (d2 > 32)
__host__ __device__ constexpr int rg_e1_tpr(int d2) { return d2 > 32 ? 2 : 1; }
constexpr int kE1Ld = 64;    // LDS row stride (doubles) of a staged history column
typedef float rg_f4u __attribute__((ext_vector_type(4), aligned(4)));


// The item's rows with NV = ceil(d2 / 4) 16-byte vectors each (a
// compile-time trip count: no per-vector guards), TPR threads per row
// (adjacent lanes, NVT = ceil(NV / TPR) vectors each, the second thread's
// last one zero when NV is odd; their partial sums are combined by a DPP
// lane swap).  Rows of more than 8 vectors take two threads: a thread
// holding up to 64 floats would leave too few registers for the LDS
// operands at four waves per SIMD (the compiler spills).
template <int STAGE, int NV, int NT, int TPR>
__device__ __forceinline__ void rg_rows(const Bufs& b, const E1Rec& r, double* s_q, double* s_t, int tid) {
  static_assert(TPR == 1 || TPR == 2, "threads per row");
  constexpr int NVT = (NV + TPR - 1) / TPR;  // vectors per thread
  // history columns per chunk: 16, stage 3 8 (its score terms of a chunk
  // sit in LDS, 8 x 256 doubles)
  constexpr int CC = STAGE == 3 ? kE3Cols : kE1Cols;
  const int O = r.O, d2 = r.d2, d01 = r.d01, N1 = r.N1;
  const int h = TPR == 1 ? 0 : (tid & 1);  // which part of the row
  const int e = r.e0 + tid / TPR;
  const bool live = e < r.e1;
  const int ec = live ? e : r.e1 - 1;
  // the row: A rows, the table's rows and Z1 rows are all FP32 [.][d2]; the
  // last vector may reach up to 3 floats past the row (the params arena
  // holds at least 3 more floats of the same agent after every A block, the
  // table and Z1 16 bytes of tail padding); those lanes are zeroed below
  const float* src;
  float* z = nullptr;
  double cval = 0.0;  // stage 3: C[o] of the row's outcome (0 for the ambiguity row)
  if (STAGE == 1) {
    // rows e < O d01 are A's rows in storage order, the rest the table's
    const long long ea = (long long)O * d01;
    src = ec < ea ? b.params + r.a_off + (long long)ec * d2 : b.ha + r.h_off + (ec - ea) * d2;
    z = b.z1 + r.z_off + ec;
  } else if (STAGE == 2) {
    const int blk = r.blk, hb2 = ec / blk;
    src = b.z1 + r.a_off + (long long)ec * d2;
    z = b.z2 + r.z_off + (long long)hb2 * N1 * blk + (ec - hb2 * blk);
  } else {
    // rows (pair, o) of Z2 (of Z1 for a rank-1 modality, d01 = 1)
    src = (d01 ? b.z1 : b.z2) + r.a_off + (long long)ec * d2;
    const int o = ec % r.blk;
    cval = o < O ? (double)b.params[r.h_off + o] : 0.0;
  }
  src += 4 * NVT * h;
  rg_f4u v[NVT];
#pragma unroll
  for (int c = 0; c < NVT; ++c) {
    if (TPR == 1 || NV % 2 == 0 || c < NVT - 1 || h == 0)
      v[c] = *reinterpret_cast<const rg_f4u*>(src + 4 * c);
    else
      v[c] = rg_f4u{0.0f, 0.0f, 0.0f, 0.0f};
  }
  const long long zstride = r.blk;  // between history columns of the output
  constexpr int kp = 4 * NVT * TPR;  // staged length of a history column
  for (int c0 = 0; c0 < N1; c0 += CC) {
    const int nb = min(CC, N1 - c0);
    // history columns [c0, c0 + nb) into s_q[n][k], zero for k in [d2, kp)
    constexpr int NQ = (CC * kp + NT - 1) / NT;
    const int tot = nb * kp;
    double hv[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int idx = min(tid + q * NT, tot - 1), n = idx / kp, k = idx - n * kp;
      const int col = c0 + n;
      int t = 0;
#pragma unroll
      for (int l = 1; l < kT; ++l) t += col >= r.ts[l] ? 1 : 0;
      // stage 3 with a padding first axis (hb < 0): one history, q = 1
      hv[q] = STAGE == 3 && r.hb[0] < 0 ? 1.0 : b.hist[r.hb[t] + (long long)(col - r.ts[t]) * d2 + min(k, d2 - 1)];
    }
    if (c0 > 0) lds_barrier();  // the previous chunk's reads are done
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int idx = tid + q * NT;
      if (idx < tot) {
        const int n = idx / kp, k = idx - n * kp;
        s_q[n * kE1Ld + k] = k < d2 ? hv[q] : 0.0;
      }
    }
    lds_barrier();
    if (c0 == 0 && h == (NV - 1) / NVT) {
      // lanes of the row's last vector past its end are zero (the staged
      // history columns are zero there too, but a finite product needs a
      // finite factor)
      constexpr int cl = (NV - 1) % NVT;
#pragma unroll
      for (int i = 1; i < 4; ++i)
        if (4 * (NV - 1) + i >= d2) v[cl][i] = 0.0f;
    }
    // two columns per pass (each row element is converted to FP64 once per
    // pass), two partial sums per column.  The empty asm keeps the
    // conversions inside the pass: hoisted out of the loop they would hold
    // the whole row as doubles (twice the row's VGPRs).
#pragma unroll 1
    for (int n = 0; n < nb; n += 2) {
#pragma unroll
      for (int c = 0; c < NVT; ++c) asm volatile("" : "+v"(v[c]));
      const double* q0 = s_q + n * kE1Ld + 4 * NVT * h;
      const double* q1 = q0 + kE1Ld;
      double a0 = 0.0, a1 = 0.0, b0 = 0.0, b1 = 0.0;
#pragma unroll
      for (int c = 0; c < NVT; ++c) {
        const double x0 = (double)v[c][0], x1 = (double)v[c][1], x2 = (double)v[c][2], x3 = (double)v[c][3];
        a0 = fma(x0, q0[4 * c], a0);
        a1 = fma(x1, q0[4 * c + 1], a1);
        a0 = fma(x2, q0[4 * c + 2], a0);
        a1 = fma(x3, q0[4 * c + 3], a1);
        b0 = fma(x0, q1[4 * c], b0);
        b1 = fma(x1, q1[4 * c + 1], b1);
        b0 = fma(x2, q1[4 * c + 2], b0);
        b1 = fma(x3, q1[4 * c + 3], b1);
      }
      double za = a0 + a1, zb = b0 + b1;
      if (TPR == 2) {
        // lanes 2r and 2r + 1 hold the two halves of row r's sums; each adds
        // its partner's (addition commutes, so both lanes get the same
        // value).  No lane-dependent select around the lane swap: the
        // compiler turns such a select into a branch and the swap inside it
        // then reads the inactive partner.
        za += dpp_mov_f64<0xB1>(za);
        zb += dpp_mov_f64<0xB1>(zb);
      } else {
        // both sums formed here: the compiler would otherwise sink the
        // second column's FMAs into its conditional store and keep every
        // converted row element alive until then
        asm volatile("" : "+v"(za), "+v"(zb));
      }
      if (STAGE == 3) {
        // the row's term of the combination score: qo C[o] - qo ln qo for an
        // outcome row, qo for the ambiguity row (pymdp/control.py:420-509)
        if (live && h == 0) {
          const int rl = tid / TPR, o = ec % r.blk;
          auto term = [&](double qo) {
            double w = qo;
            if (o < O) {
              w = qo * cval;
              if (qo != 0.0) w -= qo * ln_fast(fmax(qo, (double)kMinVal));
            }
            return w;
          };
          s_t[n * (NT / TPR) + rl] = term(za);
          if (n + 1 < nb) s_t[(n + 1) * (NT / TPR) + rl] = term(zb);
        }
      } else if (live && h == 0) {
        z[(long long)(c0 + n) * zstride] = (float)za;
        if (n + 1 < nb) z[(long long)(c0 + n + 1) * zstride] = (float)zb;
      }
    }
    if (STAGE == 3) {
      // score[pair, h0] = sum over the pair's O1 rows, in row order
      lds_barrier();
      const int O1 = r.blk, np = (r.e1 - r.e0) / O1, pair0 = r.e0 / O1;
      for (int idx = tid; idx < np * nb; idx += NT) {
        const int pl = idx / nb, n = idx - pl * nb;
        const double* tp = s_t + n * (NT / TPR) + pl * O1;
        double sum = 0.0;
        for (int oo = 0; oo < O1; ++oo) sum += tp[oo];
        b.score[r.z_off + (long long)(pair0 + pl) * N1 + c0 + n] = sum;
      }
      lds_barrier();  // s_t is rewritten by the next chunk
    }
  }
}

// Stage 3 of a modality whose first axis is padding (rank <= 2: one history,
// q = 1), so qo is the stage-2 output itself: a group of 2^lg >= O1 lanes
// (at most the wave) per (h2, h1) pair, a lane per outcome (a lane of a
// 64-lane group takes outcome 64 too), the terms of the combination score
// summed by a fixed-order DPP group sum; each wave takes groups of pairs
// four batches at a time (loads in flight together), up to 256 pairs per
// work item.
template <int NT>
__device__ __forceinline__ void rg_score_pairs(const Bufs& b, const E1Rec& r, double* s_c, int tid) {
  constexpr int NW = NT / kWave, NB = 4;
  const int O = r.O, O1 = O + 1;
  for (int o = tid; o < O; o += NT) s_c[o] = (double)b.params[r.h_off + o];
  lds_barrier();
  const int lg = ceil_log2(min(O1, kWave)), gl = 1 << lg;
  const int lane = tid & (kWave - 1), wave = tid / kWave;
  const int gpw = kWave >> lg, grp = lane >> lg, sub = lane & (gl - 1);
  const int p0 = r.e0 / O1, np = (r.e1 - r.e0) / O1;
  const float* zb = (r.d01 ? b.z1 : b.z2) + r.a_off + (long long)p0 * O1;
  // the loop bounds are uniform per wave: every lane reaches the group sums
  for (int pb = wave * gpw; pb < np; pb += NB * NW * gpw) {
    double w[NB];
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      const int p = pb + u * NW * gpw + grp;
      w[u] = 0.0;
      if (p < np) {
        const float* z = zb + (long long)p * O1;
        for (int o = sub; o < O1; o += gl) {
          const double qo = (double)z[o];
          double t = qo;
          if (o < O) {
            t = qo * s_c[o];
            if (qo != 0.0) t -= qo * ln_fast(fmax(qo, (double)kMinVal));
          }
          w[u] += t;
        }
      }
    }
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      const int p = pb + u * NW * gpw + grp;
      const double v = group_sum_low_dpp(w[u], lg);
      if (sub == 0 && p < np) b.score[r.z_off + p0 + p] = v;
    }
  }
}

// One work item per workgroup (stage 1: records b.e1, stage 2: b.e2, stage 3:
// b.e3, always 256 threads so that an item holds at least one pair's 65
// rows).  Six
// waves per SIMD (at most 80 VGPRs, no spills; eight spill): the row of a
// thread and the loads in flight per CU are what set the stage's HBM rate.
template <int STAGE, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(6))) void rg_rows_kernel(Bufs b) {
  __shared__ double s_q[kE1Cols * kE1Ld];
  __shared__ double s_t[STAGE == 3 ? kE3Cols * NT : 1];  // stage 3: per-row score terms [column][row]
  __shared__ E1Rec s_rec;
  const int tid = threadIdx.x;
  const long long tbase = STAGE == 1 ? 0 : STAGE == 2 ? b.n_t1 : (long long)b.n_t1 + b.n_t2;
  long long* tr = b.trace ? b.trace + 2 * ((long long)b.n_units + tbase + blockIdx.x) : nullptr;
  if (tr && tid == 0) tr[0] = wall_clock64();
  rg_copy_lds<E1Rec, NT>((STAGE == 1 ? b.e1 : STAGE == 2 ? b.e2 : b.e3) + blockIdx.x, &s_rec, tid);
  lds_barrier();
  if (STAGE == 3 && s_rec.hb[0] < 0) {
    rg_score_pairs<NT>(b, s_rec, s_q, tid);
    if (tr && tid == 0) tr[1] = wall_clock64();
    return;
  }
  switch ((s_rec.d2 + 3) >> 2) {
#define RG_E1_CASE(nv, tpr)                                  \
  case nv:                                                   \
    static_assert(rg_e1_tpr(4 * nv) == tpr, "threads per row"); \
    rg_rows<STAGE, nv, NT, tpr>(b, s_rec, s_q, s_t, tid);    \
    break;
    // threads per row as rg_e1_tpr(d2), which the host tiling follows
    RG_E1_CASE(1, 1) RG_E1_CASE(2, 1) RG_E1_CASE(3, 1) RG_E1_CASE(4, 1) RG_E1_CASE(5, 1) RG_E1_CASE(6, 1)
    RG_E1_CASE(7, 1) RG_E1_CASE(8, 1) RG_E1_CASE(9, 2) RG_E1_CASE(10, 2) RG_E1_CASE(11, 2) RG_E1_CASE(12, 2)
    RG_E1_CASE(13, 2) RG_E1_CASE(14, 2) RG_E1_CASE(15, 2) RG_E1_CASE(16, 2)
#undef RG_E1_CASE
    default:
      break;
  }
  if (tr && tid == 0) tr[1] = wall_clock64();
}

// sum_o A ln A of every (position, k) of a modality, summed in FP64 and
// stored in FP32 (the ambiguity row of the stage-1 product, whose output Z1
// is FP32 too): one thread per element, the outcome loop with eight loads in
// flight.
__global__ __launch_bounds__(kNT) void rg_ha_kernel(Bufs b) {
  const HaTask tk = b.hat[blockIdx.x];
  const int e = tk.e0 + (int)threadIdx.x;
  if (e >= tk.n) return;
  // element (o, pos, k) of the outcome-major A_m at o n + e, e = pos d2 + k
  const long long st = tk.n;
  const float* col = b.params + tk.a_off + e;
  double s0 = 0.0, s1 = 0.0;
  int o = 0;
  for (; o + 8 <= tk.O; o += 8) {
    float v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = col[(o + j) * st];
#pragma unroll
    for (int j = 0; j < 8; j += 2) {
      s0 += xlogx_fast(v[j]);
      s1 += xlogx_fast(v[j + 1]);
    }
  }
  for (; o < tk.O; ++o) s0 += xlogx_fast(col[o * st]);
  b.ha[tk.ha_off + e] = (float)(s0 + s1);
}

// ---------------------------------------------------------------------------
// Per-agent G, q_pi, action
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kNT) void rg_assemble_kernel(Bufs b) {
  __shared__ double s_red[kNW];
  __shared__ float s_redf[kNW];
  __shared__ int s_redi[kNW];
  __shared__ int s_best;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid >> 6;
  const Agent& a = b.ag[blockIdx.x];
  const int P = a.P, T = a.T, F = a.F;
  const int* p2h = b.iarena + a.p2h_off;
  const double* sc = b.score + a.score_off;
  double* logit = b.score + a.logit_off;
  double lmax = -INFINITY;
  // four policies per thread at a time, so that their history-index and
  // score gathers for one (t, m) are in flight together
  constexpr int kPB = 4;
  for (int p0 = tid; p0 < P; p0 += kPB * kNT) {
    double G[kPB];
#pragma unroll
    for (int q = 0; q < kPB; ++q) G[q] = 0.0;
    for (int t = 0; t < T; ++t)
      for (int m = 0; m < a.M; ++m) {
        const int g = a.mod_group[m];
        long long ix[kPB];
#pragma unroll
        for (int q = 0; q < kPB; ++q) ix[q] = a.score_base[t][m];
        for (int j = 0; j < a.gR[g]; ++j) {
          const long long st = a.score_stride[t][m][j];
          const int fj = a.gfac[g][j];
#pragma unroll
          for (int q = 0; q < kPB; ++q) {
            const int p = min(p0 + q * kNT, P - 1);
            ix[q] += (long long)p2h[((long long)p * T + t) * F + fj] * st;
          }
        }
#pragma unroll
        for (int q = 0; q < kPB; ++q) G[q] += sc[ix[q]];
      }
#pragma unroll
    for (int q = 0; q < kPB; ++q) {
      const int p = p0 + q * kNT;
      if (p < P) {
        b.out[a.G_off + p] = (float)G[q];
        double l = (double)a.gamma * G[q];
        if (a.has_E) l += log_stable_f64(b.params[a.E_off + p]);
        logit[p] = l;
        lmax = fmax(lmax, l);
      }
    }
  }
  lmax = wave_max(lmax);
  if (lane == 0) s_red[wave] = lmax;
  __syncthreads();
  double mx = s_red[0];
  for (int w = 1; w < kNW; ++w) mx = fmax(mx, s_red[w]);
  __syncthreads();
  // exp(l - max) replaces the logit (one exp per policy)
  double se = 0.0;
  for (int p = tid; p < P; p += kNT) {
    const double e = exp(logit[p] - mx);
    logit[p] = e;
    se += e;
  }
  se = wave_sum(se);
  if (lane == 0) s_red[wave] = se;
  __syncthreads();
  double tot = 0.0;
  for (int w = 0; w < kNW; ++w) tot += s_red[w];
  // q_pi in FP32 and the first index of its maximum (argmax of the returned q_pi)
  float best = -1.0f;
  int bi = 0x7fffffff;
  for (int p = tid; p < P; p += kNT) {
    const float q = (float)(logit[p] / tot);
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
    s_best = bx;
  }
  __syncthreads();
  // empirical prior for the next step (pymdp/agent.py:825-851): B_f[:, deps,
  // u_f] . (x)q of the chosen first action u, which is the level-0 rollout
  // vector of that policy's first-step history
  {
    const long long base = (long long)s_best * T * F;
    for (int f = 0; f < F; ++f) {
      const double* h = b.hist + a.hist_off + a.hist_lvl[0][f] + (long long)p2h[base + f] * a.S[f];
      for (int i = tid; i < a.S[f]; i += kNT) b.out[a.pred_off + a.S_off[f] + i] = (float)h[i];
    }
  }
  // the FPI queue starts from zero at the next step, whose team-exchange
  // tags continue the sweep sequence
  if (blockIdx.x == 0 && tid == 0) {
    b.cnt[0] = 0;
    b.cnt[1] = (int)((unsigned)b.cnt[1] + 1u);
  }
}

// ---------------------------------------------------------------------------
// Host plan
// ---------------------------------------------------------------------------
// EFE tile sizes (see build_tiles); the defaults are the LDS capacities
struct TileCfg {
  int e1_rows = 256;  // rows per stage-1 work item (= its workgroup size)
  int e2_rows = 256;  // rows per stage-2 work item (= its workgroup size)
};
constexpr int kE3Rows = 256;  // row slots of a stage-3 work item (its workgroup size)

// autotuner candidates
constexpr int kTuneE1R[] = {128, 256};
constexpr int kTuneE2R[] = {128, 256};

struct Plan {
  int device = 0;
  TileCfg tiles;                        // the tiling in use
  size_t cap_t1 = 0, cap_t2 = 0, cap_t3 = 0;  // device tile-array capacities
  int64_t n_agents = 0, sumF = 0, sumM = 0;
  std::vector<Agent> agents;
  std::vector<Unit> units;
  std::vector<Tile> t1, t2, t3;
  std::vector<E1Rec> e1, e2, e3;
  std::vector<HaTask> hat;
  std::vector<RollTask> roll;     // rollout tasks, level by level
  int roll_start[kT + 1] = {};    // first task of each level
  long long ha_total = 0;
  bool params_static = false;     // see pymdp_hip_ragged_set_params_static
  const float* ha_params = nullptr;  // params arena the sum_o A ln A table was formed from
  std::vector<int32_t> iarena;
  std::vector<int64_t> layout;  // see pymdp_hip_ragged_layout
  long long hist_total = 0, score_total = 0, z1_total = 0, z2_total = 0, part_total = 0;
  double efe_macs = 0.0, fpi_macs = 0.0;
  int team_max = 1, totS_max = 1;
  int grid1 = 0, grid_cap = 0;
  bool tuned = false;
  int carry = 0;               // see pymdp_hip_ragged_set_prior_source
  long long* trace = nullptr;  // caller-owned device buffer, see pymdp_hip_ragged_set_trace
  // device
  Agent* d_ag = nullptr;
  Unit* d_units = nullptr;
  Tile *d_t1 = nullptr, *d_t2 = nullptr, *d_t3 = nullptr;
  E1Rec *d_e1 = nullptr, *d_e2 = nullptr, *d_e3 = nullptr;
  HaTask* d_hat = nullptr;
  RollTask* d_roll = nullptr;
  float* d_ha = nullptr;
  int32_t* d_iarena = nullptr;
  double *d_hist = nullptr, *d_score = nullptr, *d_part = nullptr;
  float *d_z1 = nullptr, *d_z2 = nullptr;
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
  for (void* x : {(void*)p->d_ag, (void*)p->d_units, (void*)p->d_t1, (void*)p->d_e1, (void*)p->d_e2, (void*)p->d_e3, (void*)p->d_hat,
                  (void*)p->d_roll,
                  (void*)p->d_ha, (void*)p->d_t2, (void*)p->d_t3,
                  (void*)p->d_iarena, (void*)p->d_hist, (void*)p->d_score, (void*)p->d_z1, (void*)p->d_z2,
                  (void*)p->d_part, (void*)p->d_cnt})
    (void)hipFree(x);
  (void)hipSetDevice(prev);
  delete p;
}

// Stage costs of contracting axes last-to-first in storage order `ax`.
// Row length as the row kernels and the FPI sweeps pay for it: whole 4-float
// vectors, and at least 8 (a row of 2 costs a thread and its loads, stores
// and reductions nearly as much as a row of 8).
static double row_len(int d) { return std::max(8.0, 4.0 * ((d + 3) / 4)); }

// Cost of one level's EFE stages for a storage order (dims outermost
// first, Hs the level's history counts per axis, O outcomes): stage j
// contracts axis j with its Hs[j] histories over rows of that axis.
static double efe_cost(const std::vector<int>& dims, const std::vector<int>& Hs, int O) {
  double rows = O + 1.0, c = 0.0, mult = 1.0;
  for (int d : dims) rows *= d;
  for (int j = (int)dims.size() - 1; j >= 0; --j) {
    mult *= Hs[j];
    rows /= dims[j];
    c += mult * rows * row_len(dims[j]);
  }
  return c + mult * O;
}

// Cost of the FPI sweeps for a storage order: two passes per sweep over the
// rows of the last axis.
static double fpi_cost(const std::vector<int>& dims, int iters) {
  double rows = 1.0;
  for (size_t j = 0; j + 1 < dims.size(); ++j) rows *= dims[j];
  return 2.0 * iters * rows * row_len(dims.back());
}

}  // namespace rg
}  // namespace pmh

using namespace pmh;
using namespace pmh::rg;

// EFE tiling from the agents' layouts.  Stage 1: blocks of e1_rows rows of
// one modality's row space [O1][d01], all its last-axis history columns;
// stage 2: blocks of e2_rows rows of one level's Z1 rows (h2, o, i), all
// the level's middle-dim histories; stage 3: whole (h2, h1) pairs of a level
// (O1 rows each, at most 256 / (threads per row) rows), all first-dim
// histories.  Every tile list is ordered by decreasing cost.
static void build_tiles(Plan* pl, const TileCfg& cfg) {
  struct TileC {
    Tile t;
    double cost;
  };
  std::vector<TileC> t1, t2, t3;
  pl->efe_macs = 0.0;
  for (int ai = 0; ai < (int)pl->agents.size(); ++ai) {
    const Agent& a = pl->agents[ai];
    for (int m = 0; m < a.M; ++m) {
      const int R = a.gR[a.mod_group[m]];
      const int O1 = a.O[m] + 1;
      const int d0 = a.ed[m][0], d1 = a.ed[m][1], d2 = a.ed[m][2];
      auto Hc = [&](int t, int p) { return a.ef[m][p] < 0 ? 1 : a.H[t][a.ef[m][p]]; };
      const int N1 = a.tstart[m][a.T];
      const int d01 = d0 * d1, rows1 = O1 * d01, rpi = cfg.e1_rows / rg_e1_tpr(d2);
      for (int e0 = 0; e0 < rows1; e0 += rpi) {
        const int e1 = std::min(rows1, e0 + rpi);
        t1.push_back(TileC{Tile{ai, m, 0, e0, e1, 0, N1, 0}, (double)(e1 - e0) * d2 * (N1 + 4)});
      }
      pl->efe_macs += (double)d01 * O1 * d2 * N1;
      for (int t = 0; t < a.T; ++t) {
        const int H0 = Hc(t, 0), H1 = Hc(t, 1), H2 = Hc(t, 2);
        if (R != 1) {
          const int rows = H2 * O1 * d0, rpi = cfg.e2_rows / rg_e1_tpr(d1);
          for (int r0 = 0; r0 < rows; r0 += rpi) {
            const int r1 = std::min(rows, r0 + rpi);
            t2.push_back(TileC{Tile{ai, m, t, r0, r1, 0, H1, 0}, (double)(r1 - r0) * d1 * (H1 + 4)});
          }
          pl->efe_macs += (double)rows * d1 * H1;
        }
        // stage-3 items: whole (h2, h1) pairs of O1 rows each, every
        // first-dim history (tile r0 / r1 count rows); a padding first axis
        // (rank <= 2) takes a thread per pair
        const int pairs = H2 * H1;
        const int ppi = a.ef[m][0] < 0 ? kE3Rows : std::max(1, (kE3Rows / rg_e1_tpr(d0)) / O1);
        for (int p0 = 0; p0 < pairs; p0 += ppi) {
          const int p1 = std::min(pairs, p0 + ppi);
          t3.push_back(TileC{Tile{ai, m, t, p0 * O1, p1 * O1, 0, H0, 0}, (double)(p1 - p0) * O1 * d0 * (H0 + 8)});
        }
        pl->efe_macs += (double)pairs * H0 * O1 * d0;
      }
    }
  }
  auto by_cost = [](const TileC& x, const TileC& y) { return x.cost > y.cost; };
  for (auto* v : {&t1, &t2, &t3}) std::stable_sort(v->begin(), v->end(), by_cost);
  pl->t1.clear();
  pl->t2.clear();
  pl->t3.clear();
  pl->e1.clear();
  for (auto& x : t1) pl->t1.push_back(x.t);
  for (const Tile& tl : pl->t1) {
    const Agent& a = pl->agents[tl.agent];
    const int m = tl.m, f = a.ef[m][2];
    E1Rec r;
    std::memset(&r, 0, sizeof(r));
    r.a_off = a.A_off[m];
    r.h_off = a.ha_off[m];
    r.z_off = a.z1_off[m];
    for (int t = 0; t < kT; ++t) r.hb[t] = t < a.T ? a.hist_off + a.hist_lvl[t][f] : 0;
    for (int t = 0; t <= kT; ++t) r.ts[t] = a.tstart[m][t];
    r.O = a.O[m];
    r.d2 = a.ed[m][2];
    r.d01 = a.ed[m][0] * a.ed[m][1];
    r.e0 = tl.r0;
    r.e1 = tl.r1;
    r.N1 = a.tstart[m][a.T];
    r.blk = (a.O[m] + 1) * r.d01;
    pl->e1.push_back(r);
  }
  for (auto& x : t2) pl->t2.push_back(x.t);
  pl->e2.clear();
  for (const Tile& tl : pl->t2) {
    // rows (h2, o, i) of level t's Z1 block, contracted over the middle dim
    // with that level's middle-factor histories
    const Agent& a = pl->agents[tl.agent];
    const int m = tl.m, t = tl.t, f = a.ef[m][1];
    E1Rec r;
    std::memset(&r, 0, sizeof(r));
    const int O1 = a.O[m] + 1, d0 = a.ed[m][0], d1 = a.ed[m][1];
    r.a_off = a.z1_off[m] + (long long)a.tstart[m][t] * O1 * d0 * d1;
    r.z_off = a.z2_off[m][t];
    r.hb[0] = a.hist_off + a.hist_lvl[t][f];
    r.ts[0] = 0;
    for (int l = 1; l <= kT; ++l) r.ts[l] = a.H[t][f];
    r.O = a.O[m];
    r.d2 = d1;
    r.d01 = d0;
    r.e0 = tl.r0;
    r.e1 = tl.r1;
    r.N1 = a.H[t][f];
    r.blk = O1 * d0;
    pl->e2.push_back(r);
  }
  for (auto& x : t3) pl->t3.push_back(x.t);
  pl->e3.clear();
  for (const Tile& tl : pl->t3) {
    // rows (pair, o) of level t's stage-2 output (stage-1 output for a
    // rank-1 modality) contracted over the first dim with that level's
    // first-factor histories, scored per (pair, history)
    const Agent& a = pl->agents[tl.agent];
    const int m = tl.m, t = tl.t, f = a.ef[m][0];
    E1Rec r;
    std::memset(&r, 0, sizeof(r));
    const int H0 = f < 0 ? 1 : a.H[t][f];
    r.a_off = a.z2_off[m][t];
    r.h_off = a.C_off[m] + (a.C_time[m] ? (long long)t * a.O[m] : 0);
    r.z_off = a.score_off + a.score_base[t][m];
    r.hb[0] = f < 0 ? -1 : a.hist_off + a.hist_lvl[t][f];
    r.ts[0] = 0;
    for (int l = 1; l <= kT; ++l) r.ts[l] = H0;
    r.O = a.O[m];
    r.d2 = a.ed[m][0];
    r.d01 = a.gR[a.mod_group[m]] == 1 ? 1 : 0;
    r.e0 = tl.r0;
    r.e1 = tl.r1;
    r.N1 = H0;
    r.blk = a.O[m] + 1;
    pl->e3.push_back(r);
  }
  pl->tiles = cfg;
}

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
  if (num_iter > kTagMaxIters) return einval("num_iter must be <= 1022");
  Plan* pl = new Plan();
  pl->n_agents = n_agents;
  long long pA = 0, pO = 0, pS = 0;  // running offsets: A_dep_flat, modality, factor
  long long pB = 0, pPol = 0;
  long long params = 0, obs = 0, outf = 0, acts = 0, hist = 0, score = 0, z1 = 0, z2 = 0, part = 0;
  std::vector<int64_t> A_off, A_perm, B_off, C_off, D_off, E_off, obs_off, qs_off, G_off, qpi_off, act_off, pred_off;
  struct UnitC {
    std::vector<Unit> units;
    double cost;
  };
  std::vector<UnitC> teams, singles;
  std::vector<RollTask> rolls[kT];
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
          double c = fpi_cost(dims, num_iter);
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
    for (int g = 0; g < a.G; ++g) {
      a.gR[g] = (int)set_order[g].size();
      for (int j = 0; j < a.gR[g]; ++j) {
        a.gfac[g][j] = set_order[g][j];
        a.gdim[g][j] = a.S[set_order[g][j]];
      }
    }
    a.iters = (F == 1 && num_iter > 0) ? 1 : num_iter;
    // FPI units: segments of the groups' stage-0 rows, packed into LDS slices
    // of cap elements (kCap doubles; an agent that needs more than one such
    // unit is packed again into slices of 2 kCap floats, x32 = 1: half as
    // many units exchanging partial marginals every sweep)
    {
      double fc = 0.0;
      for (int g = 0; g < a.G; ++g) {
        const int R = a.gR[g];
        int rows = 1;
        for (int j = 0; j + 1 < R; ++j) rows *= a.gdim[g][j];
        fc += (double)rows * a.gdim[g][R - 1] * (1.0 + 2.0 * a.iters);
      }
      auto pack = [&](int cap_el, int x32) {
        std::vector<Unit> us;
        Unit cur;
        std::memset(&cur, 0, sizeof(cur));
        int used = 0;
        auto flush = [&]() {
          if (cur.nseg > 0) {
            cur.x32 = x32;
            us.push_back(cur);
            std::memset(&cur, 0, sizeof(cur));
            used = 0;
          }
        };
        for (int g = 0; g < a.G; ++g) {
          const int R = a.gR[g];
          const int n = a.gdim[g][R - 1], ld = n | 1;
          int rows = 1;
          for (int j = 0; j + 1 < R; ++j) rows *= a.gdim[g][j];
          int r = 0;
          while (r < rows) {
            const int cap = std::min((cap_el - used) / ld, kRowsCap);
            if (cap <= 0 || cur.nseg == kM) {
              flush();
              continue;
            }
            const int take = std::min(rows - r, cap);
            cur.g[cur.nseg] = g;
            cur.r0[cur.nseg] = r;
            cur.r1[cur.nseg] = r + take;
            cur.lds[cur.nseg] = used;
            cur.nseg++;
            used += take * ld;
            r += take;
          }
        }
        flush();
        return us;
      };
      std::vector<Unit> us = pack(kCap, 0);
      if (us.size() > 1) us = pack(2 * kCap, 1);
      if (us.size() > 1) {
        // a team waits every sweep for its largest unit: re-pack with the
        // smallest slice (in steps of 1/64 of a unit) that keeps the unit
        // count, so the units carry about equal shares
        const size_t k = us.size();
        for (int c = 2 * kCap / 2; c < 2 * kCap; c += 2 * kCap / 64) {
          std::vector<Unit> v = pack(c, 1);
          if (v.size() <= k) {
            us = std::move(v);
            break;
          }
        }
      }
      if (us.empty()) return fail("no FPI work");
      if ((int)us.size() > kTeamMax) return fail("log-likelihood too large for one agent (more than 128 units)");
      a.team = (int)us.size();
      for (int i = 0; i < a.team; ++i) {
        us[i].agent = (int)ai;
        us[i].slice = i;
      }
      pl->fpi_macs += fc;
      // queue order (longest first) by estimated time: every segment of a
      // unit costs about as much per sweep as 800 log-likelihood elements
      // (its barriers, reductions and setup), which dominates small units
      int nseg = 0;
      for (const Unit& un : us) nseg += un.nseg;
      const double est = fc + 800.0 * nseg * (1.0 + 2.0 * a.iters);
      if (a.team > 1) {
        a.part_off = part;
        part += 2LL * a.team * a.totS;
        teams.push_back(UnitC{us, est});
      } else {
        a.part_off = 0;
        singles.push_back(UnitC{us, est});
      }
      pl->team_max = std::max(pl->team_max, a.team);
    }
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
    a.pred_off = outf;
    pred_off.push_back(outf);
    outf += a.totS;
    a.act_off = acts;
    act_off.push_back(acts);
    acts += F;
    // histories, rollout records, policy tables
    a.hist_off = hist;
    long long hl = a.totS;
    for (int t = 0; t < T; ++t)
      for (int f = 0; f < F; ++f) {
        a.H[t][f] = hp.fac[t][f].n;
        a.hist_lvl[t][f] = (int)hl;
        hl += (long long)hp.fac[t][f].n * a.S[f];
      }
    // rollout tasks: per (level, factor) the histories sorted by action, as
    // (h, u, parents) entries in the int arena, at most kNT outputs per task
    for (int t = 0; t < T; ++t)
      for (int f = 0; f < F; ++f) {
        const FactorLevelPlan& lv = hp.fac[t][f];
        const int Sf = a.S[f], nh = lv.n, rb = a.BR[f], KB = a.BK[f], ne = 2 + rb;
        std::vector<int> ord(nh);
        for (int h = 0; h < nh; ++h) ord[h] = h;
        std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return lv.action[x] < lv.action[y]; });
        const long long ent0 = (long long)pl->iarena.size();
        for (int h : ord) {
          pl->iarena.push_back(h);
          pl->iarena.push_back(lv.action[h]);
          for (int j = 0; j < rb; ++j) pl->iarena.push_back(t == 0 ? 0 : lv.parents[(size_t)h * lv.rank + j]);
        }
        RollTask base;
        std::memset(&base, 0, sizeof(base));
        base.b_off = a.B_off[f];
        base.out = a.hist_off + a.hist_lvl[t][f];
        for (int j = 0; j < rb; ++j) {
          const int d = a.Bdep[f][j];
          base.par[j] = a.hist_off + (t == 0 ? a.S_off[d] : a.hist_lvl[t - 1][d]);
          base.dd[j] = a.S[d];
        }
        base.Sf = Sf;
        base.KB = KB;
        base.rb = rb;
        base.t = t;
        for (int e0 = 0; e0 < nh;) {
          RollTask rt = base;
          rt.ent = ent0 + (long long)e0 * ne;
          const int e1 = std::min(nh, e0 + std::max(1, kNT / Sf));
          rt.nh = e1 - e0;
          rolls[t].push_back(rt);
          e0 = e1;
        }
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
    // score layout and EFE tiles
    a.score_off = score;
    long long sl = 0;
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
      }
    a.logit_off = score + sl;
    score += sl + P;
    for (int m = 0; m < M; ++m) {
      const int g = a.mod_group[m], R = a.gR[g];
      // padded rank-3 view
      for (int p = 0; p < 3; ++p) {
        const int j = p - (3 - R);
        a.ed[m][p] = j >= 0 ? a.gdim[g][j] : 1;
        a.ef[m][p] = j >= 0 ? a.gfac[g][j] : -1;
      }
      const int O1 = a.O[m] + 1;
      const int d0 = a.ed[m][0], d1 = a.ed[m][1], d2 = a.ed[m][2];
      auto Hc = [&](int t, int p) { return a.ef[m][p] < 0 ? 1 : a.H[t][a.ef[m][p]]; };
      a.tstart[m][0] = 0;
      for (int t = 0; t < T; ++t) a.tstart[m][t + 1] = a.tstart[m][t] + Hc(t, 2);
      for (int t = T + 1; t <= kT; ++t) a.tstart[m][t] = a.tstart[m][T];
      const int N1 = a.tstart[m][T];
      a.z1_off[m] = z1;
      z1 += (long long)N1 * O1 * d0 * d1;
      // E1 tiles: blocks of (i, j) positions whose A rows and sum_o A ln A
      // rows fit the tile's LDS x blocks of at most 16 concatenated history
      // columns; the modality's sum_o A ln A table ([pos][d2], rg_ha_kernel)
      {
        const int d01 = d0 * d1;
        const long long nel = (long long)d01 * d2;
        a.ha_off[m] = pl->ha_total;
        for (long long e0 = 0; e0 < nel; e0 += kNT)
          pl->hat.push_back(HaTask{a.A_off[m], pl->ha_total, a.O[m], d2, (int)nel, (int)e0});
        pl->ha_total += nel;
      }
      for (int t = 0; t < T; ++t) {
        const int H0 = Hc(t, 0), H1 = Hc(t, 1), H2 = Hc(t, 2);
        if (R == 1) {
          // Z1 of a rank-1 modality already has the stage-2 layout
          a.z2_off[m][t] = a.z1_off[m] + (long long)a.tstart[m][t] * O1;
        } else {
          a.z2_off[m][t] = z2;
          z2 += (long long)H2 * H1 * O1 * d0;
        }
      }
    }
    pl->totS_max = std::max(pl->totS_max, a.totS);
    pl->agents.push_back(a);
    pS += F;
    pO += M;
    pPol += P * T * F;
  }
  // FPI queue: teams (largest first, units of a team consecutive), then single units
  std::stable_sort(teams.begin(), teams.end(), [](const UnitC& x, const UnitC& y) { return x.cost > y.cost; });
  std::stable_sort(singles.begin(), singles.end(), [](const UnitC& x, const UnitC& y) { return x.cost > y.cost; });
  for (auto& v : teams)
    for (auto& u : v.units) pl->units.push_back(u);
  for (auto& v : singles)
    for (auto& u : v.units) pl->units.push_back(u);
  for (int t = 0; t < kT; ++t) {
    pl->roll_start[t] = (int)pl->roll.size();
    pl->roll.insert(pl->roll.end(), rolls[t].begin(), rolls[t].end());
  }
  pl->roll_start[kT] = (int)pl->roll.size();
  build_tiles(pl, TileCfg{});
  if (pl->roll.size() > INT32_MAX || pl->t1.size() > INT32_MAX || pl->t2.size() > INT32_MAX || pl->t3.size() > INT32_MAX) {
    delete pl;
    return einval("too many EFE tiles");
  }
  pl->sumF = pS;
  pl->sumM = pO;
  pl->hist_total = hist;
  pl->score_total = score;
  pl->z1_total = z1;
  pl->z2_total = z2;
  pl->part_total = part;
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
  L.insert(L.end(), pred_off.begin(), pred_off.end());
  *plan_out = pl;
  return PYMDP_HIP_OK;
}

// Copies the host tile lists (within the device capacities) to the device.
static hipError_t upload_tiles(Plan* pl) {
  if (pl->t1.size() > pl->cap_t1 || pl->t2.size() > pl->cap_t2 || pl->t3.size() > pl->cap_t3)
    return hipErrorInvalidValue;
  auto up = [&](void* d, const void* h, size_t bytes) {
    return bytes ? hipMemcpy(d, h, bytes, hipMemcpyHostToDevice) : hipSuccess;
  };
  hipError_t e;
  if ((e = up(pl->d_t1, pl->t1.data(), sizeof(Tile) * pl->t1.size())) != hipSuccess) return e;
  if ((e = up(pl->d_e1, pl->e1.data(), sizeof(E1Rec) * pl->e1.size())) != hipSuccess) return e;
  if ((e = up(pl->d_t2, pl->t2.data(), sizeof(Tile) * pl->t2.size())) != hipSuccess) return e;
  if ((e = up(pl->d_e2, pl->e2.data(), sizeof(E1Rec) * pl->e2.size())) != hipSuccess) return e;
  if ((e = up(pl->d_e3, pl->e3.data(), sizeof(E1Rec) * pl->e3.size())) != hipSuccess) return e;
  return up(pl->d_t3, pl->t3.data(), sizeof(Tile) * pl->t3.size());
}

static int upload(Plan* pl) {
  if (pl->d_ag) return 0;
  auto fail_dev = [&](hipError_t e, const char* what) {
    set_error(std::string("ragged: ") + what + ": " + hipGetErrorString(e));
    return PYMDP_HIP_EDEVICE;
  };
  // persistent FPI grid: every workgroup co-resident (the team barriers rely
  // on it), at least one workgroup more than the largest team
  int cus = 0, per_cu = 0;
  hipError_t e;
  if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, pl->device)) != hipSuccess)
    return fail_dev(e, "hipDeviceGetAttribute");
  if ((e = pl->totS_max <= 256
               ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, rg_fpi_kernel<256>, kNT, 0)
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, rg_fpi_kernel<kF * kDim>, kNT, 0)) !=
      hipSuccess)
    return fail_dev(e, "hipOccupancyMaxActiveBlocksPerMultiprocessor");
  pl->grid_cap = std::max(1, cus * per_cu);

  if (pl->grid_cap <= pl->team_max) {
    set_error("ragged: the device holds " + std::to_string(pl->grid_cap) +
              " FPI workgroups at once, an agent needs a team of " + std::to_string(pl->team_max));
    return PYMDP_HIP_EDEVICE;
  }
  pl->grid1 = std::min<long long>(pl->grid_cap, (long long)pl->units.size());
  pl->grid1 = std::max(pl->grid1, std::min(pl->grid_cap, pl->team_max + 1));
  // tile arrays sized for the smallest candidate tiles (the most tiles), so
  // the autotuner can re-tile without reallocating
  {
    const TileCfg keep = pl->tiles;
    build_tiles(pl, TileCfg{kTuneE1R[0], kTuneE2R[0]});
    pl->cap_t1 = pl->t1.size();
    pl->cap_t2 = pl->t2.size();
    pl->cap_t3 = pl->t3.size();
    build_tiles(pl, keep);
    pl->cap_t1 = std::max(pl->cap_t1, pl->t1.size());
    pl->cap_t2 = std::max(pl->cap_t2, pl->t2.size());
    pl->cap_t3 = std::max(pl->cap_t3, pl->t3.size());
  }
  auto alloc = [&](void** p, size_t bytes) { return hipMalloc(p, std::max<size_t>(bytes, 8)); };
  if ((e = alloc((void**)&pl->d_ag, sizeof(Agent) * pl->agents.size())) != hipSuccess) return fail_dev(e, "hipMalloc");
  if ((e = alloc((void**)&pl->d_units, sizeof(Unit) * pl->units.size())) != hipSuccess) return fail_dev(e, "hipMalloc");
  if ((e = alloc((void**)&pl->d_t1, sizeof(Tile) * pl->cap_t1)) != hipSuccess) return fail_dev(e, "hipMalloc");
  if ((e = alloc((void**)&pl->d_e1, sizeof(E1Rec) * pl->cap_t1)) != hipSuccess) return fail_dev(e, "hipMalloc");
  if ((e = alloc((void**)&pl->d_e2, sizeof(E1Rec) * pl->cap_t2)) != hipSuccess) return fail_dev(e, "hipMalloc");
  if ((e = alloc((void**)&pl->d_e3, sizeof(E1Rec) * pl->cap_t3)) != hipSuccess) return fail_dev(e, "hipMalloc");
  if ((e = alloc((void**)&pl->d_hat, sizeof(HaTask) * pl->hat.size())) != hipSuccess) return fail_dev(e, "hipMalloc");
  if ((e = alloc((void**)&pl->d_roll, sizeof(RollTask) * pl->roll.size())) != hipSuccess)
    return fail_dev(e, "hipMalloc");
  // 16 bytes of tail padding: stage 1 reads the table's rows as 16-byte vectors
  if ((e = alloc((void**)&pl->d_ha, sizeof(float) * pl->ha_total + 16)) != hipSuccess) return fail_dev(e, "hipMalloc");
  if ((e = alloc((void**)&pl->d_t2, sizeof(Tile) * pl->cap_t2)) != hipSuccess) return fail_dev(e, "hipMalloc");
  if ((e = alloc((void**)&pl->d_t3, sizeof(Tile) * pl->cap_t3)) != hipSuccess) return fail_dev(e, "hipMalloc");
  if ((e = alloc((void**)&pl->d_iarena, sizeof(int32_t) * pl->iarena.size())) != hipSuccess)
    return fail_dev(e, "hipMalloc");
  // 16 bytes of tail padding: stages read these buffers as 16-byte vectors
  if ((e = alloc((void**)&pl->d_hist, sizeof(double) * pl->hist_total + 16)) != hipSuccess)
    return fail_dev(e, "hipMalloc");
  if ((e = alloc((void**)&pl->d_score, sizeof(double) * pl->score_total)) != hipSuccess)
    return fail_dev(e, "hipMalloc");
  if ((e = alloc((void**)&pl->d_z1, sizeof(float) * pl->z1_total + 16)) != hipSuccess) return fail_dev(e, "hipMalloc");
  if ((e = alloc((void**)&pl->d_z2, sizeof(float) * pl->z2_total + 16)) != hipSuccess) return fail_dev(e, "hipMalloc");
  if ((e = alloc((void**)&pl->d_part, sizeof(double) * pl->part_total)) != hipSuccess) return fail_dev(e, "hipMalloc");
  if ((e = alloc((void**)&pl->d_cnt, sizeof(int) * 4)) != hipSuccess) return fail_dev(e, "hipMalloc");
  auto up = [&](void* d, const void* h, size_t bytes) {
    return bytes ? hipMemcpy(d, h, bytes, hipMemcpyHostToDevice) : hipSuccess;
  };
  if ((e = up(pl->d_ag, pl->agents.data(), sizeof(Agent) * pl->agents.size())) != hipSuccess ||
      (e = up(pl->d_units, pl->units.data(), sizeof(Unit) * pl->units.size())) != hipSuccess ||
      (e = upload_tiles(pl)) != hipSuccess ||
      (e = up(pl->d_hat, pl->hat.data(), sizeof(HaTask) * pl->hat.size())) != hipSuccess ||
      (e = up(pl->d_roll, pl->roll.data(), sizeof(RollTask) * pl->roll.size())) != hipSuccess ||
      (e = up(pl->d_iarena, pl->iarena.data(), sizeof(int32_t) * pl->iarena.size())) != hipSuccess)
    return fail_dev(e, "hipMemcpy");
  // every partial of a fresh plan carries the fresh-buffer tag (all ones)
  if ((e = hipMemset(pl->d_part, 0xFF, sizeof(double) * std::max<long long>(1, pl->part_total))) != hipSuccess)
    return fail_dev(e, "hipMemset");
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

extern "C" int pymdp_hip_ragged_tables(void* plan, int64_t* sizes16, void* agents_out, void* units_out,
                                       void* tiles_out, int32_t* iarena_out) {
  Plan* pl = static_cast<Plan*>(plan);
  if (!pl) return einval("plan is NULL");
  if (sizes16) {
    std::fill(sizes16, sizes16 + 16, 0);
    sizes16[0] = (int64_t)sizeof(Agent);
    sizes16[1] = (int64_t)sizeof(Unit);
    sizes16[2] = (int64_t)pl->units.size();
    sizes16[3] = (int64_t)sizeof(Tile);
    sizes16[4] = (int64_t)pl->t1.size();
    sizes16[5] = (int64_t)pl->t2.size();
    sizes16[6] = (int64_t)pl->t3.size();
    sizes16[7] = (int64_t)pl->iarena.size();
  }
  if (agents_out) std::memcpy(agents_out, pl->agents.data(), sizeof(Agent) * pl->agents.size());
  if (units_out) std::memcpy(units_out, pl->units.data(), sizeof(Unit) * pl->units.size());
  if (tiles_out) {
    Tile* t = static_cast<Tile*>(tiles_out);
    std::copy(pl->t1.begin(), pl->t1.end(), t);
    std::copy(pl->t2.begin(), pl->t2.end(), t + pl->t1.size());
    std::copy(pl->t3.begin(), pl->t3.end(), t + pl->t1.size() + pl->t2.size());
  }
  if (iarena_out) std::copy(pl->iarena.begin(), pl->iarena.end(), iarena_out);
  return PYMDP_HIP_OK;
}

extern "C" int pymdp_hip_ragged_e1_records(void* plan, void* out, int64_t* count_out, int64_t* record_bytes) {
  Plan* pl = static_cast<Plan*>(plan);
  if (!pl) return einval("plan is NULL");
  if (count_out) *count_out = (int64_t)pl->e1.size();
  if (record_bytes) *record_bytes = (int64_t)sizeof(E1Rec);
  if (out) std::memcpy(out, pl->e1.data(), sizeof(E1Rec) * pl->e1.size());
  return PYMDP_HIP_OK;
}

extern "C" int pymdp_hip_ragged_e2_records(void* plan, void* out, int64_t* count_out, int64_t* record_bytes) {
  Plan* pl = static_cast<Plan*>(plan);
  if (!pl) return einval("plan is NULL");
  if (count_out) *count_out = (int64_t)pl->e2.size();
  if (record_bytes) *record_bytes = (int64_t)sizeof(E1Rec);
  if (out) std::memcpy(out, pl->e2.data(), sizeof(E1Rec) * pl->e2.size());
  return PYMDP_HIP_OK;
}

extern "C" int pymdp_hip_ragged_e3_records(void* plan, void* out, int64_t* count_out, int64_t* record_bytes) {
  Plan* pl = static_cast<Plan*>(plan);
  if (!pl) return einval("plan is NULL");
  if (count_out) *count_out = (int64_t)pl->e3.size();
  if (record_bytes) *record_bytes = (int64_t)sizeof(E1Rec);
  if (out) std::memcpy(out, pl->e3.data(), sizeof(E1Rec) * pl->e3.size());
  return PYMDP_HIP_OK;
}

extern "C" int pymdp_hip_ragged_roll_tasks(void* plan, void* out, int64_t* count_out, int64_t* record_bytes,
                                           int64_t* level_start) {
  Plan* pl = static_cast<Plan*>(plan);
  if (!pl) return einval("plan is NULL");
  if (count_out) *count_out = (int64_t)pl->roll.size();
  if (record_bytes) *record_bytes = (int64_t)sizeof(RollTask);
  if (level_start)
    for (int t = 0; t <= kT; ++t) level_start[t] = pl->roll_start[t];
  if (out) std::memcpy(out, pl->roll.data(), sizeof(RollTask) * pl->roll.size());
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

extern "C" int pymdp_hip_ragged_info(void* plan, int64_t* out16) {
  Plan* pl = static_cast<Plan*>(plan);
  if (!pl || !out16) return einval("plan or out is NULL");
  std::fill(out16, out16 + 16, 0);
  out16[0] = pl->n_agents;
  out16[1] = (int64_t)pl->units.size();
  out16[2] = (int64_t)pl->t1.size();
  out16[3] = (int64_t)pl->t2.size();
  out16[4] = (int64_t)pl->t3.size();
  out16[5] = pl->grid1;
  out16[6] = pl->hist_total;
  out16[7] = pl->score_total;
  out16[8] = pl->z1_total;
  out16[9] = pl->z2_total;
  out16[10] = pl->part_total;
  out16[11] = pl->team_max;
  out16[12] = (int64_t)pl->efe_macs;
  out16[13] = (int64_t)pl->fpi_macs;
  out16[14] = pl->grid_cap;
  out16[15] = pl->ha_total;
  return PYMDP_HIP_OK;
}

static int ragged_launch(Plan* pl, const float* params, const int32_t* obs, float* out, int32_t* act, hipStream_t s,
                         int grid1) {
  Bufs b;
  b.ag = pl->d_ag;
  b.units = pl->d_units;
  b.t1 = pl->d_t1;
  b.e1 = pl->d_e1;
  b.e2 = pl->d_e2;
  b.e3 = pl->d_e3;
  b.hat = pl->d_hat;
  b.roll = pl->d_roll;
  b.ha = pl->d_ha;
  b.t2 = pl->d_t2;
  b.t3 = pl->d_t3;
  b.iarena = pl->d_iarena;
  b.params = params;
  b.obs = obs;
  b.out = out;
  b.act = act;
  b.hist = pl->d_hist;
  b.score = pl->d_score;
  b.z1 = pl->d_z1;
  b.z2 = pl->d_z2;
  b.part = pl->d_part;
  b.cnt = pl->d_cnt;
  b.n_agents = (int)pl->n_agents;
  b.n_units = (int)pl->units.size();
  b.n_t1 = (int)pl->t1.size();
  b.n_t2 = (int)pl->t2.size();
  b.n_t3 = (int)pl->t3.size();
  b.trace = pl->trace;
  b.carry = pl->carry;
  // the sum_o A ln A table: every step, or with static parameters only when
  // the params arena changed
  if (!pl->hat.empty() && (!pl->params_static || pl->ha_params != params)) {
    hipLaunchKernelGGL(rg_ha_kernel, dim3((unsigned)pl->hat.size()), dim3(kNT), 0, s, b);
    pl->ha_params = pl->params_static ? params : nullptr;
  }
  if (pl->totS_max <= 256)
    hipLaunchKernelGGL(rg_fpi_kernel<256>, dim3(grid1), dim3(kNT), 0, s, b);
  else
    hipLaunchKernelGGL(rg_fpi_kernel<kF * kDim>, dim3(grid1), dim3(kNT), 0, s, b);
  for (int t = 0; t < kT; ++t) {
    const int n = pl->roll_start[t + 1] - pl->roll_start[t];
    if (n)
      hipLaunchKernelGGL(rg_roll_kernel, dim3((unsigned)n), dim3(kNT), 0, s, b, pl->roll_start[t]);
  }
  if (b.n_t1) {
    if (pl->tiles.e1_rows == 128)
      hipLaunchKernelGGL((rg_rows_kernel<1, 128>), dim3(b.n_t1), dim3(128), 0, s, b);
    else
      hipLaunchKernelGGL((rg_rows_kernel<1, 256>), dim3(b.n_t1), dim3(256), 0, s, b);
  }
  if (b.n_t2) {
    if (pl->tiles.e2_rows == 128)
      hipLaunchKernelGGL((rg_rows_kernel<2, 128>), dim3(b.n_t2), dim3(128), 0, s, b);
    else
      hipLaunchKernelGGL((rg_rows_kernel<2, 256>), dim3(b.n_t2), dim3(256), 0, s, b);
  }
  if (b.n_t3) hipLaunchKernelGGL((rg_rows_kernel<3, kE3Rows>), dim3(b.n_t3), dim3(kE3Rows), 0, s, b);
  hipLaunchKernelGGL(rg_assemble_kernel, dim3((unsigned)pl->n_agents), dim3(kNT), 0, s, b);
  PMH_HIP_CHECK(hipGetLastError(), "ragged: ");
  return 0;
}

extern "C" int pymdp_hip_ragged_step(void* plan, const float* params, const int32_t* obs, float* out, int32_t* act,
                                     void* stream) {
  Plan* pl = static_cast<Plan*>(plan);
  if (!pl) return einval("plan is NULL");
  if (!pl->d_ag) return einval("plan is not uploaded");
  if (!params || !obs || !out || !act) return einval("arena pointer is NULL");
  if (reinterpret_cast<uintptr_t>(params) % 16 != 0) return einval("the params arena must be 16-byte aligned");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev != pl->device) return einval("plan was created on another device");
  if (ensure_log_table()) {
    set_error("ragged: cannot initialise the log table");
    return PYMDP_HIP_EDEVICE;
  }
  return ragged_launch(pl, params, obs, out, act, static_cast<hipStream_t>(stream), pl->grid1);
}

// Times the persistent FPI grid at 1, 2, ... workgroups per CU up to the
// occupancy limit (3 steps each after one warm-up, hipEvents on `stream`)
// and keeps the fastest for later steps.  Must not be called while `stream`
// is being captured.
extern "C" int pymdp_hip_ragged_autotune(void* plan, const float* params, const int32_t* obs, float* out, int32_t* act,
                                         void* stream, float* best_ms_out) {
  Plan* pl = static_cast<Plan*>(plan);
  if (!pl) return einval("plan is NULL");
  if (!pl->d_ag) return einval("plan is not uploaded");
  if (!params || !obs || !out || !act) return einval("arena pointer is NULL");
  if (reinterpret_cast<uintptr_t>(params) % 16 != 0) return einval("the params arena must be 16-byte aligned");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)
    return einval("autotune cannot run while the stream is being captured");
  if (ensure_log_table()) {
    set_error("ragged: cannot initialise the log table");
    return PYMDP_HIP_EDEVICE;
  }
  int cus = 0;
  PMH_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, pl->device), "ragged: ");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  PMH_HIP_CHECK(hipEventCreate(&e0), "ragged: ");
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    set_error("ragged: hipEventCreate failed");
    return PYMDP_HIP_EDEVICE;
  }
  constexpr int kReps = 4;
  // mean step time over kReps steps after one warm-up step
  auto time_steps = [&](int g1, float* ms) -> int {
    int rc = ragged_launch(pl, params, obs, out, act, s, g1);
    if (rc) return rc;
    if (hipEventRecord(e0, s) != hipSuccess) return PYMDP_HIP_EDEVICE;
    for (int r = 0; r < kReps; ++r)
      if ((rc = ragged_launch(pl, params, obs, out, act, s, g1))) return rc;
    if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(ms, e0, e1) != hipSuccess) {
      set_error("ragged: autotune timing failed");
      return PYMDP_HIP_EDEVICE;
    }
    *ms /= kReps;
    return 0;
  };
  auto run = [&]() -> int {
    // 1. persistent FPI grid: 1, 2, ... workgroups per CU
    float best = 1e30f, ms = 0.f;
    int b1 = pl->grid1, rc = 0;
    const int lo = std::min(pl->grid_cap, pl->team_max + 1);
    for (int w = 1; w * cus <= pl->grid_cap + cus - 1; ++w) {
      const int g1 = std::max(lo, std::min(pl->grid_cap, w * cus));
      if ((rc = time_steps(g1, &ms))) return rc;
      if (ms < best) {
        best = ms;
        b1 = g1;
      }
    }
    pl->grid1 = b1;
    // 2. EFE tile sizes, one stage at a time from the current tiling
    TileCfg bc = pl->tiles;
    if ((rc = time_steps(pl->grid1, &best))) return rc;
    auto trial = [&](TileCfg c) -> int {
      build_tiles(pl, c);
      if (upload_tiles(pl) != hipSuccess) {
        set_error("ragged: tile upload failed");
        return PYMDP_HIP_EDEVICE;
      }
      float t = 0.f;
      const int r = time_steps(pl->grid1, &t);
      if (r) return r;
      if (t < best) {
        best = t;
        bc = c;
      }
      return 0;
    };
    for (int x : kTuneE1R) {
      TileCfg c = bc;
      c.e1_rows = x;
      if (x != bc.e1_rows && (rc = trial(c))) return rc;
    }
    for (int x : kTuneE2R) {
      TileCfg c = bc;
      c.e2_rows = x;
      if (x != bc.e2_rows && (rc = trial(c))) return rc;
    }
    build_tiles(pl, bc);
    if (upload_tiles(pl) != hipSuccess) {
      set_error("ragged: tile upload failed");
      return PYMDP_HIP_EDEVICE;
    }
    if (best_ms_out) *best_ms_out = best;
    return 0;
  };
  const int rc = run();
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  pl->tuned = true;
  return PYMDP_HIP_OK;
}

// 1 if a team barrier of a previous step gave up waiting (its agents' qs are
// NaN), else 0.  Synchronises the device.
extern "C" int pymdp_hip_ragged_status(void* plan, int32_t* flag_out) {
  Plan* pl = static_cast<Plan*>(plan);
  if (!pl || !flag_out) return einval("plan or flag_out is NULL");
  if (!pl->d_cnt) return einval("plan is not uploaded");
  int h[4] = {0, 0, 0, 0};
  PMH_HIP_CHECK(hipDeviceSynchronize(), "ragged: ");
  PMH_HIP_CHECK(hipMemcpy(h, pl->d_cnt, sizeof(h), hipMemcpyDeviceToHost), "ragged: ");
  *flag_out = h[3];
  return PYMDP_HIP_OK;
}

extern "C" int pymdp_hip_ragged_tile_config(void* plan, int32_t* out8) {
  Plan* pl = static_cast<Plan*>(plan);
  if (!pl || !out8) return einval("plan or out8 is NULL");
  const TileCfg& c = pl->tiles;
  const int32_t v[8] = {c.e1_rows, c.e2_rows, kE3Rows, pl->tuned ? 1 : 0, pl->grid1, 0, 0, 0};
  std::memcpy(out8, v, sizeof(v));
  return PYMDP_HIP_OK;
}

extern "C" int pymdp_hip_ragged_set_tiles(void* plan, const int32_t* cfg2) {
  Plan* pl = static_cast<Plan*>(plan);
  if (!pl || !cfg2) return einval("plan or cfg2 is NULL");
  TileCfg c;
  c.e1_rows = cfg2[0];
  c.e2_rows = cfg2[1];
  if ((c.e1_rows != 128 && c.e1_rows != 256) || (c.e2_rows != 128 && c.e2_rows != 256))
    return einval("tile sizes out of range (stage 1 / 2 rows per work item: 128 or 256)");
  build_tiles(pl, c);
  if (pl->d_ag && upload_tiles(pl) != hipSuccess) {
    set_error("ragged: tile upload failed");
    return PYMDP_HIP_EDEVICE;
  }
  return PYMDP_HIP_OK;
}

extern "C" int pymdp_hip_ragged_set_params_static(void* plan, int32_t is_static) {
  Plan* pl = static_cast<Plan*>(plan);
  if (!pl) return einval("plan is NULL");
  if (is_static != 0 && is_static != 1) return einval("params_static must be 0 or 1");
  pl->params_static = is_static != 0;
  pl->ha_params = nullptr;  // formed again by the next step
  return PYMDP_HIP_OK;
}

extern "C" int pymdp_hip_ragged_set_prior_source(void* plan, int32_t carry) {
  Plan* pl = static_cast<Plan*>(plan);
  if (!pl) return einval("plan is NULL");
  if (carry != 0 && carry != 1) return einval("prior source must be 0 (D) or 1 (predicted prior)");
  pl->carry = carry;
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
