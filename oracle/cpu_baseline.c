/*
 * C/OpenMP restatement of the reference's CPU FFI algorithms, used as the
 * timed CPU baseline and as a second CPU checker.  Test infrastructure only
 * (see oracle/__init__.py); never linked into the HIP library.
 *
 * FPI  follows pymdp/ffi/fpi/fpi_cpu_core.h:191-252 (fpi_one_batch) and the
 *      per-modality contractions of :30-171 (rank 1, 2 and the generic
 *      forward-fold / suffix-product chain for higher ranks), with an OpenMP
 *      loop over agents like pymdp/ffi/fpi/fpi_cpu_runner.cc:94-103, but
 *      with the JAX-path numerics: MINVAL = float64 eps floor
 *      (pymdp/maths.py:15), always num_iter sweeps (no early stop), plain
 *      softmax.  The log-likelihood is built per call from A and the
 *      observation index, as the JAX path does (pymdp/maths.py:214-235).
 * EFE  (cpu_ref_agents) follows the factor-history pipeline of
 *      pymdp/ffi/neg_efe/neg_efe_cpu_pipeline.cc:66-101: per level the
 *      factor phase rolls every distinct action history forward through
 *      B (neg_efe_cpu_core.h:500-520, score_factor_history), the modality
 *      phase scores every distinct tuple of dependency histories with the
 *      shared contraction of neg_efe_cpu_core.h:595-671
 *      (score_modality_shared: [A ; H_A] contracted with each inner-factor
 *      history once, then each tuple finished against the Kronecker
 *      product of its outer factors' histories), and the per-policy
 *      scatter of neg_efe_cpu_pipeline.cc:24-62 sums the scores.  Batch
 *      elements run one per OpenMP thread as neg_efe_cpu_entry.cc:76-111
 *      does for Bn >= 32 (omp_helpers.h:106-114).  FP32 throughout (the
 *      reference's f16 packing of the Kronecker weights is not restated:
 *      the JAX path is FP32).  The transposed B and the H_A rows are
 *      prepared once per agent (cpu_ref_prepare), as the reference caches
 *      them across calls by content; the history tables are built once on
 *      the host, as the reference memoises them (:79-86).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXR 8
/* FP32 dot products may be reassociated into SIMD lanes, as the
 * reference's hand-vectorised sgemv / sdot kernels do. */
#define SIMD_SUM _Pragma("omp simd reduction(+ : acc)")
#define MINVAL 2.220446049250313e-16f

static inline float log_stable(float x) { return logf(x > MINVAL ? x : MINVAL); }

static void softmax(const float* x, float* y, int64_t n) {
  float m = x[0];
  for (int64_t i = 1; i < n; ++i) m = x[i] > m ? x[i] : m;
  float s = 0.0f;
  for (int64_t i = 0; i < n; ++i) {
    y[i] = expf(x[i] - m);
    s += y[i];
  }
  for (int64_t i = 0; i < n; ++i) y[i] /= s;
}

/* marginal_k[s_k] += sum_{others} ll[s] prod_{j != k} q_j[s_j] for every k,
 * by the forward fold F_{k+1} = sum_{s_k} F_k q_k and suffix Kronecker
 * products, ~ (K + 1) * prod(S) flops. */
static void modality_marginals(int R, const float* ll, const int64_t* dims, const float* const* q, float* const* out,
                               float* fbuf, float* sbuf) {
  if (R == 1) {
    for (int64_t s = 0; s < dims[0]; ++s) out[0][s] += ll[s];
    return;
  }
  if (R == 2) {
    const int64_t S0 = dims[0], S1 = dims[1];
    for (int64_t a = 0; a < S0; ++a) {
      const float* row = ll + a * S1;
      float acc = 0.0f;
      SIMD_SUM
      for (int64_t c = 0; c < S1; ++c) {
        acc += row[c] * q[1][c];
        out[1][c] += q[0][a] * row[c];
      }
      out[0][a] += acc;
    }
    return;
  }
  int64_t tail[MAXR];  /* prod of dims after k */
  tail[R - 1] = 1;
  for (int k = R - 2; k >= 0; --k) tail[k] = tail[k + 1] * dims[k + 1];
  /* suffix products: suf_k = q_{k+1} x ... x q_{R-1}, stored at sbuf + soff[k] */
  int64_t soff[MAXR];
  int64_t o = 0;
  for (int k = 0; k < R - 1; ++k) {
    soff[k] = o;
    o += tail[k];
  }
  memcpy(sbuf + soff[R - 2], q[R - 1], sizeof(float) * tail[R - 2]);
  for (int k = R - 3; k >= 0; --k) {
    const float* prev = sbuf + soff[k + 1];
    float* dst = sbuf + soff[k];
    for (int64_t i = 0; i < dims[k + 1]; ++i)
      for (int64_t j = 0; j < tail[k + 1]; ++j) dst[i * tail[k + 1] + j] = q[k + 1][i] * prev[j];
  }
  /* forward folds F_k (F_0 = ll), each F_k has prod_{j>=k} dims[j] entries */
  int64_t foff[MAXR];
  o = 0;
  for (int k = 1; k < R; ++k) {
    foff[k] = o;
    o += dims[k] * tail[k];
  }
  for (int k = 0; k < R; ++k) {
    const float* Fk = (k == 0) ? ll : fbuf + foff[k];
    const int64_t inner = tail[k];
    if (k + 1 < R) {
      float* Fn = fbuf + foff[k + 1];
      memset(Fn, 0, sizeof(float) * inner);
      for (int64_t s = 0; s < dims[k]; ++s) {
        const float w = q[k][s];
        const float* src = Fk + s * inner;
        for (int64_t j = 0; j < inner; ++j) Fn[j] += w * src[j];
      }
    }
    for (int64_t s = 0; s < dims[k]; ++s) {
      const float* row = Fk + s * inner;
      float acc = 0.0f;
      if (inner == 1) {
        acc = row[0];
      } else {
        const float* suf = sbuf + soff[k];
        SIMD_SUM
        for (int64_t j = 0; j < inner; ++j) acc += row[j] * suf[j];
      }
      out[k][s] += acc;
    }
  }
}

/* Growable per-thread work buffer. */
typedef struct {
  float* p;
  size_t n;
} fbuf_t;

static float* fbuf_get(fbuf_t* b, size_t n) {
  if (n > b->n) {
    free(b->p);
    b->n = n + n / 2 + 64;
    b->p = (float*)malloc(sizeof(float) * b->n);
  }
  return b->p;
}

typedef struct {
  fbuf_t ll, fb, sb, lq, qv, lp;
} fpi_work_t;

/* One agent's FPI (pymdp/ffi/fpi/fpi_cpu_core.h:191-252 with the JAX-path
 * numerics): row[m] is the observed row of A_m (NULL: index out of range,
 * an all-zero one-hot row, ln-likelihood 0), prior[f] the prior of factor f. */
static void fpi_one(const float* const* row, const int64_t* K, int64_t M, const float* const* prior, const int64_t* S,
                    int64_t F, const int64_t* A_dep_flat, const int64_t* A_dep_off, int iters, fpi_work_t* w,
                    float* q_out) {
  int64_t soff[64 + 1], maxK = 1, llsum = 0;
  soff[0] = 0;
  for (int64_t f = 0; f < F; ++f) soff[f + 1] = soff[f] + S[f];
  const int64_t totS = soff[F];
  for (int64_t m = 0; m < M; ++m) {
    if (K[m] > maxK) maxK = K[m];
    llsum += K[m];
  }
  float* ll = fbuf_get(&w->ll, (size_t)llsum);
  float* fb = fbuf_get(&w->fb, (size_t)(2 * maxK + 16));
  float* sb = fbuf_get(&w->sb, (size_t)(2 * maxK + 16));
  float* lq = fbuf_get(&w->lq, (size_t)totS);
  float* qv = fbuf_get(&w->qv, (size_t)totS);
  float* lp = fbuf_get(&w->lp, (size_t)totS);
  int64_t off = 0;
  for (int64_t m = 0; m < M; ++m) {
    for (int64_t k = 0; k < K[m]; ++k) ll[off + k] = row[m] ? log_stable(row[m][k]) : 0.0f;
    off += K[m];
  }
  for (int64_t f = 0; f < F; ++f)
    for (int64_t s = 0; s < S[f]; ++s) lp[soff[f] + s] = log_stable(prior[f][s]);
  memset(lq, 0, sizeof(float) * (size_t)totS);
  for (int it = 0; it < iters; ++it) {
    for (int64_t f = 0; f < F; ++f) softmax(lq + soff[f], qv + soff[f], S[f]);
    memcpy(lq, lp, sizeof(float) * (size_t)totS);
    off = 0;
    for (int64_t m = 0; m < M; ++m) {
      const int R = (int)(A_dep_off[m + 1] - A_dep_off[m]);
      int64_t dims[MAXR];
      const float* qs[MAXR];
      float* outs[MAXR];
      for (int i = 0; i < R; ++i) {
        const int64_t f = A_dep_flat[A_dep_off[m] + i];
        dims[i] = S[f];
        qs[i] = qv + soff[f];
        outs[i] = lq + soff[f];
      }
      modality_marginals(R, ll + off, dims, qs, outs, fb, sb);
      off += K[m];
    }
  }
  for (int64_t f = 0; f < F; ++f) softmax(lq + soff[f], q_out + soff[f], S[f]);
}

/* State inference for `batch` agents with index observations. */
int cpu_ref_infer_states(const float* const* A, const int64_t* A_stride, const int32_t* obs_idx,
                         const float* const* prior, const int64_t* prior_stride, float* q_out, int64_t batch,
                         const int64_t* S, int64_t F, const int64_t* O, int64_t M, const int64_t* A_dep_flat,
                         const int64_t* A_dep_off, int32_t num_iter, int32_t nthreads) {
  if (F > 64 || M > 256) return 1;
  int64_t K[256], totS = 0;
  for (int64_t f = 0; f < F; ++f) totS += S[f];
  for (int64_t m = 0; m < M; ++m) {
    K[m] = 1;
    for (int64_t i = A_dep_off[m]; i < A_dep_off[m + 1]; ++i) K[m] *= S[A_dep_flat[i]];
  }
  const int iters = (F == 1) ? 1 : num_iter;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    fpi_work_t w;
    memset(&w, 0, sizeof(w));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
    for (int64_t b = 0; b < batch; ++b) {
      const float* row[256];
      const float* pr[64];
      for (int64_t m = 0; m < M; ++m) {
        const int32_t o = obs_idx[b * M + m];
        row[m] = (o >= 0 && o < O[m]) ? A[m] + b * A_stride[m] + (int64_t)o * K[m] : NULL;
      }
      for (int64_t f = 0; f < F; ++f) pr[f] = prior[f] + b * prior_stride[f];
      fpi_one(row, K, M, pr, S, F, A_dep_flat, A_dep_off, iters, &w, q_out + b * totS);
    }
    free(w.ll.p);
    free(w.fb.p);
    free(w.sb.p);
    free(w.lq.p);
    free(w.qv.p);
    free(w.lp.p);
  }
  return 0;
}

/* ---------------------------------------------------------------------------
 * One agent of the batched step: FPI (optional), factor-history neg-EFE,
 * q_pi and the argmax policy.  Field order and types are mirrored by the
 * ctypes Structure in oracle/cpu_ref.py.
 * ------------------------------------------------------------------------- */
typedef struct {
  int32_t F, M, T, P, iters, do_fpi;
  float gamma;
  int32_t pad;
  const int32_t* S;         /* [F] */
  const int32_t* U;         /* [F] */
  const int32_t* O;         /* [M] */
  const int32_t* adep;      /* modality dependencies, flat */
  const int32_t* adep_off;  /* [M + 1] */
  const int32_t* bdep;      /* transition dependencies, flat */
  const int32_t* bdep_off;  /* [F + 1] */
  const int32_t* C_t;       /* [M]: 1 when C[m] is [T, O], 0 when [O] */
  const float* const* A;    /* [M]: [O, S_dep...] */
  const float* const* B;    /* [F]: [S, S_dep..., U] */
  const float* const* C;    /* [M] */
  const float* const* D;    /* [F] */
  const float* E;           /* [P] or NULL (uniform) */
  const int32_t* obs;       /* [M] */
  const int32_t* hn;        /* [T * F]: distinct action histories of (t, f) */
  const int32_t* const* hpar; /* [T * F]: [hn x |bdep_f|] parent history (level t - 1) per dependency */
  const int32_t* const* hact; /* [T * F]: [hn] action of each history */
  const int32_t* tn;          /* [T * M]: distinct dependency-history tuples of (t, m) */
  const int32_t* const* tup;  /* [T * M]: [tn x |adep_m|] level-t history of each dependency */
  const int32_t* const* p2t;  /* [T * M]: [P] tuple of each policy */
  float* qs;    /* [sum S]: output of the FPI, or the input posterior when do_fpi == 0 */
  float* G;     /* [P] */
  float* qpi;   /* [P] */
  int32_t* best; /* [1] first index of the maximum of q_pi */
  float* ha;    /* prepared: H_A row of each modality, sum_m K_m floats */
  float* btr;   /* prepared: B transposed to [U][S][K_b] per factor */
} cpu_agent_t;

static int64_t mod_K(const cpu_agent_t* a, int m) {
  int64_t k = 1;
  for (int i = a->adep_off[m]; i < a->adep_off[m + 1]; ++i) k *= a->S[a->adep[i]];
  return k;
}

static int64_t fac_K(const cpu_agent_t* a, int f) {
  int64_t k = 1;
  for (int i = a->bdep_off[f]; i < a->bdep_off[f + 1]; ++i) k *= a->S[a->bdep[i]];
  return k;
}

/* Floats of the prepared tables of one agent (H_A rows, transposed B). */
int64_t cpu_ref_prepared_floats(const cpu_agent_t* a) {
  int64_t n = 0;
  for (int m = 0; m < a->M; ++m) n += mod_K(a, m);
  for (int f = 0; f < a->F; ++f) n += (int64_t)a->U[f] * a->S[f] * fac_K(a, f);
  return n;
}

/* H_A[k] = -sum_o A[o, k] ln max(A[o, k], MINVAL) (0 ln 0 = 0) and
 * B_tr[u][s][k] = B[s, k, u]; a->ha / a->btr must hold
 * cpu_ref_prepared_floats(a) floats together (ha first, btr after). */
int cpu_ref_prepare(cpu_agent_t* a) {
  float* h = a->ha;
  for (int m = 0; m < a->M; ++m) {
    const int64_t K = mod_K(a, m);
    for (int64_t k = 0; k < K; ++k) {
      float acc = 0.0f;
      for (int o = 0; o < a->O[m]; ++o) {
        const float x = a->A[m][(int64_t)o * K + k];
        acc -= x == 0.0f ? 0.0f : x * log_stable(x);
      }
      h[k] = acc;
    }
    h += K;
  }
  float* bt = a->btr;
  for (int f = 0; f < a->F; ++f) {
    const int64_t K = fac_K(a, f), S = a->S[f], U = a->U[f];
    for (int64_t u = 0; u < U; ++u)
      for (int64_t s = 0; s < S; ++s)
        for (int64_t k = 0; k < K; ++k) bt[(u * S + s) * K + k] = a->B[f][(s * K + k) * U + u];
    bt += U * S * K;
  }
  return 0;
}

/* out = q_{d0}[h0] x q_{d1}[h1] x ... (row-major, first dependency slowest);
 * q[f] holds the level's histories of factor f, S[f] apart. */
static int64_t kron_hist(int R, const int32_t* deps, const int32_t* hist, const float* const* q, const int32_t* S,
                         float* out, float* tmp) {
  int64_t n = 1;
  out[0] = 1.0f;
  for (int i = 0; i < R; ++i) {
    const int32_t f = deps[i];
    const float* v = q[f] + (int64_t)hist[i] * S[f];
    memcpy(tmp, out, sizeof(float) * (size_t)n);
    for (int64_t j = 0; j < n; ++j)
      for (int32_t s = 0; s < S[f]; ++s) out[j * S[f] + s] = tmp[j] * v[s];
    n *= S[f];
  }
  return n;
}

typedef struct {
  fpi_work_t fpi;
  fbuf_t lvl[2], kr, kt, T, sc, qo;
} agent_work_t;

static void agent_one(cpu_agent_t* a, agent_work_t* w) {
  const int F = a->F, M = a->M, T = a->T, P = a->P;
  int64_t S64[64], soff[65], K64[256], Aflat[256 * MAXR], Aoff[257], Kmax = 1;
  soff[0] = 0;
  for (int f = 0; f < F; ++f) {
    S64[f] = a->S[f];
    soff[f + 1] = soff[f] + a->S[f];
  }
  for (int m = 0; m <= M; ++m) Aoff[m] = a->adep_off[m];
  for (int i = 0; i < a->adep_off[M]; ++i) Aflat[i] = a->adep[i];
  for (int m = 0; m < M; ++m) {
    K64[m] = mod_K(a, m);
    if (K64[m] > Kmax) Kmax = K64[m];
  }
  for (int f = 0; f < F; ++f)
    if (fac_K(a, f) > Kmax) Kmax = fac_K(a, f);
  if (a->do_fpi) {
    const float* row[256];
    for (int m = 0; m < M; ++m) {
      const int32_t o = a->obs[m];
      row[m] = (o >= 0 && o < a->O[m]) ? a->A[m] + (int64_t)o * K64[m] : NULL;
    }
    fpi_one(row, K64, M, a->D, S64, F, Aflat, Aoff, F == 1 ? 1 : a->iters, &w->fpi, a->qs);
  }
  /* per-level score offsets and sizes */
  int64_t nsc = 0, hmax = 0, Tmax = 1;
  for (int t = 0; t < T; ++t) {
    int64_t hl = 0;
    for (int f = 0; f < F; ++f) hl += (int64_t)a->hn[t * F + f] * a->S[f];
    if (hl > hmax) hmax = hl;
    for (int m = 0; m < M; ++m) {
      nsc += a->tn[t * M + m];
      const int R = a->adep_off[m + 1] - a->adep_off[m];
      const int fi = a->adep[a->adep_off[m] + R - 1];
      const int64_t n = (int64_t)a->hn[t * F + fi] * (a->O[m] + 1) * (K64[m] / a->S[fi]);
      if (n > Tmax) Tmax = n;
    }
  }
  if ((int64_t)soff[F] > hmax) hmax = soff[F];
  float* lv[2] = {fbuf_get(&w->lvl[0], (size_t)hmax), fbuf_get(&w->lvl[1], (size_t)hmax)};
  float* kr = fbuf_get(&w->kr, (size_t)Kmax);
  float* kt = fbuf_get(&w->kt, (size_t)Kmax);
  float* Tb = fbuf_get(&w->T, (size_t)Tmax);
  float* sc = fbuf_get(&w->sc, (size_t)nsc + 1);
  int Omax = 1;
  for (int m = 0; m < M; ++m)
    if (a->O[m] > Omax) Omax = a->O[m];
  float* qo = fbuf_get(&w->qo, (size_t)Omax);
  /* level -1: the posterior, one history per factor */
  memcpy(lv[0], a->qs, sizeof(float) * (size_t)soff[F]);
  const float* prev[64];
  float* cur[64];
  for (int f = 0; f < F; ++f) prev[f] = lv[0] + soff[f];
  int64_t so = 0;
  for (int t = 0; t < T; ++t) {
    /* factor phase: qs_next[h] = B[f][:, :, u_h] . (x_{deps} prev[parent_h]) */
    float* dst = lv[(t + 1) & 1];
    const float* bt = a->btr;
    for (int f = 0; f < F; ++f) {
      cur[f] = dst;
      const int n = a->hn[t * F + f], R = a->bdep_off[f + 1] - a->bdep_off[f];
      const int64_t S = a->S[f], K = fac_K(a, f);
      const int32_t* par = a->hpar[t * F + f];
      const int32_t* act = a->hact[t * F + f];
      for (int h = 0; h < n; ++h) {
        int32_t ph[MAXR];
        for (int i = 0; i < R; ++i) ph[i] = t == 0 ? 0 : par[h * R + i];
        kron_hist(R, a->bdep + a->bdep_off[f], ph, prev, a->S, kr, kt);
        const float* Bu = bt + (int64_t)act[h] * S * K;
        float* o = dst + (int64_t)h * S;
        for (int64_t s = 0; s < S; ++s) {
          const float* br = Bu + s * K;
          float acc = 0.0f;
          SIMD_SUM
          for (int64_t k = 0; k < K; ++k) acc += br[k] * kr[k];
          o[s] = acc;
        }
      }
      dst += (int64_t)n * S;
      bt += (int64_t)a->U[f] * S * K;
    }
    /* modality phase (neg_efe_cpu_core.h:595-671) */
    const float* ha = a->ha;
    for (int m = 0; m < M; ++m) {
      const int R = a->adep_off[m + 1] - a->adep_off[m];
      const int32_t* deps = a->adep + a->adep_off[m];
      const int fi = deps[R - 1];
      const int64_t Sin = a->S[fi], K = K64[m], Krem = K / Sin, O = a->O[m], rows = O + 1;
      const int Hin = a->hn[t * F + fi];
      /* stage 1: T[h][r][j] = sum_s [A ; H_A][r][j Sin + s] q_in[h][s] */
      for (int h = 0; h < Hin; ++h) {
        const float* qin = cur[fi] + (int64_t)h * Sin;
        float* Th = Tb + (int64_t)h * rows * Krem;
        for (int64_t r = 0; r < rows; ++r) {
          const float* src = r < O ? a->A[m] + r * K : ha;
          for (int64_t j = 0; j < Krem; ++j) {
            const float* x = src + j * Sin;
            float acc = 0.0f;
            SIMD_SUM
            for (int64_t s = 0; s < Sin; ++s) acc += x[s] * qin[s];
            Th[r * Krem + j] = acc;
          }
        }
      }
      /* stage 2: each tuple against the Kronecker product of its outer histories */
      const int n = a->tn[t * M + m];
      const int32_t* tp = a->tup[t * M + m];
      const float* Cm = a->C[m] + (a->C_t[m] ? (int64_t)t * O : 0);
      for (int i = 0; i < n; ++i) {
        const int32_t* hist = tp + (int64_t)i * R;
        kron_hist(R - 1, deps, hist, (const float* const*)cur, a->S, kr, kt);
        const float* Th = Tb + (int64_t)hist[R - 1] * rows * Krem;
        float ent = 0.0f, util = 0.0f, had = 0.0f;
        for (int64_t o = 0; o < O; ++o) {
          float acc = 0.0f;
          SIMD_SUM
          for (int64_t j = 0; j < Krem; ++j) acc += Th[o * Krem + j] * kr[j];
          qo[o] = acc;
        }
        _Pragma("omp simd reduction(+ : had)")
        for (int64_t j = 0; j < Krem; ++j) had += Th[O * Krem + j] * kr[j];
        for (int64_t o = 0; o < O; ++o) {
          ent -= qo[o] == 0.0f ? 0.0f : qo[o] * log_stable(qo[o]);
          util += qo[o] * Cm[o];
        }
        sc[so + i] = ent - had + util;
      }
      so += n;
      ha += K;
    }
    for (int f = 0; f < F; ++f) prev[f] = cur[f];
  }
  /* scatter (neg_efe_cpu_pipeline.cc:24-62), q_pi = softmax(gamma G + ln E), argmax */
  for (int p = 0; p < P; ++p) a->G[p] = 0.0f;
  so = 0;
  for (int t = 0; t < T; ++t)
    for (int m = 0; m < M; ++m) {
      const int32_t* pt = a->p2t[t * M + m];
      for (int p = 0; p < P; ++p) a->G[p] += sc[so + pt[p]];
      so += a->tn[t * M + m];
    }
  const float lnu = logf(1.0f / (float)P);
  for (int p = 0; p < P; ++p) a->qpi[p] = a->gamma * a->G[p] + (a->E ? log_stable(a->E[p]) : lnu);
  softmax(a->qpi, a->qpi, P);
  int32_t bi = 0;
  for (int p = 1; p < P; ++p)
    if (a->qpi[p] > a->qpi[bi]) bi = p;
  a->best[0] = bi;
}

/* n agents, one per OpenMP thread at a time (dynamic schedule). */
int cpu_ref_agents(cpu_agent_t* agents, int64_t n, int32_t nthreads) {
  for (int64_t i = 0; i < n; ++i)
    if (agents[i].F > 64 || agents[i].M > 256 || agents[i].T < 1 || agents[i].P < 1) return 1;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    agent_work_t w;
    memset(&w, 0, sizeof(w));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
    for (int64_t i = 0; i < n; ++i) agent_one(&agents[i], &w);
    fbuf_t* all[] = {&w.fpi.ll, &w.fpi.fb, &w.fpi.sb, &w.fpi.lq, &w.fpi.qv, &w.fpi.lp,
                     &w.lvl[0], &w.lvl[1], &w.kr, &w.kt, &w.T, &w.sc, &w.qo};
    for (size_t k = 0; k < sizeof(all) / sizeof(all[0]); ++k) free(all[k]->p);
  }
  return 0;
}

int cpu_ref_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
