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
 * EFE  follows the factor-history pipeline of
 *      pymdp/ffi/neg_efe/neg_efe_cpu_pipeline.cc:66-101 with the shared
 *      contraction of pymdp/ffi/neg_efe/neg_efe_cpu_core.h:595-671: one
 *      agent per OpenMP thread, per level B-rollout per history, modality
 *      scores per distinct history tuple, scatter to policies.  The history
 *      plan is passed in (built once on the host, as the reference memoises
 *      it, :79-86).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXR 8
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
        for (int64_t j = 0; j < inner; ++j) acc += row[j] * suf[j];
      }
      out[k][s] += acc;
    }
  }
}

/* State inference for `batch` agents with index observations. */
int cpu_ref_infer_states(const float* const* A, const int64_t* A_stride, const int32_t* obs_idx,
                         const float* const* prior, const int64_t* prior_stride, float* q_out, int64_t batch,
                         const int64_t* S, int64_t F, const int64_t* O, int64_t M, const int64_t* A_dep_flat,
                         const int64_t* A_dep_off, int32_t num_iter, int32_t nthreads) {
  int64_t totS = 0, soff[64 + 1];
  if (F > 64) return 1;
  soff[0] = 0;
  for (int64_t f = 0; f < F; ++f) soff[f + 1] = soff[f] + S[f];
  totS = soff[F];
  int64_t K[256], maxK = 1, llsum = 0;
  if (M > 256) return 1;
  for (int64_t m = 0; m < M; ++m) {
    K[m] = 1;
    for (int64_t i = A_dep_off[m]; i < A_dep_off[m + 1]; ++i) K[m] *= S[A_dep_flat[i]];
    if (K[m] > maxK) maxK = K[m];
    llsum += K[m];
  }
  const int iters = (F == 1) ? 1 : num_iter;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    float* ll = (float*)malloc(sizeof(float) * (size_t)llsum);
    float* fb = (float*)malloc(sizeof(float) * (size_t)(2 * maxK + 16));
    float* sb = (float*)malloc(sizeof(float) * (size_t)(2 * maxK + 16));
    float* lq = (float*)malloc(sizeof(float) * (size_t)totS);
    float* qv = (float*)malloc(sizeof(float) * (size_t)totS);
    float* lp = (float*)malloc(sizeof(float) * (size_t)totS);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
    for (int64_t b = 0; b < batch; ++b) {
      int64_t off = 0;
      for (int64_t m = 0; m < M; ++m) {
        const int32_t o = obs_idx[b * M + m];
        const float* row = A[m] + b * A_stride[m] + (int64_t)o * K[m];
        for (int64_t k = 0; k < K[m]; ++k)
          ll[off + k] = (o >= 0 && o < O[m]) ? log_stable(row[k]) : 0.0f;
        off += K[m];
      }
      for (int64_t f = 0; f < F; ++f)
        for (int64_t s = 0; s < S[f]; ++s) lp[soff[f] + s] = log_stable(prior[f][b * prior_stride[f] + s]);
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
      for (int64_t f = 0; f < F; ++f) softmax(lq + soff[f], q_out + b * totS + soff[f], S[f]);
    }
    free(ll);
    free(fb);
    free(sb);
    free(lq);
    free(qv);
    free(lp);
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
