// Variational free energy of a one-step posterior: the return_info
// diagnostics of infer_states for the "fpi" / "exact" methods
// (pymdp/inference.py:456-483 -> pymdp/maths.py:448-762, single-step branch):
//
//   neg_entropy         = sum_f sum_s xlogy(q_f, max(q_f, MINVAL))
//   prior_cross_entropy = -sum_f sum_s xlogy(q_f, max(prior_f, MINVAL))
//   accuracy            = sum_m sum_{s over deps(m)} ll_m[s] prod_j q_{dep_j}[s_j]
//   vfe                 = neg_entropy + prior_cross_entropy - accuracy
//
// with ll_m = sum_o obs_m[o] ln max(A_m[o, s], MINVAL) (one row for an index
// observation; an index outside [0, O_m) is an all-zero one-hot row).  One
// workgroup per agent; every sum is formed in FP64 in a fixed order.  The
// accuracy re-reads the observed rows of A (one HBM pass, like the FPI's
// log-likelihood build).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>

#include "../../include/pymdp_hip.h"
#include "common.hpp"

namespace pmh {
namespace vfe {

constexpr int kNT = 256;
constexpr int kMaxF = 16, kMaxM = 16, kMaxR = 8;
constexpr int kMaxTotS = 2048;

struct Desc {
  int F, M, totS, obs_index;
  int S_off[kMaxF + 1];
  int O[kMaxM], R[kMaxM], dep[kMaxM][kMaxR], dim[kMaxM][kMaxR];
  long long K[kMaxM], obs_off[kMaxM], A_bs[kMaxM], prior_bs[kMaxF];
  long long obs_stride;
  const float* A[kMaxM];
  const float* prior[kMaxF];
  const void* obs;
  const float* qs;
  float* out;
};

__device__ __forceinline__ double ln_clip(double x) { return log(fmax(x, (double)kMinVal)); }

__global__ __launch_bounds__(kNT) void vfe_kernel(Desc d) {
  __shared__ double s_q[kMaxTotS];
  __shared__ double s_red[3][kNT / kWave];
  const int b = blockIdx.x, tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const float* q = d.qs + (long long)b * d.totS;
  for (int i = tid; i < d.totS; i += kNT) s_q[i] = (double)q[i];
  __syncthreads();
  double neg = 0.0, cross = 0.0, acc = 0.0;
  for (int f = 0; f < d.F; ++f) {
    const float* p = d.prior[f] + b * d.prior_bs[f];
    for (int i = tid; i < d.S_off[f + 1] - d.S_off[f]; i += kNT) {
      const double qi = s_q[d.S_off[f] + i];
      if (qi != 0.0) {
        neg += qi * ln_clip(qi);
        cross -= qi * ln_clip((double)p[i]);
      }
    }
  }
  for (int m = 0; m < d.M; ++m) {
    const float* Am = d.A[m] + b * d.A_bs[m];
    const long long K = d.K[m];
    const int R = d.R[m], O = d.O[m];
    int o_idx = -1;
    const float* od = nullptr;
    if (d.obs_index) {
      o_idx = static_cast<const int*>(d.obs)[(long long)b * d.M + m];
      if (o_idx < 0 || o_idx >= O) continue;  // all-zero one-hot row: ll_m = 0
    } else {
      od = static_cast<const float*>(d.obs) + b * d.obs_stride + d.obs_off[m];
    }
    for (long long s = tid; s < K; s += kNT) {
      double w = 1.0;
      long long rest = s;
      for (int j = R - 1; j >= 0; --j) {
        const int sj = (int)(rest % d.dim[m][j]);
        rest /= d.dim[m][j];
        w *= s_q[d.S_off[d.dep[m][j]] + sj];
      }
      double ll = 0.0;
      if (d.obs_index) {
        ll = ln_clip((double)__builtin_nontemporal_load(Am + (long long)o_idx * K + s));
      } else {
        for (int o = 0; o < O; ++o) {
          const float ob = od[o];
          if (ob != 0.0f) ll += (double)ob * ln_clip((double)Am[(long long)o * K + s]);
        }
      }
      acc += ll * w;
    }
  }
  neg = wave_sum(neg);
  cross = wave_sum(cross);
  acc = wave_sum(acc);
  if (lane == 0) {
    s_red[0][wave] = neg;
    s_red[1][wave] = cross;
    s_red[2][wave] = acc;
  }
  __syncthreads();
  if (tid == 0) {
    double n = 0.0, c = 0.0, a = 0.0;
    for (int w = 0; w < kNT / kWave; ++w) {
      n += s_red[0][w];
      c += s_red[1][w];
      a += s_red[2][w];
    }
    float* o = d.out + (long long)b * 4;
    o[0] = (float)n;
    o[1] = (float)c;
    o[2] = (float)a;
    o[3] = (float)(n + c - a);
  }
}

static int einval(const std::string& m) {
  set_error("fpi_ffi: vfe: " + m);
  return PYMDP_HIP_EINVAL;
}

}  // namespace vfe
}  // namespace pmh

using namespace pmh;

extern "C" int pymdp_hip_fpi_vfe(const float* const* A_ptrs, const int64_t* A_batch_stride, const void* obs,
                                 int32_t obs_is_index, const float* const* prior_ptrs,
                                 const int64_t* prior_batch_stride, const float* qs, float* vfe_out, int64_t batch,
                                 const int64_t* S, int64_t F, const int64_t* O, int64_t M,
                                 const int64_t* obs_offsets, const int64_t* A_dep_flat,
                                 const int64_t* A_dep_offsets, void* stream) {
  using namespace pmh::vfe;
  set_error("");
  if (!A_ptrs || !A_batch_stride || !obs || !prior_ptrs || !prior_batch_stride || !qs || !vfe_out || !S || !O ||
      !A_dep_flat || !A_dep_offsets)
    return einval("NULL argument");
  if (!obs_is_index && !obs_offsets) return einval("obs_offsets is NULL");
  if (batch <= 0) return einval("batch must be positive");
  if (F < 1 || F > kMaxF) return einval("number of factors must be in [1, 16]");
  if (M < 1 || M > kMaxM) return einval("number of modalities must be in [1, 16]");
  Desc d{};
  d.F = (int)F;
  d.M = (int)M;
  d.obs_index = obs_is_index ? 1 : 0;
  d.S_off[0] = 0;
  for (int f = 0; f < F; ++f) {
    if (S[f] < 1) return einval("S must be positive");
    d.S_off[f + 1] = d.S_off[f] + (int)S[f];
    if (d.S_off[f + 1] > kMaxTotS) return einval("sum of S exceeds 2048");
    d.prior[f] = prior_ptrs[f];
    d.prior_bs[f] = prior_batch_stride[f];
  }
  d.totS = d.S_off[F];
  if (A_dep_offsets[0] != 0) return einval("A_dep_offsets[0] must be 0");
  for (int m = 0; m < M; ++m) {
    const int64_t r0 = A_dep_offsets[m], r1 = A_dep_offsets[m + 1];
    if (r1 <= r0 || r1 - r0 > kMaxR) return einval("modality dependency count must be in [1, 8]");
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
    d.O[m] = (int)O[m];
    d.K[m] = K;
    d.A[m] = A_ptrs[m];
    d.A_bs[m] = A_batch_stride[m];
    d.obs_off[m] = obs_is_index ? m : obs_offsets[m];
  }
  d.obs_stride = obs_is_index ? M : obs_offsets[M];
  d.obs = obs;
  d.qs = qs;
  d.out = vfe_out;
  if (batch > INT32_MAX) return einval("batch too large");
  hipLaunchKernelGGL(vfe_kernel, dim3((unsigned)batch), dim3(kNT), 0, static_cast<hipStream_t>(stream), d);
  PMH_HIP_CHECK(hipGetLastError(), "fpi_ffi: vfe: ");
  return PYMDP_HIP_OK;
}
