/*
 * pymdp_hip.h - C ABI of the MI355X (gfx950) FPI + EFE hot-path library
 * (libpymdp_hip.so).
 *
 * Every entry point takes plain device pointers, int64 attribute arrays that
 * live in HOST memory, and a hipStream_t (passed as void*; NULL = the null
 * stream).  All work is enqueued on that stream; no entry point synchronises
 * the stream or allocates on the warm path.  Device scratch belongs to the
 * library, one grow-only buffer per (device, stream): calls on different
 * streams may overlap; a buffer that grows is retired, not freed, so a HIP
 * graph captured earlier stays valid until pymdp_hip_release_scratch().
 * Under stream capture nothing is allocated: make one call of a shape (and,
 * for the neg-EFE entries, of a policy set) on the capturing stream outside
 * capture first.
 *
 * Return value: 0 on success, PYMDP_HIP_EINVAL for an invalid argument
 * (the reference's kInvalidArgument), PYMDP_HIP_EDEVICE for a HIP runtime
 * error.  pymdp_hip_last_error() returns the message of the most recent
 * failure on the calling thread, prefixed "fpi_ffi: " or "efe_ffi: " like
 * the reference's ffi::Error messages (pymdp/ffi/common/error_helpers.h:24-27).
 *
 * Layout conventions follow the reference's ragged flat-buffer ABI
 * (pymdp/ffi/registration/xla_register.cc:30-54 and :118-127):
 *   - per-agent buffers are agent-major: [batch, per_agent_count]
 *   - *_offsets arrays are prefix sums of length N+1
 *   - dependency lists are (flat, offsets) pairs of int64
 */
#ifndef PYMDP_HIP_H_
#define PYMDP_HIP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PYMDP_HIP_OK 0
#define PYMDP_HIP_EINVAL 1
#define PYMDP_HIP_EDEVICE 2

/* Largest dependency rank of one modality / factor (reference
 * pymdp/ffi/common/modality_dispatch.h:25, kMaxFfiDependencyRank). */
#define PYMDP_HIP_MAX_DEP_RANK 8

/* Message of the last failed call on this thread ("" if none). */
const char* pymdp_hip_last_error(void);

/* Library version string and gfx target it was built for. */
const char* pymdp_hip_version(void);

/* Frees the device scratch of the calling thread's current device and every
 * retired scratch buffer and plan table; synchronises the device.  Graphs
 * captured before the call must not be replayed after it.  No counterpart in
 * the reference (its XLA handlers free per-executable state with the
 * executable, pymdp/ffi/neg_efe/neg_efe.h:30-39). */
int pymdp_hip_release_scratch(void);

/* Bytes of per-stream scratch held (all devices); *retired_count (if not
 * NULL) receives the number of retired allocations awaiting release. */
int64_t pymdp_hip_scratch_bytes(int64_t* retired_count);

/* Name of the FPI kernel instance (template arguments = innermost width,
 * rows per lane, workgroup size) used by the most recent FPI call on the
 * calling thread.  Diagnostic only; no counterpart in the reference. */
const char* pymdp_hip_last_launch(void);

/*
 * Generalised fused FPI from log-likelihoods.
 * Replaces the XLA FFI target `pymdp_fpi` / `pymdp_fpi_cuda_native`
 * (pymdp/ffi/registration/xla_register.cc:118-158, handler FpiCpu /
 * FpiCudaDevice, pymdp/ffi/fpi/fpi.h:18-35).
 *
 *   ll_flat  [batch, ll_offsets[M]]   per-modality log-likelihood, each ll_m
 *                                     row-major over S[A_dep[m][0]], ...
 *   lp_flat  [batch, lp_offsets[F]]   log prior per factor
 *   q_out    [batch, lp_offsets[F]]   posterior marginals (softmax of log_q)
 *   ll_count, lp_count, q_count       element counts of the three buffers;
 *                                     batch = lp_count / lp_offsets[F]
 *                                     (pymdp/ffi/fpi/fpi_entry.h:55-66)
 *   S[F], ll_offsets[M+1], lp_offsets[F+1], A_dep_flat[...],
 *   A_dep_offsets[M+1]                the reference's five int64 attributes
 *   num_iter                          Jacobi sweeps (always all of them: the
 *                                     reference JAX path has no early stop)
 */
int pymdp_hip_fpi(const float* ll_flat, const float* lp_flat, float* q_out,
                  int64_t ll_count, int64_t lp_count, int64_t q_count,
                  const int64_t* S, int64_t F,
                  const int64_t* ll_offsets, int64_t M,
                  const int64_t* lp_offsets,
                  const int64_t* A_dep_flat, const int64_t* A_dep_offsets,
                  int32_t num_iter, void* stream);

/* Diagnostic timing trace of the fast FPI kernel (no counterpart in the
 * reference): a device buffer of batch * (num_iter + 3) int64 values that
 * later launches of the one-dependency-tuple fast path fill with wall-clock
 * ticks (100 MHz) per agent: start, log-likelihood rows formed, end of each
 * sweep, end.  NULL switches it off.  Process-wide, not for concurrent
 * callers. */
int pymdp_hip_fpi_set_trace(int64_t* trace);
/*
 * Fused state inference: log-likelihood build + FPI in one launch.
 * Replaces the body of inference.update_posterior_states's one-step `fpi`
 * branch (pymdp/inference.py:171-191 -> pymdp/algos.py:88-155), i.e.
 * compute_log_likelihood_per_modality (pymdp/maths.py:266-287) followed by
 * the `pymdp_fpi` kernel.
 *
 *   A_ptrs[M]          device pointer of A[m], each [batch?, O_m, S_deps...]
 *   A_batch_stride[M]  elements between consecutive agents' A[m]
 *                      (O_m * K_m for per-agent A, 0 for an A shared by all)
 *   obs                obs_is_index == 0: float [batch, obs_offsets[M]]
 *                      distributional observations (distr_obs=True);
 *                      obs_is_index == 1: int32 [batch, M] indices, treated
 *                      as jax.nn.one_hot(index, O_m) (pymdp/agent.py:659-673)
 *   prior_ptrs[F]      device pointer of each factor's prior probabilities
 *   prior_batch_stride[F]  elements between consecutive agents' priors
 *                      (log_stable is applied in-kernel, pymdp/algos.py:110)
 *   q_out    [batch, lp_offsets[F]]  posterior marginals
 *   O[M], obs_offsets[M+1]           observation sizes and prefix sums
 *   method_exact       1 -> num_iter forced to 1 (pymdp/inference.py:182)
 * With F == 1 the closed form softmax(sum_m ll_m + ln prior) is returned
 * regardless of num_iter (pymdp/algos.py:101-104).  num_iter == 0 with F > 1
 * returns uniform marginals (a zero-length scan).
 */
int pymdp_hip_infer_states(const float* const* A_ptrs, const int64_t* A_batch_stride,
                           const void* obs, int32_t obs_is_index,
                           const float* const* prior_ptrs, const int64_t* prior_batch_stride,
                           float* q_out, int64_t batch,
                           const int64_t* S, int64_t F,
                           const int64_t* O, int64_t M,
                           const int64_t* obs_offsets, const int64_t* lp_offsets,
                           const int64_t* A_dep_flat, const int64_t* A_dep_offsets,
                           int32_t num_iter, int32_t method_exact, void* stream);

/*
 * pymdp_hip_infer_states with no device allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, kernels and results as pymdp_hip_infer_states.  The
 * per-agent workspace of the streaming FPI variant (dependency tuples of
 * more than about 16K joint states) lives in a fixed 64 MiB device array
 * that each GPU holds until the process ends; a call needing more is
 * rejected with PYMDP_HIP_EINVAL.  Every call on a GPU shares that array:
 * calls that overlap on one device race.
 */
int pymdp_hip_infer_states_noalloc(const float* const* A_ptrs, const int64_t* A_batch_stride,
                                   const void* obs, int32_t obs_is_index,
                                   const float* const* prior_ptrs, const int64_t* prior_batch_stride,
                                   float* q_out, int64_t batch,
                                   const int64_t* S, int64_t F,
                                   const int64_t* O, int64_t M,
                                   const int64_t* obs_offsets, const int64_t* lp_offsets,
                                   const int64_t* A_dep_flat, const int64_t* A_dep_offsets,
                                   int32_t num_iter, int32_t method_exact, void* stream);

/*
 * Variational free energy of the posterior qs [batch, sum S] (e.g. the
 * output of pymdp_hip_infer_states): the return_info diagnostics of
 * update_posterior_states for "fpi" / "exact" (pymdp/inference.py:456-483,
 * pymdp/maths.py:448-762 single-step branch).  Arguments as for
 * pymdp_hip_infer_states; vfe_out [batch, 4] = (neg_entropy,
 * prior_cross_entropy, accuracy, vfe), vfe = neg_entropy +
 * prior_cross_entropy - accuracy (no parameter KL: infer_states passes none).
 * Limits: F, M <= 16, sum S <= 2048, modality rank <= 8.
 */
int pymdp_hip_fpi_vfe(const float* const* A_ptrs, const int64_t* A_batch_stride,
                      const void* obs, int32_t obs_is_index,
                      const float* const* prior_ptrs, const int64_t* prior_batch_stride,
                      const float* qs, float* vfe_out, int64_t batch,
                      const int64_t* S, int64_t F, const int64_t* O, int64_t M,
                      const int64_t* obs_offsets, const int64_t* A_dep_flat,
                      const int64_t* A_dep_offsets, void* stream);

/*
 * Sequence inference over a window of T steps: marginal (method 0) or
 * variational (method 1) message passing, the sequence branch of
 * update_posterior_states (pymdp/inference.py:194-271 -> pymdp/algos.py:
 * 157-500), for factor-local transitions.  A as for pymdp_hip_infer_states;
 * B_ptrs[f] [batch?, S_f, S_f, U_f]; obs [batch, T, M] int32 indices or
 * [batch, T, sum O] distributions (obs_offsets); prior_ptrs[f] [batch?, S_f];
 * actions [batch, T-1, F] int32 or NULL (every U_f must then be 1);
 * valid_steps [batch] int32 or NULL (no masks: every transition counts, a
 * negative action selects the uniform transition); q_out [batch, T, sum S].
 * Limits: T <= 64, F, M <= 8, S_f <= 64, T * sum S <= 1024,
 * T * sum_m prod S[deps_m] <= 4096, modality rank <= 4.  Errors "seq_ffi: ".
 */
int pymdp_hip_sequence_infer(int32_t method, const float* const* A_ptrs, const int64_t* A_batch_stride,
                             const float* const* B_ptrs, const int64_t* B_batch_stride,
                             const void* obs, int32_t obs_is_index,
                             const float* const* prior_ptrs, const int64_t* prior_batch_stride,
                             const int32_t* actions, const int32_t* valid_steps, float* q_out,
                             int64_t batch, int64_t T, const int64_t* S, const int64_t* U, int64_t F,
                             const int64_t* O, int64_t M, const int64_t* obs_offsets,
                             const int64_t* A_dep_flat, const int64_t* A_dep_offsets,
                             int32_t num_iter, float tau, void* stream);

/* Flag bits of the neg-EFE entries; identical to the reference
 * (pymdp/ffi/neg_efe/neg_efe_layout.h:23-42, pymdp/ffi/_efe.py:26-37). */
#define PYMDP_HIP_FLAG_USE_UTILITY (1 << 0)
#define PYMDP_HIP_FLAG_USE_STATES_INFO_GAIN (1 << 1)
#define PYMDP_HIP_FLAG_USE_INDUCTIVE (1 << 2)
#define PYMDP_HIP_FLAG_USE_PARAM_INFO_GAIN (1 << 3) /* needs pA and / or pB */
/* Caller promises A, pA and pB are not modified in place between calls: the
 * tables derived from them (sum_o A ln A, the Dirichlet novelty weights) are
 * kept per (device, stream) and reused while the device pointers, strides,
 * batch, policy set and term flags match, as the reference's warm test
 * (pymdp/ffi/neg_efe/neg_efe_cuda_entry.cc:40-100).  A buffer freed and
 * reallocated at the same address is indistinguishable: call
 * pymdp_hip_param_cache_clear() when parameters change. */
#define PYMDP_HIP_FLAG_MODEL_PARAMS_STATIC (1 << 4)

/* Drops the cached parameter tables of the current device. */
int pymdp_hip_param_cache_clear(void);

/*
 * Negative expected free energy of every policy, optionally with the policy
 * posterior.  Replaces update_posterior_policies[_inductive]
 * (pymdp/control.py:198-349, :795-963) = vmap of compute_neg_efe_policy
 * [_inductive] (pymdp/control.py:604-793) + softmax(gamma G + ln E).
 *
 *   policy_host [P, T, F] int32 HOST array (policies are static agent
 *                      state; the action-history plan built from it is
 *                      cached and matched on the full content)
 *   qs_flat [batch, sum S]            current beliefs qs_init
 *   A_ptrs[M] / A_batch_stride[M]     A[m] [batch?, O_m, S_deps...]
 *   B_ptrs[F] / B_batch_stride[F]     B[f] [batch?, S_f, S_Bdeps..., U_f]
 *   C_ptrs[M] / C_batch_stride[M]     C[m] [batch?, O_m] (C_has_time[m]=0)
 *                                     or [batch?, T, O_m] (C_has_time[m]=1)
 *   pA_ptrs[M] / pA_batch_stride[M]   Dirichlet counts of A (shape of A[m])
 *                                     or pA_ptrs NULL; pB_ptrs[F] /
 *                                     pB_batch_stride[F] likewise for B.  Read
 *                                     only with PYMDP_HIP_FLAG_USE_PARAM_INFO_GAIN
 *                                     (then at least one must be given): the
 *                                     expected information gain about them,
 *                                     pymdp/control.py:511-602, is added to G
 *   E [batch, P] or NULL (uniform 1/P), E_batch_stride
 *   gamma [batch] device or NULL (then gamma_scalar)
 *   I_ptrs[F], I_batch_stride[F], I_depths[F]   inductive matrices
 *                                     [depth_f, S_f]; read only with
 *                                     PYMDP_HIP_FLAG_USE_INDUCTIVE
 *   inductive_epsilon device pointer holding epsilon_count (1 or batch)
 *                     values, or NULL (then epsilon_scalar)
 *   G_out [batch, P]                  neg_efe
 *   q_pi_out [batch, P] or NULL       softmax(gamma G + log_stable(E))
 *   S[F], O[M], U[F], A/B dependency (flat, offsets)    model shape
 *   flags                             PYMDP_HIP_FLAG_* bits
 */
int pymdp_hip_neg_efe(const int32_t* policy_host, int64_t P, int64_t T,
                      const float* qs_flat,
                      const float* const* A_ptrs, const int64_t* A_batch_stride,
                      const float* const* B_ptrs, const int64_t* B_batch_stride,
                      const float* const* C_ptrs, const int64_t* C_batch_stride,
                      const int64_t* C_has_time,
                      const float* const* pA_ptrs, const int64_t* pA_batch_stride,
                      const float* const* pB_ptrs, const int64_t* pB_batch_stride,
                      const float* E, int64_t E_batch_stride,
                      const float* gamma, float gamma_scalar,
                      const float* const* I_ptrs, const int64_t* I_batch_stride,
                      const int64_t* I_depths,
                      const float* inductive_epsilon, int64_t epsilon_count, float epsilon_scalar,
                      float* G_out, float* q_pi_out,
                      int64_t batch,
                      const int64_t* S, int64_t F, const int64_t* O, int64_t M,
                      const int64_t* U,
                      const int64_t* A_dep_flat, const int64_t* A_dep_offsets,
                      const int64_t* B_dep_flat, const int64_t* B_dep_offsets,
                      int32_t flags, void* stream);

/*
 * The reference's packed neg-EFE boundary, `pymdp_neg_efe_all_policies`
 * (pymdp/ffi/registration/xla_register.cc:30-64, pymdp/ffi/neg_efe/neg_efe.h:54-59,
 * packing at pymdp/ffi/_efe.py:583-711): ragged flat buffers with the 13
 * int64 attribute spans and the flags word.  Buffers are agent-major
 * [batch, per-agent size]; C is the (T, O_m)-broadcast form; the output is
 * neg_efe [batch, P] (the caller applies gamma, ln E and the softmax, as in
 * the reference).  policy_host is the [P, T, F] int32 policy matrix in host
 * memory (identical for every agent).  pA_flat / pB_flat share the offsets
 * of A / B; pA_count / pB_count = 0 means absent (as the reference's
 * zero-element buffers, pymdp/ffi/_efe.py:602-613).
 */
int pymdp_hip_neg_efe_all_policies(const int32_t* policy_host, const float* qs_flat,
                                   const float* A_flat, const float* B_flat,
                                   const float* C_flat, const float* I_flat,
                                   const float* pA_flat, int64_t pA_count,
                                   const float* pB_flat, int64_t pB_count,
                                   const float* inductive_epsilon, int64_t epsilon_count,
                                   float* neg_efe_out, int64_t batch,
                                   int64_t P, int64_t T,
                                   const int64_t* S, int64_t F, const int64_t* O, int64_t M,
                                   const int64_t* U,
                                   const int64_t* qs_offsets, const int64_t* A_offsets,
                                   const int64_t* B_offsets, const int64_t* C_offsets,
                                   const int64_t* I_offsets, const int64_t* I_depths,
                                   const int64_t* A_dep_flat, const int64_t* A_dep_offsets,
                                   const int64_t* B_dep_flat, const int64_t* B_dep_offsets,
                                   int32_t flags, void* stream);

/*
 * Adjoint (vector-Jacobian product) of pymdp_hip_infer_states with index
 * observations.  Replaces the backward of the reference's with_jax_grad
 * wrapper around the FPI FFI call (pymdp/ffi/_core.py:282-310, which runs
 * jax.vjp of pymdp/algos.py:88-155); no HIP counterpart in the reference.
 *
 *   q_bar [batch, sum S]            dL/dq of the posterior
 *   A_bar_ptrs[m] [batch, O_m*K_m]  dL/dA_m, written at the observed rows
 *                                   only (caller zero-fills the rest)
 *   prior_bar [batch, sum S]        dL/dprior
 *   q_out, mode_out (optional)      recomputed posterior; 1 where the
 *                                   implicit adjoint was used, 0 unrolled
 *   mode                            0 = unrolled reverse sweep; 1 = implicit
 *                                   differentiation at the fixed point when
 *                                   max|q^(N) - q^(N-1)| <= implicit_tol,
 *                                   else (or if the adjoint iteration does
 *                                   not converge) the unrolled sweep
 * Limits: F, M <= 8, sum S <= 256, ranks <= 3, sum over dependency groups
 * of the joint state counts <= 2048, per-agent A (batch stride > 0).
 */
int pymdp_hip_fpi_adjoint(const float* const* A_ptrs, const int64_t* A_batch_stride,
                          const int32_t* obs_idx,
                          const float* const* prior_ptrs, const int64_t* prior_batch_stride,
                          const float* q_bar, float* const* A_bar_ptrs, float* prior_bar,
                          float* q_out, int32_t* mode_out,
                          int64_t batch, const int64_t* S, int64_t F, const int64_t* O, int64_t M,
                          const int64_t* A_dep_flat, const int64_t* A_dep_offsets,
                          int32_t num_iter, int32_t mode, float implicit_tol, void* stream);

/*
 * Adjoint of the neg-EFE of every policy (horizon 1, factor-local B,
 * utility and state information gain) w.r.t. qs, A, B and C given
 * G_bar = dL/dG [batch, P].  Replaces the backward of with_jax_grad around
 * the neg-EFE FFI call (pymdp/ffi/_core.py:282-310 over
 * pymdp/control.py:604-690).  policy_dev: int32 [P, 1, F] in DEVICE memory.
 * A_bar / B_bar / C_bar / qs_bar are written in full.
 */
int pymdp_hip_neg_efe_adjoint(const int32_t* policy_dev, int64_t P, int64_t T,
                              const float* qs_flat,
                              const float* const* A_ptrs, const int64_t* A_batch_stride,
                              const float* const* B_ptrs, const int64_t* B_batch_stride,
                              const float* const* C_ptrs, const int64_t* C_batch_stride,
                              const float* G_bar, float* qs_bar,
                              float* const* A_bar_ptrs, float* const* B_bar_ptrs,
                              float* const* C_bar_ptrs,
                              int64_t batch, const int64_t* S, int64_t F,
                              const int64_t* O, int64_t M, const int64_t* U,
                              const int64_t* A_dep_flat, const int64_t* A_dep_offsets,
                              const int64_t* B_dep_flat, const int64_t* B_dep_offsets,
                              int32_t flags, void* stream);

/*
 * Dirichlet-parameter learning (pymdp/learning.py:11-234, as driven by
 * Agent.infer_parameters, pymdp/agent.py:416-620), batched over agents; one
 * fused pass over the count tensor (p read once, q and E[q] written once).
 *
 * pymdp_hip_update_pA replaces update_obs_likelihood_dirichlet_m
 * (pymdp/learning.py:11-55) for one modality:
 *   qA[b] = pA[b] + lr[b] * sum_t obs[b, t] (x) qs_dep0[b, t] (x) ... (x) qs_dep{n-1}[b, t]
 *   EA[b] = max(qA[b], MINVAL) normalised over the outcome axis
 *           (dirichlet_expected_value, pymdp/maths.py:843-868)
 * pA [batch?, O, S_dep0, ..., S_dep{n-1}] with batch stride pA_batch_stride
 * (0 = one tensor shared by every agent); obs [batch, T, O] distributional
 * (one-hot rows for index observations); qs_deps[j] [batch, T, S_depj]
 * (batch stride qs_batch_stride[j]); lr [batch] on the device or NULL (then
 * lr_scalar); qA_out, EA_out [batch, O, S_dep0, ...] contiguous.
 */
int pymdp_hip_update_pA(const float* pA, int64_t pA_batch_stride, float* qA_out, float* EA_out,
                        int64_t batch, int64_t T, int64_t O, int32_t n_deps, const int64_t* dep_sizes,
                        const float* obs, int64_t obs_batch_stride, const float* const* qs_deps,
                        const int64_t* qs_batch_stride, const float* lr, float lr_scalar, void* stream);
/*
 * pymdp_hip_update_pB replaces update_state_transition_dirichlet_f
 * (pymdp/learning.py:119-157) for factor `factor`:
 *   qB[b] = pB[b] + lr[b] * sum_{t < T1} qs_next[b, t] (x) qs_dep0[b, t] (x) ... (x) onehot(actions[b, t, factor])
 *   EB[b] = max(qB[b], MINVAL) normalised over the next-state axis
 * pB [batch?, S, S_dep0, ..., U] (reference layout, control axis last);
 * qs_next [batch, T1, S] (the beliefs at t + 1) and qs_deps[j] [batch, T1,
 * S_depj] (at t) with their batch and time strides; actions [batch, T1,
 * num_factors] int32.  An action outside [0, U) contributes nothing
 * (jax.nn.one_hot); mask_all = 1 also drops every t at which any factor's
 * action is negative (the transition mask of pymdp/agent.py:478-520).
 */
int pymdp_hip_update_pB(const float* pB, int64_t pB_batch_stride, float* qB_out, float* EB_out,
                        int64_t batch, int64_t T1, int64_t S, int32_t n_deps, const int64_t* dep_sizes,
                        int64_t U, const float* qs_next, int64_t qs_next_batch_stride,
                        int64_t qs_next_t_stride, const float* const* qs_deps,
                        const int64_t* qs_deps_batch_stride, const int64_t* qs_deps_t_stride,
                        const int32_t* actions, int64_t actions_batch_stride, int32_t num_factors,
                        int32_t factor, int32_t mask_all, const float* lr, float lr_scalar, void* stream);

/*
 * Batched ragged agents: one infer_states -> infer_policies -> sample_action
 * -> update_empirical_prior step for many independent agents of different
 * shapes.  No counterpart in the reference (its FFI handles one agent shape
 * per call); it composes the reference semantics of pymdp/algos.py:88-155
 * (FPI), pymdp/control.py:604-793 (neg-EFE per policy), :963 (q_pi),
 * :115-150 (deterministic action selection, sampling_mode "full") and
 * pymdp/agent.py:825-851 (predicted prior B . q of the chosen action) per
 * agent.
 *
 * Per agent a (in order), hdr[4a..4a+3] = (F, M, T, P).  Concatenated over
 * agents: S, U [F each], O [M each], A_dep_rank [M each] with A_dep_flat,
 * B_dep_rank [F each] with B_dep_flat, C_time [M each] (1 = C_m is [T, O_m]),
 * E_present [1 each] (0 = uniform E), gamma [1 each], policies [P*T*F each,
 * int32].  Limits: F, M <= 8, T <= 4, S, O <= 64, dependency ranks <= 3,
 * num_iter <= 1022 (the sweep number tags the partial marginals exchanged
 * between the workgroups of a large agent).
 * The plan is built on the host; pymdp_hip_ragged_plan_upload allocates and
 * fills its device tables on the current device (the only allocation).
 */
int pymdp_hip_ragged_plan_create(int64_t n_agents, const int64_t* hdr,
                                 const int64_t* S, const int64_t* U, const int64_t* O,
                                 const int64_t* A_dep_rank, const int64_t* A_dep_flat,
                                 const int64_t* B_dep_rank, const int64_t* B_dep_flat,
                                 const int64_t* C_time, const int32_t* E_present,
                                 const float* gamma, const int32_t* policies,
                                 int32_t num_iter, void** plan_out);
int pymdp_hip_ragged_plan_upload(void* plan);

/* Packed layout chosen by the plan, as int64 values:
 *   [params_floats, obs_ints, out_floats, act_ints]
 *   A_off [sum M]            offset of A_m in the params arena,
 *                            outcome-major, storage axes in the plan's
 *                            order: [O_m][leading storage axes][last
 *                            storage axis]
 *   A_perm [sum M * 3]       storage axis j = position A_perm[j] of the
 *                            modality's given dependency list (-1 = unused)
 *   B_off [sum F]            B_f stored [U_f][S_deps...][S_f] (control axis
 *                            first, next-state axis last; given as
 *                            [S_f][S_deps...][U_f])
 *   C_off [sum M]            C_m [O_m] or [T, O_m]
 *   D_off [sum F]            prior D_f [S_f]
 *   E_off [n]                E [P] (or -1)
 *   obs_off [n]              M int32 observation indices in the obs arena
 *   qs_off, G_off, qpi_off [n each]   outputs in the float out arena
 *   act_off [n]              F int32 actions in the action arena
 *   pred_off [n]             predicted prior [sum_f S_f] for the next step in
 *                            the out arena (B_f[:, deps, u_f] . q, u = the
 *                            chosen action)
 * With out == NULL returns the number of values. */
int pymdp_hip_ragged_layout(void* plan, int64_t* out, int64_t n_out);
/* out16 = (agents, FPI units, stage-1 tiles, stage-2 tiles, stage-3 tiles,
 *          FPI grid, history doubles, score doubles, stage-1 floats,
 *          stage-2 floats, team-exchange doubles, largest FPI team,
 *          EFE multiply-adds per step, FPI multiply-adds per step, FPI grid
 *          capacity (workgroups resident at once), sum_o A ln A floats). */
int pymdp_hip_ragged_info(void* plan, int64_t* out16);
/* Host copies of the plan's device tables (for host-side verification):
 * sizes16 = (sizeof agent record, sizeof unit, units, sizeof tile, stage-1
 * tiles, stage-2 tiles, stage-3 tiles, int arena, 0...); tiles_out receives
 * the three tile lists concatenated. */
int pymdp_hip_ragged_tables(void* plan, int64_t* sizes16, void* agents_out,
                            void* units_out, void* tiles_out, int32_t* iarena_out);
/* Host copies of the stage-1 work-item records (per item, a range of rows of
 * one modality: params offset of its A, offset of its sum_o A ln A table,
 * stage-1 output offset, history offset and first column of each level,
 * sizes and the row range); count_out
 * / record_bytes receive the count and the record size; out may be NULL. */
int pymdp_hip_ragged_e1_records(void* plan, void* out, int64_t* count_out, int64_t* record_bytes);
/* Host copies of the stage-2 work-item records (same record layout: per
 * item, a range of one level's stage-1 output rows, the z1 offset of the
 * level's first row, the stage-2 output offset, the history offset of the
 * level's middle-factor histories, their count, the row length and the rows
 * per output block); count_out / record_bytes / out as for the stage-1
 * records. */
int pymdp_hip_ragged_e2_records(void* plan, void* out, int64_t* count_out, int64_t* record_bytes);
/* Host copies of the stage-3 work-item records (same layout: per item, a
 * range of whole (h2, h1) pairs of one level's stage-2 rows, their offset and
 * buffer (d01 = 1: the stage-1 output of a rank-1 modality), the params
 * offset of C at that level, the score offset, the history offset of the
 * level's first-factor histories (-1: padding axis, q = 1), their count, the
 * row length and O + 1 rows per pair). */
int pymdp_hip_ragged_e3_records(void* plan, void* out, int64_t* count_out, int64_t* record_bytes);
/* Host copies of the rollout task records (per task: params offset of B_f,
 * history offsets of its outputs and of each dependency's parent level,
 * int-arena offset of its history records, sizes and level); level_start
 * (kT + 1 = 5 values, may be NULL) receives the first task of each level and
 * the total; count_out / record_bytes receive the count and the record size;
 * out may be NULL. */
int pymdp_hip_ragged_roll_tasks(void* plan, void* out, int64_t* count_out, int64_t* record_bytes,
                                int64_t* level_start);
/* Static parameters (no counterpart in the reference's ragged-free API; the
 * counterpart of PYMDP_HIP_FLAG_MODEL_PARAMS_STATIC): with is_static = 1 the
 * caller promises the A blocks of the params arena are not modified in place
 * between steps, so the sum_o A ln A table is formed by the first step with a
 * given arena pointer and reused; with 0 (the default) every step forms it.
 * Either call invalidates the table.  Graphs captured while the table was
 * valid do not form it. */
int pymdp_hip_ragged_set_params_static(void* plan, int32_t is_static);
/* One step on `stream`: five or six launches (the sum_o A ln A table unless
 * it is valid under static parameters, FPI units, three EFE stages, per-agent
 * assembly), no allocation, no synchronisation.  Steps that overlap on one
 * device race on the plan's scratch.  `params` must be 16-byte aligned
 * (EINVAL otherwise; device allocations are). */
int pymdp_hip_ragged_step(void* plan, const float* params, const int32_t* obs,
                          float* out, int32_t* actions, void* stream);
/* The plan's autotuner (no counterpart in the reference): times the
 * persistent FPI grid (1, 2, ... workgroups per CU up to the occupancy
 * limit), then the rows per stage-1 and per stage-2 work item (128 / 256),
 * each over whole steps, and keeps the fastest; *best_ms = that step time.
 * Not while `stream` is being captured; graphs captured before the call
 * must be captured again (the tile tables are rewritten). */
int pymdp_hip_ragged_autotune(void* plan, const float* params, const int32_t* obs,
                              float* out, int32_t* actions, void* stream, float* best_ms);
/* out8 = (stage-1 rows per work item, stage-2 rows per work item, stage-3
 *         rows per work item (fixed), 1 if autotuned, FPI grid, 0, 0, 0). */
int pymdp_hip_ragged_tile_config(void* plan, int32_t* out8);
/* Sets the rows per stage-1 and per stage-2 work item (cfg2, each 128 or
 * 256; EINVAL otherwise) and rewrites the work-item tables.  Results do not
 * depend on them.  Not while a graph that replays the plan may run. */
int pymdp_hip_ragged_set_tiles(void* plan, const int32_t* cfg2);
/* *flag_out = 1 if a team exchange (the per-sweep wait of a large agent's
 * workgroups for each other's partial marginals) of an earlier step gave up
 * waiting (the posteriors of that agent are NaN), else 0.  Synchronises the
 * device. */
int pymdp_hip_ragged_status(void* plan, int32_t* flag_out);
/* Prior of later steps' FPI: carry = 0 reads D from the params arena (the
 * default), carry = 1 reads the predicted prior the previous step wrote at
 * pred_off in the out arena, as the reference's rollout passes
 * update_empirical_prior's result to the next infer_states
 * (pymdp/envs/rollout.py:65-82).  Applies to steps launched (or graphs
 * captured) afterwards. */
int pymdp_hip_ragged_set_prior_source(void* plan, int32_t carry);
/* Optional timing trace: a device buffer of 2 * (units + stage-1 tiles +
 * stage-2 tiles + stage-3 tiles) int64 values that later steps fill with
 * wall-clock ticks (100 MHz): [2u, 2u+1] = start / end of FPI unit u, then
 * the tiles of the three EFE stages in order.  NULL switches the trace off. */
int pymdp_hip_ragged_set_trace(void* plan, int64_t* trace);
int pymdp_hip_ragged_plan_destroy(void* plan);

#ifdef __cplusplus
}
#endif

#endif /* PYMDP_HIP_H_ */
