"""NumPy restatement of the reference's JAX math for FPI and EFE.

Test infrastructure only (see oracle/__init__.py). Every function takes a
`dt` argument: np.float32 restates the reference JAX path in FP32 (the
parity target), np.float64 is the high-precision oracle that errors are
reported against.

Citations are to the reference tree (pymdp 1.0.1 plus its ffi fork).
"""
from __future__ import annotations

import itertools
from typing import Sequence

import numpy as np

# pymdp/maths.py:15  MINVAL = jnp.finfo(float).eps  (float64 eps; used as a
# weakly-typed python scalar, so in FP32 it is the float32 value nearest it).
MINVAL = 2.220446049250313e-16


def _a(x, dt):
    return np.asarray(x, dtype=dt)


def log_stable(x, dt=np.float32):
    """pymdp/maths.py:64-79  log(clip(x, MINVAL))."""
    x = _a(x, dt)
    return np.log(np.maximum(x, dt(MINVAL)))


def stable_xlogx(x, dt=np.float32):
    """pymdp/maths.py:17-30  xlogy(x, clip(x, MINVAL)); exactly 0 where x == 0."""
    x = _a(x, dt)
    out = x * np.log(np.maximum(x, dt(MINVAL)))
    return np.where(x == 0, dt(0), out).astype(dt)


def stable_entropy(x, dt=np.float32):
    """pymdp/maths.py:32-45."""
    return -stable_xlogx(x, dt).sum()


def softmax(x, axis=-1):
    """jax.nn.softmax: exp(x - max) / sum."""
    m = np.max(x, axis=axis, keepdims=True)
    e = np.exp(x - m)
    return e / e.sum(axis=axis, keepdims=True)


def factor_dot(M, xs, keep_dims=None):
    """pymdp/maths.py:81-106 (factor_dot) + :179-211 (factor_dot_flex).

    Contracts every axis of M not in keep_dims with xs, in axis order.
    """
    keep_dims = () if keep_dims is None else tuple(keep_dims)
    dims = [i for i in range(M.ndim) if i not in keep_dims]
    assert len(dims) == len(xs)
    letters = "abcdefghijklmnopqrstuvwxyz"
    m_sub = letters[: M.ndim]
    ops = [M]
    subs = [m_sub]
    for d, x in zip(dims, xs):
        ops.append(x)
        subs.append(m_sub[d])
    out = "".join(m_sub[k] for k in keep_dims)
    return np.einsum(",".join(subs) + "->" + out, *ops, optimize=True)


# ----------------------------------------------------------------------------
# Log-likelihoods  (pymdp/maths.py:214-287)
# ----------------------------------------------------------------------------

def compute_log_likelihood_single_modality(o_m, A_m, dt=np.float32, distr_obs=True):
    """pymdp/maths.py:214-235."""
    if distr_obs:
        log_A = log_stable(A_m, dt)
        return np.tensordot(_a(o_m, dt), log_A, axes=([0], [0]))
    return log_stable(_a(A_m, dt)[int(o_m)], dt)


def compute_log_likelihood(obs, A, dt=np.float32, distr_obs=True):
    """pymdp/maths.py:238-263: per-modality ll, then summed over modalities."""
    res = [compute_log_likelihood_single_modality(o, a, dt, distr_obs) for o, a in zip(obs, A)]
    return np.sum(np.stack(res), 0)


def compute_log_likelihood_per_modality(obs, A, dt=np.float32, distr_obs=True):
    """pymdp/maths.py:266-287."""
    return [compute_log_likelihood_single_modality(o, a, dt, distr_obs) for o, a in zip(obs, A)]


# ----------------------------------------------------------------------------
# Fixed-point iteration  (pymdp/algos.py:26-155)
# ----------------------------------------------------------------------------

def marginal_log_likelihood(qs, log_likelihood, i):
    """pymdp/algos.py:26-28: leave-one-factor-out contraction kept on axis i."""
    xs = [q for j, q in enumerate(qs) if j != i]
    return factor_dot(log_likelihood, xs, keep_dims=(i,))


def all_marginal_log_likelihood(qs, log_likelihoods, all_factor_lists, dt=np.float32):
    """pymdp/algos.py:30-44. Factors no modality touches receive 0."""
    num_factors = len(qs)
    qL_all = [np.zeros(1, dtype=dt)] * num_factors
    for m, factor_list_m in enumerate(all_factor_lists):
        rel = [qs[f] for f in factor_list_m]
        for l, f in enumerate(factor_list_m):
            qL_all[f] = qL_all[f] + marginal_log_likelihood(rel, log_likelihoods[m], l)
    return qL_all


def run_vanilla_fpi(A, obs, prior, num_iter=1, dt=np.float32, distr_obs=True):
    """pymdp/algos.py:53-86: single summed ll over the full joint."""
    nf = len(prior)
    ll = compute_log_likelihood(obs, A, dt, distr_obs)
    log_prior = [log_stable(p, dt) for p in prior]
    log_q = [np.zeros_like(lp) for lp in log_prior]
    for _ in range(num_iter):
        q = [softmax(l) for l in log_q]
        log_q = [marginal_log_likelihood(q, ll, f) + log_prior[f] for f in range(nf)]
    return [softmax(l) for l in log_q]


def run_factorized_fpi(A, obs, prior, A_dependencies, num_iter=1, dt=np.float32, distr_obs=True):
    """pymdp/algos.py:88-155.

    F == 1: one-pass closed form softmax(ll + ln prior) (:101-104).
    Otherwise log_q starts at 0 and num_iter Jacobi sweeps follow (:123-154):
    every factor is updated from the same q = softmax(log_q).
    """
    if len(prior) == 1:
        ll = compute_log_likelihood(obs, A, dt, distr_obs)
        return [softmax(ll + log_stable(prior[0], dt))]
    lls = compute_log_likelihood_per_modality(obs, A, dt, distr_obs)
    log_prior = [log_stable(p, dt) for p in prior]
    log_q = [np.zeros_like(lp) for lp in log_prior]
    for _ in range(num_iter):
        q = [softmax(l) for l in log_q]
        mll = all_marginal_log_likelihood(q, lls, A_dependencies, dt)
        log_q = [m + lp for m, lp in zip(mll, log_prior)]
    return [softmax(l) for l in log_q]


def fpi_from_ll(ll_list, lp_list, A_dependencies, num_iter, dt=np.float32):
    """The body the reference FFI replaces (`_jax_fpi` at pymdp/algos.py:124-133):
    FPI from per-modality log-likelihoods and log-priors."""
    ll_list = [_a(x, dt) for x in ll_list]
    lp_list = [_a(x, dt) for x in lp_list]
    log_q = [np.zeros_like(lp) for lp in lp_list]
    for _ in range(num_iter):
        q = [softmax(l) for l in log_q]
        mll = all_marginal_log_likelihood(q, ll_list, A_dependencies, dt)
        log_q = [m + lp for m, lp in zip(mll, lp_list)]
    return [softmax(l) for l in log_q]


def update_posterior_states_fpi(A, obs, prior, A_dependencies, num_iter=16, dt=np.float32,
                                method="fpi"):
    """pymdp/inference.py:171-191 + :274-295 (one-step branch, no history):
    returns qs[f] with a leading time axis of length 1."""
    n_it = 1 if method == "exact" else num_iter
    qs = run_factorized_fpi(A, obs, prior, A_dependencies, num_iter=n_it, dt=dt)
    return [q[None] for q in qs]


# ----------------------------------------------------------------------------
# Policies and EFE  (pymdp/control.py)
# ----------------------------------------------------------------------------

def construct_policies(num_states, num_controls=None, policy_len=1, control_fac_idx=None):
    """pymdp/control.py:152-195."""
    num_factors = len(num_states)
    if control_fac_idx is None:
        if num_controls is not None:
            control_fac_idx = [f for f, n_c in enumerate(num_controls) if n_c > 1]
        else:
            control_fac_idx = list(range(num_factors))
    if num_controls is None:
        num_controls = [num_states[c] if c in control_fac_idx else 1 for c in range(num_factors)]
    x = list(num_controls) * policy_len
    pols = list(itertools.product(*[list(range(i)) for i in x]))
    return np.array(pols, dtype=np.int32).reshape(len(pols), policy_len, num_factors)


def compute_expected_state(qs_prior, B, u_t, B_dependencies=None):
    """pymdp/control.py:351-389: qs'_f = B_f[..., u_f] . (x parents)."""
    if B_dependencies is None:
        B_dependencies = [[f] for f in range(len(B))]
    out = []
    for B_f, u_f, deps in zip(B, u_t, B_dependencies):
        out.append(factor_dot(B_f[..., int(u_f)], [qs_prior[d] for d in deps], keep_dims=(0,)))
    return out


def compute_expected_obs(qs, A, A_dependencies):
    """pymdp/control.py:420-447."""
    return [factor_dot(A_m, [qs[d] for d in A_dependencies[m]], keep_dims=(0,)) for m, A_m in enumerate(A)]


def compute_info_gain(qs, qo, A, A_dependencies, dt=np.float32):
    """pymdp/control.py:449-481: sum_m [H(qo_m) - H_A_m . (x qs)]."""
    total = None
    for m, (qo_m, A_m) in enumerate(zip(qo, A)):
        H_qo = stable_entropy(qo_m, dt)
        H_A_m = -stable_xlogx(A_m, dt).sum(0)
        qs_H_A_m = factor_dot(H_A_m, [qs[d] for d in A_dependencies[m]])
        v = H_qo - qs_H_A_m
        total = v if total is None else total + v
    return total


def compute_expected_utility(qo, C, t=0, dt=np.float32):
    """pymdp/control.py:483-509: C_m[t] when C_m is time-indexed."""
    util = dt(0.0)
    for o_m, C_m in zip(qo, C):
        C_m = _a(C_m, dt)
        util = util + ((o_m * (C_m[t] if C_m.ndim > 1 else C_m)).sum())
    return util


def exact_wnorm(A, dt=np.float32):
    """pymdp/maths.py:783-809 (negated eq. D.15 of Da Costa et al. 2020)."""
    from scipy.special import digamma
    A = _a(A, dt)
    safe_A = np.maximum(A, dt(MINVAL))
    safe_sum = np.maximum(safe_A.sum(axis=0), dt(MINVAL))
    wA = (np.log(safe_sum) - np.log(safe_A) + dt(1.0) / safe_A - dt(1.0) / safe_sum
          + digamma(safe_A).astype(dt) - digamma(safe_sum).astype(dt))
    return (-wA).astype(dt)


def calc_negative_pA_info_gain(pA, qo, qs, A_dependencies, dt=np.float32):
    """pymdp/control.py:511-552."""
    total = None
    for m, (pa_m, qo_m) in enumerate(zip(pA, qo)):
        pa_m = _a(pa_m, dt)
        wa_m = exact_wnorm(pa_m, dt) * (pa_m > 0)
        fd = factor_dot(wa_m, [qs[d] for d in A_dependencies[m]], keep_dims=(0,))
        v = qo_m.dot(fd)
        total = v if total is None else total + v
    return total


def calc_negative_pB_info_gain(pB, qs_t, qs_t_minus_1, B_dependencies, u_t_minus_1, dt=np.float32):
    """pymdp/control.py:554-602."""
    total = None
    for f, pb in enumerate(pB):
        pb_u = _a(pb, dt)[..., int(u_t_minus_1[f])]
        w = exact_wnorm(pb_u, dt) * (pb_u > 0)
        fd = factor_dot(w, [qs_t_minus_1[d] for d in B_dependencies[f]], keep_dims=(0,))
        v = qs_t[f].dot(fd)
        total = v if total is None else total + v
    return total


def calc_inductive_value_t(qs, qs_next, I, epsilon=1e-3, dt=np.float32):
    """pymdp/control.py:1021-1061."""
    val = dt(0.0)
    log_eps = log_stable(np.array(epsilon, dtype=dt), dt)
    for f in range(len(qs)):
        idx = int(np.argmax(qs[f]))
        m = max(int(np.argmax(I[f][:, idx])) - 1, 0)
        I_m = (dt(1.0) - _a(I[f][m, :], dt)) * log_eps
        path_available = np.clip(_a(I[f][:, idx], dt).sum(0), 0, 1)
        val = val + path_available * I_m.dot(qs_next[f])
    return val


def generate_I_matrix(H, B, threshold, depth):
    """pymdp/control.py:965-1019 (factor-local B)."""
    if depth < 1:
        raise ValueError("`depth` must be >= 1")
    out = []
    for f in range(len(H)):
        reach = (np.asarray(B[f]) > threshold).astype(np.float32).sum(axis=-1)
        reach = (reach > 0).astype(np.float32)
        pred = reach.T
        rows = [np.asarray(H[f], dtype=np.float32)]
        cur = rows[0]
        for _ in range(depth - 1):
            cur = (pred @ cur > 0).astype(np.float32)
            rows.append(cur)
        out.append(np.stack(rows, 0))
    return out


def compute_neg_efe_policy(qs_init, A, B, C, pA, pB, A_dependencies, B_dependencies, policy_i,
                           use_utility=True, use_states_info_gain=True, use_param_info_gain=False,
                           I=None, inductive_epsilon=1e-3, use_inductive=False, dt=np.float32):
    """pymdp/control.py:604-690 and the inductive variant :692-793 (one policy)."""
    qs = [_a(q, dt) for q in qs_init]
    A = [_a(a, dt) for a in A]
    B = [_a(b, dt) for b in B]
    neg_efe = dt(0.0)
    for t in range(policy_i.shape[0]):
        qs_next = compute_expected_state(qs, B, policy_i[t], B_dependencies)
        qo = compute_expected_obs(qs_next, A, A_dependencies)
        info_gain = compute_info_gain(qs_next, qo, A, A_dependencies, dt) if use_states_info_gain else dt(0)
        utility = compute_expected_utility(qo, C, t, dt) if use_utility else dt(0)
        ind = (calc_inductive_value_t(qs_init, qs_next, I, inductive_epsilon, dt)
               if use_inductive else dt(0))
        neg_param = dt(0)
        if pA is not None and use_param_info_gain:
            neg_param = neg_param + calc_negative_pA_info_gain(pA, qo, qs_next, A_dependencies, dt)
        if pB is not None and use_param_info_gain:
            neg_param = neg_param + calc_negative_pB_info_gain(pB, qs_next, qs, B_dependencies,
                                                               policy_i[t], dt)
        neg_efe = neg_efe + (info_gain + utility - neg_param + ind)
        qs = qs_next
    return dt(neg_efe)


# ---- vectorised over policies (same arithmetic, one row per policy) -----------

def _bdot(M, xs, lead="z"):
    """Per-policy factor_dot with keep_dims=(0,): M [P, S0, d1..dk], xs[i] [P, di]."""
    letters = "abcdefghijklmnopqrstuvwxy"
    sub = letters[: M.ndim - 1]
    spec = lead + sub + "," + ",".join(lead + sub[i + 1] for i in range(len(xs))) + "->" + lead + sub[0]
    return np.einsum(spec, M, *xs, optimize=True)


def _adot(A, xs, keep0=True):
    """A [O, d1..dk] (shared), xs[i] [P, di] -> [P, O] (keep0) or [P]."""
    letters = "abcdefghijklmnopqrstuvwxy"
    sub = letters[: A.ndim]
    if keep0:
        spec = sub + "," + ",".join("z" + sub[i + 1] for i in range(len(xs))) + "->z" + sub[0]
    else:
        spec = sub + "," + ",".join("z" + sub[i] for i in range(len(xs))) + "->z"
    return np.einsum(spec, A, *xs, optimize=True)


def neg_efe_all_policies(policy_matrix, qs_init, A, B, C, A_dependencies, B_dependencies,
                         pA=None, pB=None, use_utility=True, use_states_info_gain=True,
                         use_param_info_gain=False, I=None, inductive_epsilon=1e-3,
                         use_inductive=False, dt=np.float32, chunk=4096):
    """`vmap(compute_neg_efe_policy[_inductive])(policy_matrix)` of
    pymdp/control.py:324-349 / :935-961, vectorised over policy rows.

    Returns G [P] (neg_efe = -EFE)."""
    pm = np.asarray(policy_matrix)
    P, T, F = pm.shape
    A = [_a(a, dt) for a in A]
    B = [_a(b, dt) for b in B]
    C = [_a(c, dt) for c in C]
    H_A = [-stable_xlogx(a, dt).sum(0) for a in A]
    wA = None
    if use_param_info_gain and pA is not None:
        wA = [exact_wnorm(p, dt) * (_a(p, dt) > 0) for p in pA]
    out = np.zeros(P, dtype=dt)
    for p0 in range(0, P, chunk):
        pol = pm[p0:p0 + chunk]
        n = pol.shape[0]
        qs = [np.broadcast_to(_a(q, dt), (n, q.shape[-1])).copy() for q in qs_init]
        G = np.zeros(n, dtype=dt)
        for t in range(T):
            qs_next = []
            for f, (B_f, deps) in enumerate(zip(B, B_dependencies)):
                Bu = np.moveaxis(B_f[..., pol[:, t, f]], -1, 0)  # [n, S_f, deps...]
                qs_next.append(_bdot(Bu, [qs[d] for d in deps]))
            qo = [_adot(A[m], [qs_next[d] for d in A_dependencies[m]]) for m in range(len(A))]
            ig = np.zeros(n, dtype=dt)
            if use_states_info_gain:
                tot = None
                for m in range(len(A)):
                    Hq = -stable_xlogx(qo[m], dt).sum(-1)
                    v = Hq - _adot(H_A[m], [qs_next[d] for d in A_dependencies[m]], keep0=False)
                    tot = v if tot is None else tot + v
                ig = tot
            ut = np.zeros(n, dtype=dt)
            if use_utility:
                u = dt(0.0)
                for m in range(len(A)):
                    Cm = C[m][t] if C[m].ndim > 1 else C[m]
                    u = u + (qo[m] * Cm).sum(-1)
                ut = u
            neg_param = np.zeros(n, dtype=dt)
            if use_param_info_gain:
                if wA is not None:
                    for m in range(len(A)):
                        fd = _adot(wA[m], [qs_next[d] for d in A_dependencies[m]])
                        neg_param = neg_param + (qo[m] * fd).sum(-1)
                if pB is not None:
                    for f, pb in enumerate(pB):
                        pbu = np.moveaxis(_a(pb, dt)[..., pol[:, t, f]], -1, 0)  # [n, S, deps]
                        w = np.stack([exact_wnorm(x, dt) * (x > 0) for x in pbu], 0)
                        fd = _bdot(w, [qs[d] for d in B_dependencies[f]])
                        neg_param = neg_param + (qs_next[f] * fd).sum(-1)
            ind = np.zeros(n, dtype=dt)
            if use_inductive:
                log_eps = log_stable(np.array(inductive_epsilon, dtype=dt), dt)
                for f in range(F):
                    idx = int(np.argmax(qs_init[f]))
                    mm = max(int(np.argmax(I[f][:, idx])) - 1, 0)
                    I_m = (dt(1.0) - _a(I[f][mm, :], dt)) * log_eps
                    pa = np.clip(_a(I[f][:, idx], dt).sum(0), 0, 1)
                    ind = ind + pa * (qs_next[f] @ I_m)
            G = G + (ig + ut - neg_param + ind)
            qs = qs_next
        out[p0:p0 + n] = G
    return out


def update_posterior_policies(policy_matrix, qs_init, A, B, C, E, A_dependencies, B_dependencies,
                              gamma=16.0, dt=np.float32, **kw):
    """pymdp/control.py:198-349 / :795-963: (q_pi, G) with q_pi = softmax(gamma G + ln E)."""
    G = neg_efe_all_policies(policy_matrix, qs_init, A, B, C, A_dependencies, B_dependencies, dt=dt, **kw)
    q_pi = softmax(dt(gamma) * G + log_stable(E, dt))
    return q_pi.astype(dt), G


def sample_policy(policies, q_pi):
    """pymdp/control.py:115-150, deterministic mode: first-index argmax."""
    return np.asarray(policies)[int(np.argmax(q_pi)), 0]


def get_marginals(q_pi, policies, num_controls):
    """pymdp/control.py:42-67."""
    pol = np.asarray(policies)
    return [np.array([q_pi[pol[:, 0, f] == a].sum() for a in range(n)]) for f, n in enumerate(num_controls)]


def sample_action(policies, num_controls, q_pi):
    """pymdp/control.py:69-113, deterministic mode."""
    return np.array([int(np.argmax(m)) for m in get_marginals(q_pi, policies, num_controls)])


# ---------------------------------------------------------------------------
# Dirichlet-parameter learning (one agent; leading time axis on obs / beliefs)
# ---------------------------------------------------------------------------
def multidimensional_outer(arrs):
    """pymdp/maths.py:763-781: outer product by iterative expansion."""
    x = arrs[0]
    for q in arrs[1:]:
        x = np.expand_dims(x, -1) * q
    return x


def dirichlet_expected_value(dir_arr, dt=np.float32):
    """pymdp/maths.py:843-868: clip at MINVAL, normalise over axis 0."""
    d = np.maximum(_a(dir_arr, dt), dt(MINVAL))
    return (d / d.sum(axis=0, keepdims=True)).astype(dt)


def update_obs_likelihood_dirichlet_m(pA_m, obs_m, qs, dependencies_m, lr=1.0, dt=np.float32):
    """pymdp/learning.py:11-55: obs_m [T, O] (distributional), qs[f] [T, S_f];
    returns (pA_m + lr * sum_t obs_t (x) qs_deps_t, its expected value)."""
    T = obs_m.shape[0]
    dfda = sum(multidimensional_outer([_a(obs_m[t], dt)] + [_a(qs[f][t], dt) for f in dependencies_m])
               for t in range(T)) if T else np.zeros_like(_a(pA_m, dt))
    new = (_a(pA_m, dt) + dt(lr) * dfda).astype(dt)
    return new, dirichlet_expected_value(new, dt)


def update_obs_likelihood_dirichlet(pA, A, obs, qs, A_dependencies, categorical_obs, num_obs, lr, dt=np.float32):
    """pymdp/learning.py:57-117 (one agent): obs[m] is [T] indices or [T, O_m]."""
    qA, EA = [], []
    for m, pA_m in enumerate(pA):
        if pA_m is None:
            qA.append(None)
            EA.append(A[m])
            continue
        o = np.asarray(obs[m])
        o_m = o.astype(dt) if categorical_obs else np.eye(num_obs[m], dtype=dt)[o.astype(np.int64)]
        a, e = update_obs_likelihood_dirichlet_m(pA_m, o_m, qs, A_dependencies[m], lr, dt)
        qA.append(a)
        EA.append(e)
    return qA, EA


def _one_hot(idx, n, dt):
    """jax.nn.one_hot: all-zero rows for indices outside [0, n)."""
    idx = np.asarray(idx, dtype=np.int64)
    out = np.zeros(idx.shape + (n,), dt)
    ok = (idx >= 0) & (idx < n)
    out[ok, idx[ok]] = 1
    return out


def update_state_transition_dirichlet_f(pB_f, actions_f, joint_qs_f, lr=1.0, dt=np.float32):
    """pymdp/learning.py:119-157: joint_qs_f = [qs_f[t+1], qs_dep[t], ...]
    each [T1, .], actions_f [T1, U_f] one-hot."""
    T1 = actions_f.shape[0]
    dfdb = sum(multidimensional_outer([_a(x[t], dt) for x in joint_qs_f] + [_a(actions_f[t], dt)])
               for t in range(T1)) if T1 else np.zeros_like(_a(pB_f, dt))
    new = (_a(pB_f, dt) + dt(lr) * dfdb).astype(dt)
    return new, dirichlet_expected_value(new, dt)


def update_state_transition_dirichlet(pB, B, joint_beliefs, actions, num_controls, lr, factors_to_update="all",
                                      dt=np.float32):
    """pymdp/learning.py:159-234 (one agent): actions [T1, F] int."""
    F = len(pB)
    upd = list(range(F)) if factors_to_update == "all" else sorted(factors_to_update)
    qB, EB = [], []
    for f in range(F):
        if f in upd:
            b, e = update_state_transition_dirichlet_f(pB[f], _one_hot(np.asarray(actions)[:, f], num_controls[f], dt),
                                                       joint_beliefs[f], lr, dt)
        else:
            b, e = _a(pB[f], dt), dirichlet_expected_value(pB[f], dt)
        qB.append(b)
        EB.append(e)
    return qB, EB


def infer_parameters(pA, A, pB, B, beliefs_A, observations, actions, B_dependencies, A_dependencies, num_obs,
                     num_controls, learn_A, learn_B, categorical_obs=False, beliefs_B=None, lr_pA=1.0, lr_pB=1.0,
                     any_valid=None, dt=np.float32):
    """pymdp/agent.py:416-620 for inference_algo "fpi" (no smoothing), one
    agent: returns (pA, A, pB, B) after the update.  beliefs [T, S_f];
    actions [T, F] or [F].  any_valid: the batch-wide "some transition is
    valid" predicate (default: this agent's)."""
    seq = beliefs_A if beliefs_B is None else beliefs_B
    T = seq[0].shape[0]
    act = None if actions is None else np.asarray(actions)
    if act is not None and act.ndim == 1:
        act = act[None]
    act1 = act[:max(T - 1, 0)] if act is not None else None
    valid = None if act1 is None else np.all(act1 >= 0, axis=-1)
    if any_valid is None:
        any_valid = valid is not None and bool(np.any(valid))
    can_B = learn_B and act1 is not None and T > 1 and act1.shape[0] == T - 1 and any_valid
    if learn_A:
        pA, A = update_obs_likelihood_dirichlet(pA, A, observations, beliefs_A, A_dependencies, categorical_obs,
                                                num_obs, lr_pA, dt)
    if learn_B and can_B:
        mask = valid.astype(dt)
        joints = [[_a(seq[f][1:], dt) * mask[:, None]] + [_a(seq[d][:-1], dt) * mask[:, None]
                                                         for d in B_dependencies[f]] for f in range(len(pB))]
        pB, B = update_state_transition_dirichlet(pB, B, joints, act1, num_controls, lr_pB, dt=dt)
    return pA, A, pB, B


# ---------------------------------------------------------------------------
# Variational free energy of a one-step posterior (return_info diagnostics)
# ---------------------------------------------------------------------------
def stable_cross_entropy(x, y, dt=np.float32):
    """pymdp/maths.py:47-62: -sum x ln max(y, MINVAL) (0 where x == 0)."""
    x, y = _a(x, dt), _a(y, dt)
    return -np.sum(np.where(x == 0, dt(0), x * np.log(np.maximum(y, dt(MINVAL)))))


def calc_vfe_one_step(qs, prior, obs, A, A_dependencies, distr_obs=True, dt=np.float32):
    """pymdp/maths.py:448-762, single-step branch (`qs[f]` of shape [S_f]; the
    kwargs pymdp/inference.py:314-322 builds for "fpi" / "exact"): returns
    (neg_entropy, prior_cross_entropy, accuracy, vfe) with
    vfe = sum_f -H(q_f) + sum_f CE(q_f, prior_f) - sum_m E_q[ll_m]."""
    neg_entropy = sum(-stable_entropy(q, dt) for q in qs)
    cross = sum(stable_cross_entropy(q, p, dt) for q, p in zip(qs, prior))
    lls = compute_log_likelihood_per_modality(obs, A, dt, distr_obs)
    accuracy = sum(factor_dot(ll, [_a(qs[f], dt) for f in deps]) for ll, deps in zip(lls, A_dependencies))
    return neg_entropy, cross, accuracy, neg_entropy + cross - accuracy
