"""NumPy restatement of the reference's sequence inference (marginal and
variational message passing over a window of T time steps), for one agent.

Test infrastructure only (see oracle/__init__.py).  Follows
  pymdp/inference.py:56-73    _truncate_for_horizon
  pymdp/inference.py:100-126  _build_sequence_validity_masks
  pymdp/inference.py:128-171  _condition_transitions_on_actions ("neutral")
  pymdp/inference.py:194-271  _run_sequence_inference
  pymdp/algos.py:157-239      mirror_gradient_descent_step, update_marginals
  pymdp/algos.py:311-399      get_vmp_messages
  pymdp/algos.py:435-500      get_mmp_messages
restricted to factor-local transitions (B_dependencies[f] == [f], the
reference default, pymdp/utils.py:69-76).
"""
from __future__ import annotations

import numpy as np

from .numpy_ref import MINVAL, _a, all_marginal_log_likelihood, compute_log_likelihood_per_modality, log_stable, softmax


def validity_masks(T, past_actions, valid_steps):
    """pymdp/inference.py:100-126 when valid_steps is given; without it the
    reference builds no masks (pymdp/inference.py:225-228) and every
    transition counts (pymdp/algos.py:329-331, :455-457), a negative action
    then selecting the neutral transition of condition_transitions."""
    if valid_steps is None:
        return np.ones(T, bool), np.ones(max(T - 1, 0), bool)
    k = int(np.clip(valid_steps, 1, T))
    obs_valid = np.arange(T) >= (T - k)
    if T <= 1:
        return obs_valid, np.zeros(0, bool)
    trans = obs_valid[:-1] & obs_valid[1:]
    if past_actions is not None:
        trans = trans & np.all(np.asarray(past_actions) >= 0, axis=-1)
    return obs_valid, trans


def condition_transitions(B, past_actions, dt=np.float32):
    """pymdp/inference.py:128-171, invalid_action_mode "neutral": B_f[..., u_t]
    per step, a uniform column (1 / S_f) where the action is negative."""
    out = []
    for f, b in enumerate(B):
        b = _a(b, dt)
        sel = []
        for u in np.asarray(past_actions)[:, f]:
            if u < 0:
                sel.append(np.full(b.shape[:-1], dt(1.0) / dt(b.shape[0]), dtype=dt))
            else:
                sel.append(b[..., int(u)])
        out.append(np.stack(sel, 0) if sel else np.zeros((0,) + b.shape[:-1], dt))
    return out


def _messages(method, B, qs, ln_prior, trans_valid, dt):
    """Forward / backward messages per factor, [T, S_f] each (factor-local B)."""
    T = qs[0].shape[0]
    start = np.concatenate([[True], ~trans_valid])
    outgoing = np.concatenate([trans_valid, [False]])
    fwd, bwd = [], []
    for f, b in enumerate(B):
        S = qs[f].shape[1]
        fm = np.zeros((T, S), dt)
        bm = np.zeros((T, S), dt)
        for t in range(1, T):
            if not trans_valid[t - 1]:
                continue
            Bt = b[t - 1]  # [S_f (next), S_f (previous)]
            if method == "mmp":
                fm[t] = log_stable(Bt @ qs[f][t - 1], dt)
            else:
                fm[t] = log_stable(Bt, dt) @ qs[f][t - 1]
        for t in range(T - 1):
            if not trans_valid[t]:
                continue
            Bt = b[t]
            if method == "mmp":
                bn = Bt / np.maximum(Bt.sum(-1, keepdims=True), dt(MINVAL))
                bm[t] = log_stable(qs[f][t + 1] @ bn, dt) * dt(0.5)
            else:
                bm[t] = qs[f][t + 1] @ log_stable(Bt, dt)
        for t in range(T):
            if start[t]:
                fm[t] = ln_prior[f]
        if method == "mmp":
            fm = fm * np.where(outgoing[:, None], dt(0.5), dt(1.0))
        fwd.append(fm.astype(dt))
        bwd.append(bm.astype(dt))
    return fwd, bwd


def run_sequence(method, A, B, obs, prior, A_dependencies, past_actions=None, num_iter=16, tau=1.0,
                 distr_obs=True, valid_steps=None, dt=np.float32):
    """Posterior q[f] [T, S_f] of the window: obs[m] is [T] indices or [T, O_m];
    B[f] [S_f, S_f, U_f]; past_actions [T-1, F] or None (then every B has one
    control)."""
    T = len(obs[0])
    if past_actions is None and T > 1:
        if any(b.shape[-1] != 1 for b in B):
            raise ValueError("past_actions required when B has more than one control")
        past_actions = np.zeros((T - 1, len(B)), np.int64)
    if past_actions is None:
        past_actions = np.zeros((0, len(B)), np.int64)
    obs_valid, trans_valid = validity_masks(T, past_actions if T > 1 else None, valid_steps)
    Bc = condition_transitions(B, past_actions, dt)
    lls = [compute_log_likelihood_per_modality([o[t] for o in obs], A, dt, distr_obs) for t in range(T)]
    lls = [[(ll * dt(obs_valid[t])).astype(dt) for ll in lls[t]] for t in range(T)]
    ln_prior = [log_stable(p, dt) for p in prior]
    qs = [np.full((T, len(p)), dt(1.0) / dt(len(p)), dt) for p in prior]
    for _ in range(num_iter):
        ln_qs = [log_stable(q, dt) for q in qs]
        fwd, bwd = _messages(method, Bc, qs, ln_prior, trans_valid, dt)
        ln_A = [np.zeros_like(q) for q in qs]
        for t in range(T):
            m = all_marginal_log_likelihood([q[t] for q in qs], lls[t], A_dependencies, dt)
            for f in range(len(qs)):
                ln_A[f][t] = ln_A[f][t] + m[f]
        new = []
        for f in range(len(qs)):
            err = ln_A[f] - ln_qs[f] + bwd[f] + fwd[f]
            l = ln_qs[f] + dt(tau) * err
            new.append(softmax(l - l.mean(-1, keepdims=True)).astype(dt))
        qs = new
    return qs


def truncate_for_horizon(obs, past_actions, inference_horizon):
    """pymdp/inference.py:56-73."""
    if inference_horizon is None:
        return obs, past_actions
    obs = [np.asarray(o)[-inference_horizon:] for o in obs]
    if past_actions is None:
        return obs, None
    h = max(inference_horizon - 1, 0)
    return obs, (np.asarray(past_actions)[:0] if h == 0 else np.asarray(past_actions)[-h:])
