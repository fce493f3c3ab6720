"""FP64 NumPy restatement of the trajectory loss of
pymdp_amd.differentiable.trajectory_loss for one agent, and central finite
differences of it (the gradient oracle).  Test infrastructure only.

Per step t (pymdp/agent.py:675-906, :825-851 restated): qs_t = FPI of the
observation with prior_t (pymdp/algos.py:88-155), G_t = neg-EFE of every
policy (pymdp/control.py:604-690), loss += log_softmax(gamma G_t + ln E)
[a_t] (pymdp/control.py:963 with E uniform), prior_{t+1}[f] =
B_f[:, :, u_f(a_t)] qs_t[f] (pymdp/control.py:351-389).  A, B, D = softmax
of their logits over axis 0, C = log_softmax.
"""
from __future__ import annotations

import numpy as np

from . import numpy_ref as R


def _softmax(x, axis):
    e = np.exp(x - x.max(axis=axis, keepdims=True))
    return e / e.sum(axis=axis, keepdims=True)


def trajectory_loss(A_logits, B_logits, C_logits, D_logits, obs, actions, policies, A_deps, B_deps,
                    num_iter=16, gamma=1.0):
    """One agent: A_logits[m] [O, S...], B_logits[f] [S, S, U], C_logits[m]
    [O], D_logits[f] [S]; obs int [T, M]; actions int [T]."""
    dt = np.float64
    A = [_softmax(np.asarray(x, dt), 0) for x in A_logits]
    B = [_softmax(np.asarray(x, dt), 0) for x in B_logits]
    C = [np.asarray(x, dt) - np.log(np.exp(np.asarray(x, dt) - np.max(x)).sum()) - np.max(x) for x in C_logits]
    prior = [_softmax(np.asarray(x, dt), 0) for x in D_logits]
    pm = np.asarray(policies)
    P = pm.shape[0]
    loss = 0.0
    for t in range(len(actions)):
        ob = [np.eye(a.shape[0], dtype=dt)[int(o)] for a, o in zip(A, obs[t])]
        qs = R.run_factorized_fpi(A, ob, prior, A_deps, num_iter=num_iter, dt=dt)
        _, G = R.update_posterior_policies(pm, qs, A, B, C, np.full(P, 1.0 / P), A_deps, B_deps, gamma=gamma,
                                           dt=dt)
        logits = gamma * G + np.log(1.0 / P)
        mx = logits.max()
        loss += logits[int(actions[t])] - mx - np.log(np.exp(logits - mx).sum())
        u = pm[int(actions[t]), 0]
        prior = [B[f][:, :, u[f]] @ qs[f] for f in range(len(qs))]
    return loss


def central_difference(fn, params, coords, h=1e-6):
    """d fn / d params[i][idx] for (i, idx) in coords, (f(x+h) - f(x-h)) / 2h."""
    out = []
    for i, idx in coords:
        p = [np.array(x, dtype=np.float64, copy=True) for x in params]
        p[i][idx] += h
        fp = fn(p)
        p[i][idx] -= 2 * h
        fm = fn(p)
        out.append((fp - fm) / (2 * h))
    return np.array(out)
