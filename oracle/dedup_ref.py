"""NumPy restatement of the factor-history (action-history dedup) neg-EFE.

Test infrastructure only.  Follows the reference's native CPU algorithm:
per-(t, f) history dedup (pymdp/ffi/neg_efe/factor_history_tree.h:57-126),
modality scoring per distinct tuple of dependency histories
(pymdp/ffi/neg_efe/factor_history_tables.h:187-242,
pymdp/ffi/neg_efe/neg_efe_cpu_core.h:327-373) and the per-policy scatter
G[p] = sum_t sum_m score[t, m, tuple(p)] (pymdp/ffi/neg_efe/neg_efe_cpu_pipeline.cc:24-62),
but with the reference's JAX numerics (MINVAL floor, no f16 packing).

Used as the oracle where the per-policy vectorised form (numpy_ref) is too
slow: 32768 policies x 4 modalities x 4096-state contractions.  Pinned
against numpy_ref.neg_efe_all_policies on small cases in tests.
"""
from __future__ import annotations

import numpy as np

from .numpy_ref import log_stable, softmax, stable_xlogx


def factor_histories(policy_matrix, B_dependencies):
    """Returns levels[t][f] = (parents [n, rank], action [n], p2h [P])."""
    pm = np.asarray(policy_matrix)
    P, T, F = pm.shape
    levels = []
    for t in range(T):
        row = []
        for f in range(F):
            deps = B_dependencies[f]
            if t == 0:
                par = np.zeros((P, len(deps)), dtype=np.int64)
            else:
                par = np.stack([levels[t - 1][d][2] for d in deps], 1)
            key = np.concatenate([par, pm[:, t, f:f + 1].astype(np.int64)], 1)
            uniq, inv = np.unique(key, axis=0, return_inverse=True)
            row.append((uniq[:, :-1], uniq[:, -1], inv.reshape(-1)))
        levels.append(row)
    return levels


def _staged_contract(X, tup, cur, deps):
    """Contract X [O, S_d0, ..., S_d(R-1)] with cur[d_i][h_i] for every
    distinct history tuple (rows of `tup`), last dependency first, sharing
    the partial contraction between tuples with a common suffix."""
    R = len(deps)
    Xs = X[None]  # [1, O, S...]
    prev_keys = np.zeros((1, 0), dtype=np.int64)
    for i in reversed(range(R)):
        keys, inv = np.unique(tup[:, i:], axis=0, return_inverse=True)
        # parent row: suffix keys[:, 1:] located among prev_keys
        if i == R - 1:
            par = np.zeros(len(keys), dtype=np.int64)
        else:
            lookup = {tuple(k): j for j, k in enumerate(prev_keys)}
            par = np.array([lookup[tuple(k[1:])] for k in keys], dtype=np.int64)
        q = cur[deps[i]][keys[:, 0]]  # [n_i, S_di]
        Xs = np.einsum("n...k,nk->n...", Xs[par], q)
        prev_keys = keys
        last_inv = inv.reshape(-1)
    return Xs[last_inv]


def neg_efe_dedup(policy_matrix, qs_init, A, B, C, A_dependencies, B_dependencies, use_utility=True,
                  use_states_info_gain=True, dt=np.float64):
    """G[p] for one agent via history dedup.  C[m] is (O,) or (T, O)."""
    pm = np.asarray(policy_matrix)
    P, T, F = pm.shape
    A = [np.asarray(a, dtype=dt) for a in A]
    B = [np.asarray(b, dtype=dt) for b in B]
    C = [np.asarray(c, dtype=dt) for c in C]
    levels = factor_histories(pm, B_dependencies)
    prev = [np.asarray(q, dtype=dt)[None] for q in qs_init]  # [H, S]
    HA = [-stable_xlogx(a, dt).sum(0) for a in A]
    G = np.zeros(P, dtype=dt)
    for t in range(T):
        cur = []
        for f in range(F):
            par, act, _ = levels[t][f]
            deps = B_dependencies[f]
            Bu = np.moveaxis(B[f][..., act], -1, 0)  # [n, S_f, deps...]
            x = Bu
            for i in reversed(range(len(deps))):
                q = prev[deps[i]][par[:, i]]  # [n, S_d]
                x = np.einsum("n...k,nk->n...", x, q)
            cur.append(x)
        for m, A_m in enumerate(A):
            deps = A_dependencies[m]
            idx = np.stack([levels[t][d][2] for d in deps], 1)
            tup, inv = np.unique(idx, axis=0, return_inverse=True)
            qo = _staged_contract(A_m, tup, cur, deps)
            hq = _staged_contract(HA[m][None], tup, cur, deps)[:, 0]
            score = np.zeros(len(tup), dtype=dt)
            if use_states_info_gain:
                score = score + (-stable_xlogx(qo, dt).sum(-1) - hq)
            if use_utility:
                Cm = C[m][t] if C[m].ndim > 1 else C[m]
                score = score + qo @ Cm
            G = G + score[inv.reshape(-1)]
        prev = cur
    return G


def q_pi_from_G(G, gamma, E=None, dt=np.float64):
    G = np.asarray(G, dtype=dt)
    lnE = log_stable(np.full(G.shape, 1.0 / G.shape[-1]) if E is None else E, dt)
    return softmax(dt(gamma) * G + lnE)
