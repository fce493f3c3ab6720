"""Brute-force FP64 restatements used to pin the oracle.

Independent code paths (explicit loops over joint states, no einsum), small
inputs only.  They restate the same reference algorithms as oracle/numpy_ref.py
so that an index, sign or ordering mistake in either shows up as a mismatch.
"""
from __future__ import annotations

import itertools
import math

import numpy as np

MINVAL = 2.220446049250313e-16


def _log(x):
    return math.log(max(float(x), MINVAL))


def _softmax(v):
    m = max(v)
    e = [math.exp(x - m) for x in v]
    s = sum(e)
    return [x / s for x in e]


def _xlogx(x):
    x = float(x)
    return 0.0 if x == 0.0 else x * math.log(max(x, MINVAL))


def fpi(A, obs, prior, deps, num_iter):
    """pymdp/algos.py:88-155 with explicit joint-state loops (one agent).
    obs[m] is a distribution over O_m."""
    F = len(prior)
    S = [len(p) for p in prior]
    ll = []
    for m, A_m in enumerate(A):
        A_m = np.asarray(A_m, dtype=np.float64)
        d = {}
        for s in itertools.product(*[range(S[f]) for f in deps[m]]):
            d[s] = sum(float(obs[m][o]) * _log(A_m[(o,) + s]) for o in range(A_m.shape[0]))
        ll.append(d)
    lp = [[_log(x) for x in p] for p in prior]
    if F == 1:
        tot = [0.0] * S[0]
        for m in range(len(A)):
            for s, v in ll[m].items():
                tot[s[0]] += v
        return [_softmax([t + l for t, l in zip(tot, lp[0])])]
    log_q = [[0.0] * n for n in S]
    for _ in range(num_iter):
        q = [_softmax(l) for l in log_q]
        acc = [[0.0] * n for n in S]
        for m in range(len(A)):
            for s, v in ll[m].items():
                for i, f in enumerate(deps[m]):
                    prod = 1.0
                    for j, g in enumerate(deps[m]):
                        if j != i:
                            prod *= q[g][s[j]]
                    acc[f][s[i]] += v * prod
        log_q = [[a + l for a, l in zip(acc[f], lp[f])] for f in range(F)]
    return [_softmax(l) for l in log_q]


def digamma(x):
    """psi(x), x > 0: upward recurrence to x >= 10, then the asymptotic series
    (Abramowitz & Stegun 6.3.5, 6.3.18)."""
    x = float(x)
    r = 0.0
    while x < 10.0:
        r -= 1.0 / x
        x += 1.0
    f = 1.0 / (x * x)
    return r + math.log(x) - 0.5 / x - f * (1 / 12 - f * (1 / 120 - f * (1 / 252 - f * (1 / 240 - f / 132))))


def wnorm(p):
    """pymdp/maths.py:783-809 times the (p > 0) mask of pymdp/control.py:546,
    element by element; the sum runs over the leading axis."""
    p = np.asarray(p, dtype=np.float64)
    out = np.zeros_like(p)
    for idx in itertools.product(*[range(n) for n in p.shape[1:]]):
        col = [max(float(p[(o,) + idx]), MINVAL) for o in range(p.shape[0])]
        s = max(sum(col), MINVAL)
        for o in range(p.shape[0]):
            if p[(o,) + idx] > 0:
                a = col[o]
                out[(o,) + idx] = -(math.log(s) - math.log(a) + 1 / a - 1 / s + digamma(a) - digamma(s))
    return out


def neg_efe_policy(qs_init, A, B, C, A_deps, B_deps, policy, use_utility=True, use_states_info_gain=True,
                   pA=None, pB=None):
    """pymdp/control.py:604-690 for one policy, explicit loops (with pA / pB:
    the parameter information gain of :511-602 is subtracted)."""
    F = len(qs_init)
    qs = [list(map(float, q)) for q in qs_init]
    G = 0.0
    for t in range(len(policy)):
        nxt = []
        for f in range(F):
            Bf = np.asarray(B[f], dtype=np.float64)
            u = int(policy[t][f])
            out = [0.0] * Bf.shape[0]
            for s in range(Bf.shape[0]):
                for par in itertools.product(*[range(len(qs[d])) for d in B_deps[f]]):
                    w = 1.0
                    for i, d in enumerate(B_deps[f]):
                        w *= qs[d][par[i]]
                    out[s] += Bf[(s,) + par + (u,)] * w
            nxt.append(out)
        ig = 0.0
        ut = 0.0
        for m, A_m in enumerate(A):
            A_m = np.asarray(A_m, dtype=np.float64)
            O = A_m.shape[0]
            qo = [0.0] * O
            ha = 0.0
            for s in itertools.product(*[range(len(nxt[d])) for d in A_deps[m]]):
                w = 1.0
                for i, d in enumerate(A_deps[m]):
                    w *= nxt[d][s[i]]
                h = 0.0
                for o in range(O):
                    qo[o] += A_m[(o,) + s] * w
                    h -= _xlogx(A_m[(o,) + s])
                ha += h * w
            Hq = -sum(_xlogx(x) for x in qo)
            ig += Hq - ha
            Cm = np.asarray(C[m], dtype=np.float64)
            Cv = Cm[t] if Cm.ndim > 1 else Cm
            ut += sum(qo[o] * Cv[o] for o in range(O))
        neg_param = 0.0
        if pA is not None:
            for m, pa in enumerate(pA):
                W = wnorm(pa)
                A_m = np.asarray(A[m], dtype=np.float64)
                for o in range(A_m.shape[0]):
                    qo_o = 0.0
                    wq_o = 0.0
                    for s in itertools.product(*[range(len(nxt[d])) for d in A_deps[m]]):
                        w = 1.0
                        for i, d in enumerate(A_deps[m]):
                            w *= nxt[d][s[i]]
                        qo_o += A_m[(o,) + s] * w
                        wq_o += W[(o,) + s] * w
                    neg_param += qo_o * wq_o
        if pB is not None:
            for f, pb in enumerate(pB):
                pb = np.asarray(pb, dtype=np.float64)
                u = int(policy[t][f])
                W = wnorm(pb[..., u])
                for s in range(pb.shape[0]):
                    acc = 0.0
                    for par in itertools.product(*[range(len(qs[d])) for d in B_deps[f]]):
                        w = 1.0
                        for i, d in enumerate(B_deps[f]):
                            w *= qs[d][par[i]]
                        acc += W[(s,) + par] * w
                    neg_param += nxt[f][s] * acc
        G += (ig if use_states_info_gain else 0.0) + (ut if use_utility else 0.0) - neg_param
        qs = nxt
    return G
