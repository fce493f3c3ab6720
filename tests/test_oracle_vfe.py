"""Pins of the one-step VFE oracle (oracle/numpy_ref.calc_vfe_one_step,
pymdp/maths.py:448-762 single-step branch), CPU only:
  * one factor: the FPI posterior is exact, so its VFE equals the negative
    log evidence -ln sum_s prior(s) prod_m A_m[o_m, s] (textbook identity),
  * two factors: the mean-field VFE is an upper bound on -ln p(o) (Gibbs),
    and the three terms equal explicit loops over joint states,
  * index and one-hot observations give the same value,
  * a planted sign error in the prior cross-entropy breaks the identity.
"""
import itertools

import numpy as np

from oracle import numpy_ref as R


def _model(rng, S, O, deps):
    A = []
    for o, d in zip(O, deps):
        a = rng.gamma(0.7, size=(o,) + tuple(S[f] for f in d))
        A.append(a / a.sum(0))
    prior = [rng.dirichlet(np.ones(s)) for s in S]
    obs = [int(rng.integers(0, o)) for o in O]
    return A, prior, obs


def _log_evidence(A, prior, obs, S, deps):
    tot = 0.0
    for s in itertools.product(*[range(n) for n in S]):
        p = np.prod([prior[f][s[f]] for f in range(len(S))])
        lik = np.prod([A[m][(obs[m],) + tuple(s[f] for f in deps[m])] for m in range(len(A))])
        tot += p * lik
    return np.log(tot)


def test_single_factor_vfe_is_negative_log_evidence():
    rng = np.random.default_rng(0)
    for S, O in (([5], [4, 3]), ([9], [2]), ([3], [6, 6, 2])):
        deps = [[0]] * len(O)
        A, prior, obs = _model(rng, S, O, deps)
        ob = [np.eye(o)[x] for o, x in zip(O, obs)]
        q = R.run_factorized_fpi(A, ob, prior, deps, num_iter=16, dt=np.float64)
        vfe = R.calc_vfe_one_step(q, prior, ob, A, deps, dt=np.float64)[3]
        assert abs(vfe + _log_evidence(A, prior, obs, S, deps)) < 1e-12


def test_two_factor_terms_match_loops_and_bound():
    rng = np.random.default_rng(1)
    S, O, deps = [3, 4], [5, 2], [[0, 1], [1]]
    A, prior, obs = _model(rng, S, O, deps)
    ob = [np.eye(o)[x] for o, x in zip(O, obs)]
    q = R.run_factorized_fpi(A, ob, prior, deps, num_iter=16, dt=np.float64)
    neg, cross, acc, vfe = R.calc_vfe_one_step(q, prior, ob, A, deps, dt=np.float64)
    neg_l = sum(x * np.log(x) for qf in q for x in qf if x > 0)
    cross_l = -sum(x * np.log(p) for qf, pf in zip(q, prior) for x, p in zip(qf, pf))
    acc_l = 0.0
    for s in itertools.product(range(3), range(4)):
        w = q[0][s[0]] * q[1][s[1]]
        acc_l += w * (np.log(A[0][obs[0], s[0], s[1]]) + np.log(A[1][obs[1], s[1]]))
    assert abs(neg - neg_l) < 1e-12 and abs(cross - cross_l) < 1e-12 and abs(acc - acc_l) < 1e-12
    assert vfe >= -_log_evidence(A, prior, obs, S, deps) - 1e-12
    # index observations give the same value
    v_idx = R.calc_vfe_one_step(q, prior, obs, A, deps, distr_obs=False, dt=np.float64)[3]
    assert abs(v_idx - vfe) < 1e-12


def test_planted_sign_error_is_caught():
    rng = np.random.default_rng(2)
    S, O, deps = [6], [3], [[0]]
    A, prior, obs = _model(rng, S, O, deps)
    ob = [np.eye(3)[obs[0]]]
    q = R.run_factorized_fpi(A, ob, prior, deps, num_iter=16, dt=np.float64)
    neg, cross, acc, _ = R.calc_vfe_one_step(q, prior, ob, A, deps, dt=np.float64)
    wrong = neg - cross - acc
    assert abs(wrong + _log_evidence(A, prior, obs, S, deps)) > 1e-3
