"""Pins of the Dirichlet-learning oracle (CPU only).

oracle/numpy_ref.py restates pymdp/learning.py:11-234 and the filtering
branch of Agent.infer_parameters (pymdp/agent.py:416-620).  Pinned by
  * explicit index loops (no outer-product helper) on small tensors,
  * closed forms: one-hot observation and one-hot beliefs add lr to exactly
    one count; the added mass is lr * (number of contributing steps),
  * dirichlet_expected_value: columns sum to 1, an all-zero column becomes
    uniform (the MINVAL clip, pymdp/maths.py:843-868),
  * transitions: an action outside [0, U) adds nothing (jax.nn.one_hot), the
    agent's mask drops a step at which any factor's action is negative, a
    batch with no valid transition keeps B itself (not E[pB]),
  * a planted mistake (beliefs of t + 1 and t swapped) is caught.
"""
import itertools

import numpy as np
import pytest

from oracle import numpy_ref as R


def _loop_pA(pA, obs, qs, deps, lr):
    """sum_t obs[t, o] * prod_j qs[deps[j]][t, k_j] by explicit loops."""
    out = np.array(pA, dtype=np.float64)
    for t in range(obs.shape[0]):
        for idx in itertools.product(*[range(n) for n in out.shape]):
            w = obs[t, idx[0]]
            for j, d in enumerate(deps):
                w *= qs[d][t, idx[1 + j]]
            out[idx] += lr * w
    return out


def _loop_pB(pB, nxt, prev, deps, act_f, lr, valid):
    out = np.array(pB, dtype=np.float64)
    U = out.shape[-1]
    for t in range(nxt.shape[0]):
        u = int(act_f[t])
        if not valid[t] or u < 0 or u >= U:
            continue
        for idx in itertools.product(*[range(n) for n in out.shape[:-1]]):
            w = nxt[t, idx[0]]
            for j, d in enumerate(deps):
                w *= prev[d][t, idx[1 + j]]
            out[idx + (u,)] += lr * w
    return out


def _norm(x):
    x = np.maximum(x, R.MINVAL)
    return x / x.sum(0, keepdims=True)


@pytest.mark.parametrize("T", [1, 3])
def test_update_pA_matches_loops(T):
    rng = np.random.default_rng(0)
    S = [3, 2, 4]
    deps = [0, 2]
    O = 5
    pA = rng.gamma(2.0, size=(O, 3, 4))
    obs = rng.dirichlet(np.ones(O), size=T)
    qs = [rng.dirichlet(np.ones(s), size=T) for s in S]
    qA, EA = R.update_obs_likelihood_dirichlet_m(pA, obs, qs, deps, lr=0.7, dt=np.float64)
    ref = _loop_pA(pA, obs, qs, deps, 0.7)
    np.testing.assert_allclose(qA, ref, rtol=1e-13, atol=0)
    np.testing.assert_allclose(EA, _norm(ref), rtol=1e-13, atol=0)
    # every step adds lr of mass (obs and beliefs are distributions)
    assert abs((qA.sum() - pA.sum()) - 0.7 * T) < 1e-12


def test_update_pA_one_hot_closed_form():
    O, S = 4, [3, 5]
    pA = np.ones((O, 3, 5))
    obs = np.eye(O)[[2]]
    qs = [np.eye(3)[[1]], np.eye(5)[[4]]]
    qA, _ = R.update_obs_likelihood_dirichlet_m(pA, obs, qs, [0, 1], lr=2.5, dt=np.float64)
    d = qA - pA
    assert d[2, 1, 4] == 2.5 and np.count_nonzero(d) == 1


def test_index_and_categorical_observations_agree():
    rng = np.random.default_rng(1)
    A = [rng.dirichlet(np.ones(4), size=(3,)).T]
    pA = [rng.gamma(1.0, size=(4, 3))]
    qs = [rng.dirichlet(np.ones(3), size=2)]
    idx = np.array([1, 3])
    a1, e1 = R.update_obs_likelihood_dirichlet(pA, A, [idx], qs, [[0]], False, [4], 1.0, np.float64)
    a2, e2 = R.update_obs_likelihood_dirichlet(pA, A, [np.eye(4)[idx]], qs, [[0]], True, [4], 1.0, np.float64)
    np.testing.assert_array_equal(a1[0], a2[0])
    np.testing.assert_array_equal(e1[0], e2[0])
    # a modality without pA keeps its A
    a3, e3 = R.update_obs_likelihood_dirichlet([None], A, [idx], qs, [[0]], False, [4], 1.0, np.float64)
    assert a3[0] is None and e3[0] is A[0]


def test_dirichlet_expected_value_zero_column():
    x = np.zeros((3, 2))
    x[:, 1] = [1.0, 2.0, 1.0]
    e = R.dirichlet_expected_value(x, np.float64)
    np.testing.assert_allclose(e[:, 0], 1 / 3)
    np.testing.assert_allclose(e[:, 1], [0.25, 0.5, 0.25])


def test_update_pB_matches_loops_and_masks():
    rng = np.random.default_rng(2)
    S, U = [3, 2], [2, 3]
    B_deps = [[0, 1], [1]]
    T = 5
    qs = [rng.dirichlet(np.ones(s), size=T) for s in S]
    acts = np.array([[0, 1], [1, -1], [5, 2], [1, 0]])  # t=1 invalid (negative), t=2 factor 0 out of range
    pB = [rng.gamma(1.0, size=(3, 3, 2, 2)), rng.gamma(1.0, size=(2, 2, 3))]
    B = [_norm(p) for p in pB]
    # learning-level call: no agent mask, only one_hot's range rule
    joints = [[qs[f][1:]] + [qs[d][:-1] for d in B_deps[f]] for f in range(2)]
    qB, EB = R.update_state_transition_dirichlet(pB, B, joints, acts, U, 0.5, dt=np.float64)
    for f in range(2):
        ref = _loop_pB(pB[f], qs[f][1:], [q[:-1] for q in qs], B_deps[f], acts[:, f], 0.5, np.ones(T - 1, bool))
        np.testing.assert_allclose(qB[f], ref, rtol=1e-13, atol=0)
        np.testing.assert_allclose(EB[f], _norm(ref), rtol=1e-13)
    # factor 1 at t=1 has action -1 -> nothing; factor 0 at t=2 has 5 >= U -> nothing
    n_contrib = [np.sum((acts[:, f] >= 0) & (acts[:, f] < U[f])) for f in range(2)]
    for f in range(2):
        assert abs((qB[f].sum() - pB[f].sum()) - 0.5 * n_contrib[f]) < 1e-12
    # agent level: step t=1 is masked for every factor
    pA = [None]
    _, _, qB2, EB2 = R.infer_parameters(pA, [None], pB, B, qs, [np.zeros(T, int)], acts, B_deps, [[0]], [2], U,
                                        learn_A=False, learn_B=True, dt=np.float64)
    valid = np.all(acts >= 0, axis=-1)
    for f in range(2):
        ref = _loop_pB(pB[f], qs[f][1:], [q[:-1] for q in qs], B_deps[f], acts[:, f], 1.0, valid)
        np.testing.assert_allclose(qB2[f], ref, rtol=1e-13, atol=0)


def test_factors_to_update_and_no_valid_transition():
    rng = np.random.default_rng(3)
    S, U = [3, 2], [2, 2]
    qs = [rng.dirichlet(np.ones(s), size=3) for s in S]
    pB = [rng.gamma(1.0, size=(3, 3, 2)), rng.gamma(1.0, size=(2, 2, 2))]
    B = [rng.dirichlet(np.ones(3), size=(3, 2)).transpose(2, 0, 1), rng.dirichlet(np.ones(2), size=(2, 2)).transpose(2, 0, 1)]
    acts = np.array([[1, 0], [0, 1]])
    joints = [[qs[f][1:], qs[f][:-1]] for f in range(2)]
    qB, EB = R.update_state_transition_dirichlet(pB, B, joints, acts, U, 1.0, factors_to_update=[1], dt=np.float64)
    np.testing.assert_array_equal(qB[0], pB[0])
    np.testing.assert_allclose(EB[0], _norm(pB[0]), rtol=1e-14)
    assert not np.array_equal(qB[1], pB[1])
    # no valid transition anywhere: B stays B (not E[pB])
    bad = np.full((2, 2), -1)
    _, _, qB3, B3 = R.infer_parameters([None], [None], pB, B, qs, [np.zeros(3, int)], bad, [[0], [1]], [[0]], [2], U,
                                       learn_A=False, learn_B=True, dt=np.float64)
    assert qB3[0] is pB[0] and B3[0] is B[0]


def test_planted_swap_is_caught():
    """Swapping the beliefs of t + 1 and t (a plausible indexing mistake)
    disagrees with the loops."""
    rng = np.random.default_rng(4)
    qs = [rng.dirichlet(np.ones(3), size=3)]
    pB = [rng.gamma(1.0, size=(3, 3, 2))]
    acts = np.array([[1], [0]])
    wrong = [[qs[0][:-1], qs[0][1:]]]
    qB, _ = R.update_state_transition_dirichlet(pB, [None], wrong, acts, [2], 1.0, dt=np.float64)
    ref = _loop_pB(pB[0], qs[0][1:], [qs[0][:-1]], [0], acts[:, 0], 1.0, np.ones(2, bool))
    assert np.abs(qB[0] - ref).max() > 1e-3
