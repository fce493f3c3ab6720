"""Pins of the sequence-inference oracle (oracle/seq_ref.py: marginal and
variational message passing, pymdp/algos.py:157-500 with the masks and
action conditioning of pymdp/inference.py:56-271), CPU only:
  * a one-step window (T = 1) is the FPI: MMP and VMP both equal the FPI
    oracle (pymdp/algos.py:88-155) for the same number of iterations,
  * MMP with identity transitions and an uninformative likelihood has every
    q_t = prior as its fixed point (the half-weighted forward and backward
    messages balance); a planted full weight on both moves it,
  * cuts: a window with valid_steps = k is, on its last k steps, the
    window of those k steps alone; with validity masks on, a negative action
    splits the window into two independent chains; without them it selects
    the neutral (uniform) transition,
  * posteriors are normalised.
"""
import numpy as np
import pytest

from oracle import numpy_ref as R
from oracle import seq_ref as Q


def _model(rng, S, U, O, deps, T, a_alpha=0.7):
    A = []
    for o, d in zip(O, deps):
        a = rng.gamma(a_alpha, size=(o,) + tuple(S[f] for f in d))
        A.append(a / a.sum(0))
    B = []
    for s, u in zip(S, U):
        b = rng.gamma(1.0, size=(s, s, u))
        B.append(b / b.sum(0))
    prior = [rng.dirichlet(np.ones(s)) for s in S]
    obs = [rng.integers(0, o, size=T) for o in O]
    acts = np.stack([rng.integers(0, u, size=T - 1) for u in U], -1) if T > 1 else np.zeros((0, len(S)), int)
    return A, B, prior, obs, acts


@pytest.mark.parametrize("method", ["mmp", "vmp"])
def test_one_step_window_is_fpi(method):
    rng = np.random.default_rng(0)
    S, U, O, deps = [4, 3, 5], [2, 2, 1], [6, 4], [[0, 1], [1, 2]]
    A, B, prior, obs, _ = _model(rng, S, U, O, deps, 1)
    ob = [np.eye(o)[x[0]] for o, x in zip(O, obs)]
    for n in (1, 3, 16):
        q = Q.run_sequence(method, A, B, [x[:1] for x in obs], prior, deps, None, num_iter=n, distr_obs=False,
                           dt=np.float64)
        f = R.run_factorized_fpi(A, ob, prior, deps, num_iter=n, dt=np.float64)
        for a, b in zip(q, f):
            np.testing.assert_allclose(a[0], b, atol=1e-12)


def test_mmp_identity_transitions_fixed_point_is_prior():
    rng = np.random.default_rng(1)
    S, T = [5], 4
    A = [np.full((3, 5), 1.0 / 3)]
    B = [np.eye(5)[:, :, None]]
    prior = [rng.dirichlet(np.ones(5))]
    obs = [np.zeros(T, int)]
    q = Q.run_sequence("mmp", A, B, obs, prior, [[0]], None, num_iter=500, distr_obs=False, dt=np.float64)
    np.testing.assert_allclose(q[0], np.tile(prior[0], (T, 1)), atol=1e-9)


def test_planted_full_weight_moves_the_fixed_point(monkeypatch):
    rng = np.random.default_rng(1)
    T = 4
    A = [np.full((3, 5), 1.0 / 3)]
    B = [np.eye(5)[:, :, None]]
    prior = [rng.dirichlet(np.ones(5))]
    obs = [np.zeros(T, int)]
    orig = Q._messages

    def planted(method, B_, qs, ln_prior, trans_valid, dt):
        fwd, bwd = orig(method, B_, qs, ln_prior, trans_valid, dt)
        return fwd, [2 * b for b in bwd]  # backward at full weight

    monkeypatch.setattr(Q, "_messages", planted)
    q = Q.run_sequence("mmp", A, B, obs, prior, [[0]], None, num_iter=200, distr_obs=False, dt=np.float64)
    assert np.abs(q[0] - prior[0]).max() > 1e-3


@pytest.mark.parametrize("method", ["mmp", "vmp"])
def test_valid_steps_equals_truncated_window(method):
    rng = np.random.default_rng(2)
    S, U, O, deps, T = [4, 3], [3, 2], [5, 3], [[0, 1], [1]], 6
    A, B, prior, obs, acts = _model(rng, S, U, O, deps, T)
    k = 3
    q = Q.run_sequence(method, A, B, obs, prior, deps, acts, num_iter=8, distr_obs=False, valid_steps=k,
                       dt=np.float64)
    q2 = Q.run_sequence(method, A, B, [x[-k:] for x in obs], prior, deps, acts[-(k - 1):], num_iter=8,
                        distr_obs=False, dt=np.float64)
    for a, b in zip(q, q2):
        np.testing.assert_allclose(a[T - k:], b, atol=1e-12)
        np.testing.assert_allclose(a.sum(-1), 1.0, atol=1e-12)


@pytest.mark.parametrize("method", ["mmp", "vmp"])
def test_negative_action_splits_the_window(method):
    rng = np.random.default_rng(3)
    S, U, O, deps, T = [4, 3], [3, 2], [5, 3], [[0, 1], [1]], 6
    A, B, prior, obs, acts = _model(rng, S, U, O, deps, T)
    acts = acts.copy()
    cut = 2  # transition 2 -> 3 invalid
    acts[cut, 1] = -1
    q = Q.run_sequence(method, A, B, obs, prior, deps, acts, num_iter=8, distr_obs=False, valid_steps=T,
                       dt=np.float64)
    qa = Q.run_sequence(method, A, B, [x[:cut + 1] for x in obs], prior, deps, acts[:cut], num_iter=8,
                        distr_obs=False, dt=np.float64)
    qb = Q.run_sequence(method, A, B, [x[cut + 1:] for x in obs], prior, deps, acts[cut + 1:], num_iter=8,
                        distr_obs=False, dt=np.float64)
    for a, b, c in zip(q, qa, qb):
        np.testing.assert_allclose(a[:cut + 1], b, atol=1e-12)
        np.testing.assert_allclose(a[cut + 1:], c, atol=1e-12)
    # without masks the negative action selects the uniform transition of factor 1
    q0 = Q.run_sequence(method, A, B, obs, prior, deps, acts, num_iter=8, distr_obs=False, dt=np.float64)
    B2 = [B[0], np.concatenate([B[1], np.full((3, 3, 1), 1.0 / 3)], -1)]
    acts2 = acts.copy()
    acts2[cut, 1] = 2
    q2 = Q.run_sequence(method, A, B2, obs, prior, deps, acts2, num_iter=8, distr_obs=False, dt=np.float64)
    for a, b in zip(q0, q2):
        np.testing.assert_allclose(a, b, atol=1e-12)
