"""Pins of the EFE oracle (CPU only).

oracle/numpy_ref.py restates pymdp/control.py:198-793; oracle/dedup_ref.py
restates the factor-history algorithm of the reference's native layer.
Pinned by
  * an explicit-loop brute force (tests/brute.py) on tiny models,
  * dedup == per-policy on small models (the two oracles agree),
  * closed forms: identity A and B give qo = qs, info gain = H(qs),
    utility = qs . C, so G = T (H(qs) + qs . C),
  * horizon additivity: G(u0, u1) = G1(u0 | qs) + G1(u1 | B(u0) qs),
  * q_pi = softmax(gamma G + ln E) sums to 1              test/test_agent_jax.py:452-478
  * inductive value KATs 0 / ln eps / 0.25 ln eps          test/test_inductive_inference_jax.py:114-133
  * exact_wnorm closed form                                test/test_param_info_gain_jax.py:30-55
  * parameter information gain (pA / pB) against explicit loops with an
    independent digamma (known values psi(1), psi(1/2), recurrence), the
    precision ordering of test/test_param_info_gain_jax.py:60-114 and the
    large-count asymptote of the weight
  * T-maze ordering cue > arm > start                      test/test_control_jax.py:76-127
  * a planted mistake (dropped H_A term) is caught by the brute force.
"""
import numpy as np
import pytest

from oracle import dedup_ref as D
from oracle import numpy_ref as R
from tests import brute


def _model(rng, S, U, O, A_deps, B_deps, timeC=False, T=1, a_alpha=0.5, b_alpha=0.5):
    A = []
    for m, o in enumerate(O):
        a = rng.gamma(a_alpha, size=(o,) + tuple(S[d] for d in A_deps[m]))
        A.append(a / a.sum(0))
    B = []
    for f, s in enumerate(S):
        b = rng.gamma(b_alpha, size=(s,) + tuple(S[d] for d in B_deps[f]) + (U[f],))
        B.append(b / b.sum(0))
    C = [rng.normal(size=(T, o)) if (timeC and m % 2 == 0) else rng.normal(size=o) for m, o in enumerate(O)]
    qs = [rng.dirichlet(np.ones(s)) for s in S]
    return A, B, C, qs


EFE_CASES = [
    ([3], [3], [4], [[0]], [[0]], 1),
    ([3, 2], [2, 3], [3, 4], [[0, 1], [1]], [[0], [1]], 2),
    ([3, 2, 4], [2, 3, 1], [3, 4], [[0, 1, 2], [2, 0]], [[0], [1, 0], [2]], 2),
    ([2, 3], [2, 2], [2, 3, 2], [[1], [0, 1], [1, 0]], [[0, 1], [1]], 3),
]


@pytest.mark.parametrize("case", range(len(EFE_CASES)))
def test_neg_efe_matches_bruteforce(case):
    S, U, O, Ad, Bd, T = EFE_CASES[case]
    rng = np.random.default_rng(case)
    A, B, C, qs = _model(rng, S, U, O, Ad, Bd, timeC=True, T=T)
    pm = R.construct_policies(S, U, T)
    G = R.neg_efe_all_policies(pm, qs, A, B, C, Ad, Bd, dt=np.float64)
    for p in range(pm.shape[0]):
        g = brute.neg_efe_policy(qs, A, B, C, Ad, Bd, pm[p])
        assert abs(G[p] - g) < 1e-10
        g1 = R.compute_neg_efe_policy(qs, A, B, C, None, None, Ad, Bd, pm[p], dt=np.float64)
        assert abs(G[p] - g1) < 1e-10
    Gd = D.neg_efe_dedup(pm, qs, A, B, C, Ad, Bd)
    np.testing.assert_allclose(Gd, G, atol=1e-10)


@pytest.mark.parametrize("flags", [(True, False), (False, True)])
def test_term_switches_match_bruteforce(flags):
    use_u, use_ig = flags
    S, U, O, Ad, Bd, T = EFE_CASES[2]
    rng = np.random.default_rng(9)
    A, B, C, qs = _model(rng, S, U, O, Ad, Bd, T=T)
    pm = R.construct_policies(S, U, T)
    G = R.neg_efe_all_policies(pm, qs, A, B, C, Ad, Bd, use_utility=use_u, use_states_info_gain=use_ig,
                               dt=np.float64)
    Gd = D.neg_efe_dedup(pm, qs, A, B, C, Ad, Bd, use_utility=use_u, use_states_info_gain=use_ig)
    for p in range(pm.shape[0]):
        g = brute.neg_efe_policy(qs, A, B, C, Ad, Bd, pm[p], use_utility=use_u, use_states_info_gain=use_ig)
        assert abs(G[p] - g) < 1e-10
    np.testing.assert_allclose(Gd, G, atol=1e-10)


def test_identity_model_closed_form():
    rng = np.random.default_rng(1)
    S = [5]
    A = [np.eye(5)]
    B = [np.repeat(np.eye(5)[:, :, None], 3, axis=2)]
    C = [rng.normal(size=5)]
    qs = [rng.dirichlet(np.ones(5))]
    for T in (1, 2, 3):
        pm = R.construct_policies(S, [3], T)
        G = R.neg_efe_all_policies(pm, qs, A, B, C, [[0]], [[0]], dt=np.float64)
        H = -(qs[0] * np.log(qs[0])).sum()
        np.testing.assert_allclose(G, T * (H + qs[0] @ C[0]), atol=1e-12)


def test_horizon_additivity():
    rng = np.random.default_rng(2)
    S, U, O, Ad, Bd = [3, 2], [2, 3], [3, 4], [[0, 1], [1]], [[0], [1]]
    A, B, C, qs = _model(rng, S, U, O, Ad, Bd)
    pm2 = R.construct_policies(S, U, 2)
    G2 = R.neg_efe_all_policies(pm2, qs, A, B, C, Ad, Bd, dt=np.float64)
    for p in range(pm2.shape[0]):
        u0, u1 = pm2[p, 0], pm2[p, 1]
        g0 = R.compute_neg_efe_policy(qs, A, B, C, None, None, Ad, Bd, u0[None], dt=np.float64)
        qs1 = R.compute_expected_state(qs, B, u0, Bd)
        g1 = R.compute_neg_efe_policy(qs1, A, B, C, None, None, Ad, Bd, u1[None], dt=np.float64)
        assert abs(G2[p] - (g0 + g1)) < 1e-12


def test_q_pi_is_softmax_of_gamma_G_plus_lnE():
    rng = np.random.default_rng(3)
    S, U, O, Ad, Bd = [3], [3], [3], [[0]], [[0]]
    A, B, C, qs = _model(rng, S, U, O, Ad, Bd)
    pm = R.construct_policies(S, U, 2)
    E = rng.dirichlet(np.ones(pm.shape[0]))
    q_pi, G = R.update_posterior_policies(pm, qs, A, B, C, E, Ad, Bd, gamma=16.0, dt=np.float64)
    assert abs(q_pi.sum() - 1) < 1e-12
    expect = np.exp(16 * G + np.log(E))
    np.testing.assert_allclose(q_pi, expect / expect.sum(), atol=1e-12)


def _chain_I():
    return [np.array([[0.0, 0.0, 1.0], [0.0, 1.0, 1.0], [1.0, 1.0, 1.0]], dtype=np.float32)]


def test_inductive_value_kats():
    I = _chain_I()
    qs = [np.eye(3, dtype=np.float32)[0]]
    log_eps = float(np.log(np.float32(1e-3)))
    on = R.calc_inductive_value_t(qs, [np.eye(3, dtype=np.float32)[1]], I, 1e-3)
    off = R.calc_inductive_value_t(qs, [np.eye(3, dtype=np.float32)[0]], I, 1e-3)
    assert abs(float(on)) < 1e-6
    assert abs(float(off) - log_eps) < 1e-6
    v = R.calc_inductive_value_t(qs, [np.array([0.25, 0.75, 0.0], dtype=np.float32)], I, 1e-3)
    assert abs(float(v) - 0.25 * log_eps) < 1e-6


def test_exact_wnorm_closed_form():
    from scipy.special import digamma
    A = np.array([[1.0, 2.0], [3.0, 4.0]])
    s = A.sum(0)
    expect = -(np.log(s) - np.log(A) + 1 / A - 1 / s + digamma(A) - digamma(s))
    np.testing.assert_allclose(R.exact_wnorm(A, np.float64), expect, rtol=1e-6, atol=1e-8)


def test_tmaze_info_gain_ordering():
    num_states = [2, 3]
    num_obs = [3, 3, 3]
    Ad = [[0, 1], [0, 1], [1]]
    A = []
    for m, o in enumerate(num_obs):
        shape = [o] + [num_states[f] for f in Ad[m]]
        a = np.zeros(shape)
        if m == 0:
            a[:, :, 0] = 1 / 3
            a[:, :, 1] = 1 / 3
            a[:, :, 2] = np.array([[0.9, 0.1], [0.0, 0.0], [0.1, 0.9]])
        if m == 1:
            a[2, :, 0] = 1
            a[0:2, :, 1] = np.array([[0.6, 0.4], [0.6, 0.4]])
            a[2, :, 2] = 1
        if m == 2:
            a = np.eye(o)
        A.append(a)
    gains = []
    for loc in range(3):
        qs = [np.ones(2) / 2, np.eye(3)[loc]]
        qo = R.compute_expected_obs(qs, A, Ad)
        gains.append(R.compute_info_gain(qs, qo, A, Ad, np.float64))
    start, arm, cue = gains
    assert arm > start and cue > arm


def test_bruteforce_catches_dropped_HA_term():
    S, U, O, Ad, Bd, T = EFE_CASES[1]
    rng = np.random.default_rng(5)
    A, B, C, qs = _model(rng, S, U, O, Ad, Bd, T=T)
    pm = R.construct_policies(S, U, T)
    # planted mistake: info gain without the ambiguity (H_A) term
    bad = []
    for p in range(pm.shape[0]):
        q = [np.asarray(x) for x in qs]
        g = 0.0
        for t in range(T):
            q = R.compute_expected_state(q, B, pm[p, t], Bd)
            qo = R.compute_expected_obs(q, A, Ad)
            g += sum(R.stable_entropy(x, np.float64) for x in qo) + R.compute_expected_utility(qo, C, t, np.float64)
        bad.append(g)
    good = [brute.neg_efe_policy(qs, A, B, C, Ad, Bd, pm[p]) for p in range(pm.shape[0])]
    assert np.abs(np.array(bad) - np.array(good)).max() > 1e-3


# ---- parameter information gain (pymdp/control.py:511-602) ------------------

def _counts(rng, like, zero_frac=0.15):
    p = like * rng.uniform(1.0, 20.0) + rng.uniform(0.0, 2.0, size=like.shape)
    p[rng.uniform(size=like.shape) < zero_frac] = 0.0
    return p


@pytest.mark.parametrize("case", [1, 2])
def test_param_info_gain_oracle_matches_brute_force(case):
    """neg_efe_all_policies with pA / pB against explicit loops (its own
    digamma) on every policy of a small model."""
    rng = np.random.default_rng(30 + case)
    S, U, O, Ad, Bd, T = EFE_CASES[case]
    A, B, C, qs = _model(rng, S, U, O, Ad, Bd, T=T)
    pA = [_counts(rng, a) for a in A]
    pB = [_counts(rng, b) for b in B]
    pm = R.construct_policies(S, U, T)
    for kw in (dict(pA=pA), dict(pB=pB), dict(pA=pA, pB=pB)):
        G = R.neg_efe_all_policies(pm, qs, A, B, C, Ad, Bd, use_param_info_gain=True, dt=np.float64, **kw)
        for p in range(pm.shape[0]):
            ref = brute.neg_efe_policy(qs, A, B, C, Ad, Bd, pm[p], **kw)
            assert abs(G[p] - ref) <= 1e-9 * max(1.0, abs(ref)), (kw.keys(), p, G[p], ref)


def test_brute_digamma_known_values():
    """psi(1) = -Euler gamma, psi(1/2) = -gamma - 2 ln 2, psi(x + 1) = psi(x) + 1/x."""
    g = 0.5772156649015329
    assert abs(brute.digamma(1.0) + g) < 1e-12
    assert abs(brute.digamma(0.5) + g + 2 * np.log(2.0)) < 1e-12
    for x in (1e-6, 0.3, 2.5, 17.0):
        assert abs(brute.digamma(x + 1) - brute.digamma(x) - 1 / x) < 1e-9 * max(1.0, 1 / x)


@pytest.mark.parametrize("seed", [0, 1, 2, 3, 42])
def test_pA_info_gain_precision_ordering(seed):
    """Scaling Dirichlet counts up makes the (negative) information gain less
    negative (test/test_param_info_gain_jax.py:60-114)."""
    rng = np.random.default_rng(seed)
    shapes = [(5, 5), (3, 5, 2), (3, 5, 2)]
    deps = [[0], [0, 1], [0, 1]]
    qs = [rng.uniform(size=n) + 1e-16 for n in (5, 2)]
    qs = [q / q.sum() for q in qs]
    qo = [rng.uniform(size=n) + 1e-16 for n in (5, 3, 3)]
    qo = [q / q.sum() for q in qo]
    pA = [rng.uniform(size=s) * 2.0 + 1e-16 for s in shapes]
    scalar = 1.0 + rng.uniform() * 0.1
    lo = R.calc_negative_pA_info_gain(pA, qo, qs, deps, np.float64)
    hi = R.calc_negative_pA_info_gain([p * scalar for p in pA], qo, qs, deps, np.float64)
    assert hi > lo


def test_wnorm_large_count_asymptote():
    """For large counts the weight tends to -(1/(2a) - 1/(2s)) (psi(x) ~ ln x - 1/(2x))."""
    p = np.array([[1.0, 3.0], [2.0, 5.0], [4.0, 1.0]]) * 1e4
    w = R.exact_wnorm(p, np.float64)
    s = p.sum(0)
    approx = -(0.5 / p - 0.5 / s)
    np.testing.assert_allclose(w, approx, rtol=1e-3)
