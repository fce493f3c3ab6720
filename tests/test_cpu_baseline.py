"""The C/OpenMP CPU baseline (oracle/cpu_baseline.c) agrees with the NumPy
oracle, so the timed CPU number is the same computation."""
import numpy as np
import pytest

from oracle import cpu_ref
from oracle import dedup_ref as DR
from oracle import numpy_ref as R

CASES = [
    ([4, 3], [5, 6], [[0, 1], [0, 1]]),
    ([3, 2, 4], [3, 4, 2], [[0, 2], [1], [2, 1, 0]]),
    ([3, 2, 2, 2], [3, 2], [[0, 1, 2, 3], [3, 0]]),
    ([5], [3, 4], [[0], [0]]),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_cpu_baseline_matches_oracle(case):
    S, O, deps = CASES[case]
    rng = np.random.default_rng(case)
    batch = 6
    A = []
    for m, o in enumerate(O):
        a = rng.gamma(0.5, size=(batch, o) + tuple(S[d] for d in deps[m])).astype(np.float32)
        A.append(a / a.sum(1, keepdims=True))
    prior = [rng.dirichlet(np.ones(s), size=batch).astype(np.float32) for s in S]
    obs = np.stack([rng.integers(0, o, size=batch) for o in O], 1)
    q = cpu_ref.infer_states(A, obs, prior, deps, num_iter=16, nthreads=2)
    off = np.concatenate([[0], np.cumsum(S)])
    for b in range(batch):
        ob = [np.eye(o, dtype=np.float32)[obs[b, m]] for m, o in enumerate(O)]
        q64 = R.run_factorized_fpi([a[b] for a in A], ob, [p[b] for p in prior], deps, 16, np.float64)
        for f in range(len(S)):
            assert np.abs(q[b, off[f]:off[f + 1]] - q64[f]).max() < 2e-5


def _agents_case(seed, S, U, O, T, A_deps, B_deps, n=3):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        A = []
        for o, dl in zip(O, A_deps):
            a = rng.gamma(0.5, size=(o,) + tuple(S[d] for d in dl)).astype(np.float32)
            A.append(a / a.sum(0, keepdims=True))
        B = []
        for f, dl in enumerate(B_deps):
            b = rng.gamma(0.3, size=(S[f],) + tuple(S[d] for d in dl) + (U[f],)).astype(np.float32)
            B.append(b / b.sum(0, keepdims=True))
        C = [rng.normal(size=o).astype(np.float32) for o in O]
        D = [rng.dirichlet(np.ones(s)).astype(np.float32) for s in S]
        qs = [rng.dirichlet(np.ones(s)).astype(np.float32) for s in S]
        out.append(dict(S=S, U=U, O=O, A=A, B=B, C=C, D=D, qs=qs, obs=[int(rng.integers(o)) for o in O],
                        policies=R.construct_policies(S, U, T), A_dependencies=A_deps, B_dependencies=B_deps,
                        gamma=16.0))
    return out


AGENT_CASES = [
    # rank-3 modalities, rank-1 transitions, horizon 2 (the plain-EFE shape, reduced)
    (([4, 3, 2], [3, 2, 1], [3, 4], 2, [[0, 1, 2], [0, 1, 2]], [[0], [1], [2]]), True),
    # mixed ranks, a rank-2 transition, horizon 3, posterior given (no FPI)
    (([3, 2, 4], [2, 2, 2], [3, 5, 2], 3, [[0, 2], [1], [2, 1, 0]], [[0], [1, 0], [2]]), False),
    # one factor (FPI is one sweep)
    (([5], [3], [4, 2], 2, [[0], [0]], [[0]]), True),
]


@pytest.mark.parametrize("case", range(len(AGENT_CASES)))
def test_cpu_agents_match_oracle(case):
    """cpu_ref_agents (FPI + factor-history neg-EFE + q_pi + argmax, FP32)
    against the FP64 oracle: the per-policy neg-EFE of numpy_ref (the
    restatement of pymdp/control.py) and the FPI of run_factorized_fpi."""
    (S, U, O, T, A_deps, B_deps), do_fpi = AGENT_CASES[case]
    ags = _agents_case(case, S, U, O, T, A_deps, B_deps)
    out = cpu_ref.Agents(ags, 16, do_fpi).run(2)
    for ag, o in zip(ags, out):
        if do_fpi:
            ob = [np.eye(n, dtype=np.float32)[x] for n, x in zip(O, ag["obs"])]
            q64 = R.run_factorized_fpi(ag["A"], ob, ag["D"], A_deps, 16, np.float64)
            assert np.abs(np.concatenate(q64) - o["qs"]).max() < 2e-5
        qs = np.split(o["qs"].astype(np.float64), np.cumsum(S)[:-1])
        pm = ag["policies"]
        q64, G64 = R.update_posterior_policies(pm, qs, ag["A"], ag["B"], ag["C"], np.full(pm.shape[0], 1.0 / pm.shape[0]),
                                               A_deps, B_deps, gamma=16.0, dt=np.float64)
        assert np.abs(o["G"] - G64).max() < 2e-5
        assert np.abs(o["q_pi"] - q64).max() < 1e-4
        assert q64[o["best"][0]] >= q64.max() - 1e-6


def test_cpu_agents_planted_mistake_detected():
    """The comparison is a real check: the oracle with a planted mistake (B
    transposed in its two state axes, or the utility term dropped) differs
    from the C restatement by far more than the tolerance."""
    S, U, O, T, A_deps, B_deps = AGENT_CASES[0][0]
    ag = _agents_case(7, S, U, O, T, A_deps, B_deps, n=1)[0]
    o = cpu_ref.Agents([ag], 16, True).run(1)[0]
    qs = np.split(o["qs"].astype(np.float64), np.cumsum(S)[:-1])
    pm = ag["policies"]
    G_ok = DR.neg_efe_dedup(pm, qs, ag["A"], ag["B"], ag["C"], A_deps, B_deps, dt=np.float64)
    assert np.abs(o["G"] - G_ok).max() < 2e-5
    Bt = [np.ascontiguousarray(np.swapaxes(b, 0, 1)) for b in ag["B"]]
    G_bt = DR.neg_efe_dedup(pm, qs, ag["A"], Bt, ag["C"], A_deps, B_deps, dt=np.float64)
    assert np.abs(o["G"] - G_bt).max() > 1e-3
    G_nu = DR.neg_efe_dedup(pm, qs, ag["A"], ag["B"], ag["C"], A_deps, B_deps, use_utility=False, dt=np.float64)
    assert np.abs(o["G"] - G_nu).max() > 1e-3
