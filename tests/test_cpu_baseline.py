"""The C/OpenMP CPU baseline (oracle/cpu_baseline.c) agrees with the NumPy
oracle, so the timed CPU number is the same computation."""
import numpy as np
import pytest

from oracle import cpu_ref
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
