"""Writes tests/golden/fpi_efe_small.npz: small seeded FPI and neg-EFE
cases with the FP64 oracle's outputs.  The reference holds no stored
vectors for these functions (its tests recompute them), so these fixtures
pin the oracle and the HIP path against regressions.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import numpy_ref as R  # noqa: E402


def _dirichlet(rng, shape, alpha):
    x = rng.gamma(alpha, size=shape)
    return (x / x.sum(axis=0, keepdims=True)).astype(np.float32)


def main():
    rng = np.random.default_rng(2024)
    out = {}
    # FPI: S=[4,3], O=[5,6], both modalities on both factors; 4 agents, 16 sweeps
    S, O, deps = [4, 3], [5, 6], [[0, 1], [0, 1]]
    for b in range(4):
        A = [_dirichlet(rng, (o, *S), 1.0) for o in O]
        D = [_dirichlet(rng, (s,), 1.0) for s in S]
        obs = [int(rng.integers(0, o)) for o in O]
        ob = [np.eye(o, dtype=np.float32)[x] for o, x in zip(O, obs)]
        q = R.run_factorized_fpi(A, ob, D, deps, num_iter=16, dt=np.float64)
        for m in range(2):
            out[f"fpi{b}_A{m}"] = A[m]
        for f in range(2):
            out[f"fpi{b}_D{f}"] = D[f]
            out[f"fpi{b}_q{f}"] = q[f]
        out[f"fpi{b}_obs"] = np.asarray(obs, np.int32)
    # sparse dependencies: S=[3,2,4], O=[3,4,2], A_deps=[[0,2],[1],[2,1,0]]
    S2, O2, deps2 = [3, 2, 4], [3, 4, 2], [[0, 2], [1], [2, 1, 0]]
    A = [_dirichlet(rng, (o, *[S2[d] for d in dl]), 0.5) for o, dl in zip(O2, deps2)]
    D = [_dirichlet(rng, (s,), 1.0) for s in S2]
    obs = [int(rng.integers(0, o)) for o in O2]
    ob = [np.eye(o, dtype=np.float32)[x] for o, x in zip(O2, obs)]
    q = R.run_factorized_fpi(A, ob, D, deps2, num_iter=16, dt=np.float64)
    for m in range(3):
        out[f"sparse_A{m}"] = A[m]
    for f in range(3):
        out[f"sparse_D{f}"] = D[f]
        out[f"sparse_q{f}"] = q[f]
    out["sparse_obs"] = np.asarray(obs, np.int32)
    # neg-EFE: S=[3,2], U=[2,3], horizon 2, O=[4,3], gamma 16
    S3, U3, O3 = [3, 2], [2, 3], [4, 3]
    Ad, Bd = [[0, 1], [1]], [[0], [1]]
    A = [_dirichlet(rng, (o, *[S3[d] for d in dl]), 0.5) for o, dl in zip(O3, Ad)]
    B = [_dirichlet(rng, (s, s, u), 0.5) for s, u in zip(S3, U3)]
    C = [rng.normal(size=o).astype(np.float32) for o in O3]
    qs = [_dirichlet(rng, (s,), 1.0) for s in S3]
    pm = R.construct_policies(S3, U3, 2)
    q_pi, G = R.update_posterior_policies(pm, qs, A, B, C, np.full(pm.shape[0], 1.0 / pm.shape[0]), Ad, Bd,
                                          gamma=16.0, dt=np.float64)
    for m in range(2):
        out[f"efe_A{m}"] = A[m]
        out[f"efe_C{m}"] = C[m]
    for f in range(2):
        out[f"efe_B{f}"] = B[f]
        out[f"efe_qs{f}"] = qs[f]
    out["efe_policies"] = pm.astype(np.int32)
    out["efe_G"] = G
    out["efe_q_pi"] = q_pi
    out["efe_action"] = np.asarray(R.sample_policy(pm, q_pi), np.int32)
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "fpi_efe_small.npz"), **out)


if __name__ == "__main__":
    main()
