"""Committed fixtures (tests/golden/fpi_efe_small.npz, written by
tests/golden/make_golden.py): the oracle reproduces them (CPU), and the HIP
path matches them to 1e-5 (GPU)."""
import os

import numpy as np
import pytest

from oracle import numpy_ref as R

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fpi_efe_small.npz")


def _load():
    return np.load(GOLD)


def test_oracle_reproduces_fixtures():
    z = _load()
    for b in range(4):
        A = [z[f"fpi{b}_A{m}"] for m in range(2)]
        D = [z[f"fpi{b}_D{f}"] for f in range(2)]
        ob = [np.eye(a.shape[0], dtype=np.float32)[int(x)] for a, x in zip(A, z[f"fpi{b}_obs"])]
        q = R.run_factorized_fpi(A, ob, D, [[0, 1], [0, 1]], num_iter=16, dt=np.float64)
        for f in range(2):
            np.testing.assert_allclose(q[f], z[f"fpi{b}_q{f}"], rtol=0, atol=1e-12)
    A = [z[f"efe_A{m}"] for m in range(2)]
    B = [z[f"efe_B{f}"] for f in range(2)]
    C = [z[f"efe_C{m}"] for m in range(2)]
    qs = [z[f"efe_qs{f}"] for f in range(2)]
    pm = z["efe_policies"]
    q_pi, G = R.update_posterior_policies(pm, qs, A, B, C, np.full(pm.shape[0], 1.0 / pm.shape[0]),
                                          [[0, 1], [1]], [[0], [1]], gamma=16.0, dt=np.float64)
    np.testing.assert_allclose(G, z["efe_G"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(q_pi, z["efe_q_pi"], rtol=0, atol=1e-12)


@pytest.mark.gpu
def test_hip_path_matches_fixtures(dev):
    import torch

    from pymdp_amd.ffi import infer_states_ffi
    from pymdp_amd.ffi._efe import update_posterior_policies_ffi
    z = _load()
    t = lambda x: torch.as_tensor(np.ascontiguousarray(x), device=dev)  # noqa: E731
    A = [torch.stack([t(z[f"fpi{b}_A{m}"]) for b in range(4)]) for m in range(2)]
    D = [torch.stack([t(z[f"fpi{b}_D{f}"]) for b in range(4)]) for f in range(2)]
    obs = [torch.tensor([int(z[f"fpi{b}_obs"][m]) for b in range(4)], device=dev) for m in range(2)]
    q = infer_states_ffi(A, obs, D, [[0, 1], [0, 1]], 16, distr_obs=False)
    for f in range(2):
        ref = np.stack([z[f"fpi{b}_q{f}"] for b in range(4)])
        assert np.abs(q[f].cpu().numpy() - ref).max() <= 1e-5
    As = [t(z[f"sparse_A{m}"])[None] for m in range(3)]
    Ds = [t(z[f"sparse_D{f}"])[None] for f in range(3)]
    os_ = [torch.tensor([int(x)], device=dev) for x in z["sparse_obs"]]
    qs_ = infer_states_ffi(As, os_, Ds, [[0, 2], [1], [2, 1, 0]], 16, distr_obs=False)
    for f in range(3):
        assert np.abs(qs_[f][0].cpu().numpy() - z[f"sparse_q{f}"]).max() <= 1e-5
    q_pi, G = update_posterior_policies_ffi(z["efe_policies"], [t(z[f"efe_qs{f}"])[None] for f in range(2)],
                                            [t(z[f"efe_A{m}"])[None] for m in range(2)],
                                            [t(z[f"efe_B{f}"])[None] for f in range(2)],
                                            [t(z[f"efe_C{m}"])[None] for m in range(2)], None, [[0, 1], [1]],
                                            [[0], [1]], gamma=16.0)
    assert np.abs(G[0].cpu().numpy() - z["efe_G"]).max() <= 1e-5
    assert np.abs(q_pi[0].cpu().numpy() - z["efe_q_pi"]).max() <= 1e-5
    assert int(np.argmax(q_pi[0].cpu().numpy())) == int(np.argmax(z["efe_q_pi"]))
