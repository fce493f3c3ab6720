"""GPU parity of the Dirichlet-learning kernels (pymdp_hip_update_pA /
pymdp_hip_update_pB through pymdp_amd.learning and Agent.infer_parameters)
against the oracle (oracle/numpy_ref.py, pymdp/learning.py:11-234,
pymdp/agent.py:416-620).

Bar: posterior counts |q - q64| <= 1e-5 * max(1, |q64|) (FP32 storage of
counts that grow with learning), expected values |E - E64| <= 1e-5, and the
same within the FP32-oracle's own distance from FP64."""
import numpy as np
import pytest
import torch

from oracle import numpy_ref as R
from pymdp_amd import learning
from pymdp_amd.agent import Agent

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _np(t):
    return t.detach().float().cpu().numpy()


def _check(x, x64, rel):
    x = np.asarray(x, np.float64)
    scale = np.maximum(1.0, np.abs(x64)) if rel else 1.0
    err = float(np.max(np.abs(x - x64) / scale))
    assert err <= TOL, err
    return err


@pytest.mark.parametrize("O,S,deps,T,shared,categorical", [
    (5, [3, 4], [0, 1], 1, False, False),
    (32, [16, 8, 4], [0, 1, 2], 3, False, True),
    (100, [6, 5], [1], 2, True, False),      # more outcomes than registers hold: chunked path
    (64, [7], [0], 1, False, True),
])
def test_update_pA_parity(dev, O, S, deps, T, shared, categorical):
    print(f"update_pA O={O} S={S} deps={deps} T={T} shared={shared}", flush=True)
    rng = np.random.default_rng(O + T)
    batch = 6
    shape = (O,) + tuple(S[d] for d in deps)
    pA = rng.gamma(1.5, size=shape if shared else (batch,) + shape).astype(np.float32)
    qs = [rng.dirichlet(np.ones(s), size=(batch, T)).astype(np.float32) for s in S]
    if categorical:
        obs = rng.dirichlet(np.ones(O), size=(batch, T)).astype(np.float32)
    else:
        obs = rng.integers(0, O, size=(batch, T))
    lr = rng.uniform(0.5, 2.0, size=batch).astype(np.float32)
    A = [torch.zeros(1, device=dev)]
    qA, EA = learning.update_obs_likelihood_dirichlet(
        [torch.as_tensor(pA, device=dev)], A, [torch.as_tensor(obs, device=dev)],
        [torch.as_tensor(q, device=dev) for q in qs], A_dependencies=[deps], categorical_obs=categorical,
        num_obs=[O], lr=torch.as_tensor(lr, device=dev))
    torch.cuda.synchronize()
    for b in range(batch):
        pa = pA if shared else pA[b]
        ob = obs[b] if categorical else np.eye(O)[obs[b]]
        q64, e64 = R.update_obs_likelihood_dirichlet_m(pa, ob, [q[b] for q in qs], deps, lr=float(lr[b]),
                                                       dt=np.float64)
        _check(_np(qA[0][b]), q64, True)
        _check(_np(EA[0][b]), e64, False)


@pytest.mark.parametrize("mask,S,U,B_deps", [
    (False, [6, 3, 70], [4, 1, 2], [[0, 1], [1], [2]]),
    (True, [6, 3, 70], [4, 1, 2], [[0, 1], [1], [2]]),
    (True, [40, 12, 3], [4, 1, 3], [[0, 1], [1], [2]]),  # full 128-column tiles, 16-byte path
])
def test_update_pB_parity(dev, mask, S, U, B_deps):
    print(f"update_pB mask={mask} S={S}", flush=True)
    rng = np.random.default_rng(9)
    batch, T = 5, 4
    qs = [rng.dirichlet(np.ones(s), size=(batch, T)).astype(np.float32) for s in S]
    acts = rng.integers(-1, 4, size=(batch, T - 1, 3)).astype(np.int32)
    acts[:, :, 1] = np.where(acts[:, :, 1] < 0, -1, 0)
    pB = [rng.gamma(1.0, size=(batch, S[f]) + tuple(S[d] for d in B_deps[f]) + (U[f],)).astype(np.float32)
          for f in range(3)]
    joints = [[torch.as_tensor(qs[f][:, 1:], device=dev)] + [torch.as_tensor(qs[d][:, :-1], device=dev)
                                                              for d in B_deps[f]] for f in range(3)]
    qB, EB = learning.update_state_transition_dirichlet(
        [torch.as_tensor(p, device=dev) for p in pB], [None] * 3, joints, torch.as_tensor(acts, device=dev),
        num_controls=U, lr=0.8, mask_invalid=mask)
    torch.cuda.synchronize()
    for b in range(batch):
        valid = np.all(acts[b] >= 0, axis=-1).astype(np.float64) if mask else np.ones(T - 1)
        jb = [[qs[f][b, 1:] * valid[:, None]] + [qs[d][b, :-1] * valid[:, None] for d in B_deps[f]]
              for f in range(3)]
        q64, e64 = R.update_state_transition_dirichlet([p[b] for p in pB], [None] * 3, jb, acts[b], U, 0.8,
                                                       dt=np.float64)
        for f in range(3):
            _check(_np(qB[f][b]), q64[f], True)
            _check(_np(EB[f][b]), e64[f], False)


def test_update_pB_factors_to_update(dev):
    rng = np.random.default_rng(3)
    batch = 3
    qs = [torch.as_tensor(rng.dirichlet(np.ones(4), size=(batch, 3)), dtype=torch.float32, device=dev)]
    pB = [torch.as_tensor(rng.gamma(1.0, size=(batch, 4, 4, 2)), dtype=torch.float32, device=dev)]
    acts = torch.ones((batch, 2, 1), dtype=torch.int32, device=dev)
    qB, EB = learning.update_state_transition_dirichlet(pB, [None], [[qs[0][:, 1:], qs[0][:, :-1]]], acts,
                                                        num_controls=[2], lr=1.0, factors_to_update=[])
    torch.cuda.synchronize()
    assert torch.equal(qB[0], pB[0])
    p64 = _np(pB[0]).astype(np.float64)
    for b in range(batch):
        _check(_np(EB[0][b]), R.dirichlet_expected_value(p64[b], np.float64), False)


def test_learning_errors(dev):
    p = torch.ones((2, 3, 4), device=dev)
    with pytest.raises(ValueError, match="pA"):
        learning.update_obs_likelihood_dirichlet_m(p, torch.ones((2, 1, 5), device=dev),
                                                   [torch.ones((2, 1, 4), device=dev) / 4], [0])
    with pytest.raises(ValueError, match="actions"):
        learning.update_state_transition_dirichlet(
            [torch.ones((2, 3, 3, 2), device=dev)], [None], [[torch.ones((2, 2, 3), device=dev) / 3] * 2],
            torch.zeros((2, 1, 1), dtype=torch.int32, device=dev), num_controls=[2], lr=1.0)


def _agent_model(rng, batch):
    S, U, O = [5, 3], [3, 2], [6, 4]
    A_deps = [[0, 1], [1]]
    A = [rng.dirichlet(np.ones(o), size=tuple(S[d] for d in A_deps[m])).transpose(-1, *range(len(A_deps[m])))
         for m, o in enumerate(O)]
    B = [rng.dirichlet(np.ones(s), size=(s, U[f])).transpose(2, 0, 1) for f, s in enumerate(S)]
    pA = [rng.gamma(2.0, size=(batch,) + a.shape) for a in A]
    pB = [rng.gamma(2.0, size=(batch,) + b.shape) for b in B]
    return S, U, O, A_deps, A, B, pA, pB


@pytest.mark.parametrize("all_invalid", [False, True])
def test_agent_infer_parameters(dev, all_invalid):
    """Agent.infer_parameters (learn_A and learn_B) on a two-step belief
    sequence against the oracle's agent-level update, per agent; with no
    valid transition in the batch B stays B."""
    print(f"agent infer_parameters all_invalid={all_invalid}", flush=True)
    rng = np.random.default_rng(5)
    batch = 4
    S, U, O, A_deps, A, B, pA, pB = _agent_model(rng, batch)
    f32 = lambda x: torch.as_tensor(np.asarray(x), dtype=torch.float32, device=dev)
    ag = Agent([f32(a) for a in A], [f32(b) for b in B], pA=[f32(p) for p in pA], pB=[f32(p) for p in pB],
               A_dependencies=A_deps, batch_size=batch, learn_A=True, learn_B=True, device=dev)
    qs = [rng.dirichlet(np.ones(s), size=(batch, 2)).astype(np.float32) for s in S]
    obs = [rng.integers(0, o, size=(batch, 2)) for o in O]
    acts = rng.integers(0, 2, size=(batch, 2, 2)).astype(np.int32)
    acts[1, 0, 1] = -1  # one agent's transition is masked
    if all_invalid:
        acts[:, 0, 0] = -1
    new = ag.infer_parameters([f32(q) for q in qs], [torch.as_tensor(o, device=dev) for o in obs],
                              torch.as_tensor(acts, device=dev))
    torch.cuda.synchronize()
    any_valid = bool(np.all(acts[:, :1] >= 0, axis=-1).any())
    for b in range(batch):
        pa, a, pb, bb = R.infer_parameters(
            [p[b] for p in pA], A, [p[b] for p in pB], B, [q[b] for q in qs], [o[b] for o in obs], acts[b],
            [[0], [1]], A_deps, O, U, learn_A=True, learn_B=True, any_valid=any_valid, dt=np.float64)
        for m in range(2):
            _check(_np(new.pA[m][b]), pa[m], True)
            _check(_np(new.A[m][b]), a[m], False)
        for f in range(2):
            nb = new.B[f][b] if new.B[f].dim() == 4 else new.B[f]
            npb = new.pB[f][b]
            _check(_np(npb), pb[f], True)
            _check(_np(nb), bb[f], False)
    if all_invalid:
        assert new.B is ag.B and new.pB is ag.pB
    # the learned model feeds the next infer_policies call
    q_pi, G = new.infer_policies([f32(q[:, -1:]) for q in qs])
    torch.cuda.synchronize()
    assert torch.isfinite(G).all() and torch.allclose(q_pi.sum(-1), torch.ones(batch, device=dev), atol=1e-5)
