"""GPU parity of sequence inference (pymdp_hip_sequence_infer: marginal and
variational message passing over a window) against the oracle
(oracle/seq_ref.py, pymdp/algos.py:157-500, pymdp/inference.py:56-271).
Bar per agent: |q - q_fp64| <= 1e-5 and |q - q_fp32| <= 1e-5 + |q_fp32 -
q_fp64| (the kernel accumulates in FP64)."""
import numpy as np
import pytest
import torch

from oracle import seq_ref as Q
from pymdp_amd import inference, utils
from pymdp_amd.agent import Agent
from pymdp_amd.ffi._seq import sequence_infer_ffi

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _np(t):
    return t.detach().double().cpu().numpy()


def _model(dev, seed, S, U, O, deps, batch, shared=False):
    g = utils._gen(seed, dev)
    A = utils.random_A_array(O, S, deps, batch=None if shared else batch, alpha=0.7, generator=g, device=dev)
    B = utils.random_B_array(S, U, [[f] for f in range(len(S))], batch=None if shared else batch, alpha=1.0,
                             generator=g, device=dev)
    prior = utils.random_factorized_categorical(S, batch=batch, alpha=1.0, generator=g, device=dev)
    return A, B, prior


def _check(method, q, A, B, prior, obs, deps, acts, valid_steps, distr, batch, num_iter):
    worst = 0.0
    for b in range(batch):
        Ab = [_np(a[b]) if a.dim() == len(d) + 2 else _np(a) for a, d in zip(A, deps)]
        Bb = [_np(x[b]) if x.dim() == 4 else _np(x) for x in B]
        pb = [_np(p[b]) for p in prior]
        ob = [_np(o[b]) if distr else o[b].cpu().numpy() for o in obs]
        ab = None if acts is None else acts[b].cpu().numpy()
        vs = None if valid_steps is None else int(valid_steps[b])
        r64 = Q.run_sequence(method, Ab, Bb, ob, pb, deps, ab, num_iter=num_iter, distr_obs=distr, valid_steps=vs,
                             dt=np.float64)
        r32 = Q.run_sequence(method, Ab, Bb, ob, pb, deps, ab, num_iter=num_iter, distr_obs=distr, valid_steps=vs,
                             dt=np.float32)
        for f, x in enumerate(q):
            g = _np(x[b])
            e64 = float(np.abs(g - r64[f]).max())
            r = float(np.abs(r32[f] - r64[f]).max())
            e32 = float(np.abs(g - r32[f]).max())
            assert e64 <= TOL, (b, f, e64)
            assert e32 <= TOL + r, (b, f, e32, r)
            worst = max(worst, e64)
    return worst


@pytest.mark.parametrize("method", ["mmp", "vmp"])
@pytest.mark.parametrize("case", ["index", "distr_masks", "shared_negative_action", "one_step"])
def test_sequence_parity(dev, method, case):
    print(f"sequence {method} {case}", flush=True)
    S, U, O, deps = [4, 3, 5], [3, 2, 1], [6, 4, 3], [[0, 1], [1, 2], [2]]
    batch, T = 8, (1 if case == "one_step" else 6)
    A, B, prior = _model(dev, 11, S, U, O, deps, batch, shared=(case == "shared_negative_action"))
    g = torch.Generator(device=dev)
    g.manual_seed(5)
    idx = [torch.randint(0, o, (batch, T), generator=g, device=dev) for o in O]
    distr = case == "distr_masks"
    obs = [torch.nn.functional.one_hot(i, o).float() for i, o in zip(idx, O)] if distr else idx
    acts = torch.stack([torch.randint(0, u, (batch, max(T - 1, 0)), generator=g, device=dev) for u in U], -1)
    valid_steps = None
    if case == "distr_masks":
        valid_steps = torch.tensor([6, 5, 4, 3, 2, 1, 6, 3], device=dev)
        acts[1, 3, 0] = -1
    if case == "shared_negative_action":
        acts[0, 2, 1] = -1  # no masks: the neutral transition
    q = sequence_infer_ffi(method, A, B, obs, prior, deps, [[0], [1], [2]], past_actions=acts,
                           valid_steps=valid_steps, num_iter=16, distr_obs=distr)
    torch.cuda.synchronize()
    worst = _check(method, q, A, B, prior, obs, deps, acts, valid_steps, distr, batch, 16)
    print(f"sequence {method} {case} max|q-q64|={worst:.3e}")


def test_update_posterior_states_and_agent_sequence(dev):
    S, U, O, deps = [3, 4], [2, 2], [4, 3], [[0], [0, 1]]
    batch, T = 4, 5
    A, B, prior = _model(dev, 3, S, U, O, deps, batch)
    g = torch.Generator(device=dev)
    g.manual_seed(9)
    obs = [torch.randint(0, o, (batch, T), generator=g, device=dev) for o in O]
    acts = torch.stack([torch.randint(0, u, (batch, T - 1), generator=g, device=dev) for u in U], -1)
    q = inference.update_posterior_states(A, B, obs, acts, prior=prior, A_dependencies=deps, num_iter=8,
                                          method="mmp", distr_obs=False, inference_horizon=3)
    torch.cuda.synchronize()
    assert tuple(q[0].shape) == (batch, 3, 3)
    _check("mmp", q, A, B, prior, [o[:, -3:] for o in obs], deps, acts[:, -2:], None, False, batch, 8)
    ag = Agent(A, B, D=prior, A_dependencies=deps, batch_size=batch, inference_algo="vmp", num_iter=8, device=dev)
    qa = ag.infer_states(obs, ag.D, past_actions=acts)
    torch.cuda.synchronize()
    _check("vmp", qa, A, B, prior, obs, deps, acts, None, False, batch, 8)
    assert ag.update_empirical_prior(acts[:, -1], qa) is ag.D


def test_sequence_errors(dev):
    A, B, prior = _model(dev, 1, [3, 2], [2, 2], [3], [[0, 1]], 2)
    obs = [torch.zeros((2, 3), dtype=torch.long, device=dev)]
    with pytest.raises(NotImplementedError, match="factor-local"):
        sequence_infer_ffi("mmp", A, B, obs, prior, [[0, 1]], [[0, 1], [1]], num_iter=2, distr_obs=False)
    with pytest.raises(ValueError, match="seq_ffi"):  # several controls need actions
        sequence_infer_ffi("mmp", A, B, obs, prior, [[0, 1]], num_iter=2, distr_obs=False)
