"""GPU parity of the one-step VFE diagnostics (pymdp_hip_fpi_vfe through
ffi.fpi_vfe_ffi and Agent.infer_states(return_info=True)) against the
oracle's calc_vfe_one_step (pymdp/maths.py:448-762) evaluated on the same
posterior.  Bar: |x - x64| <= 1e-5 * max(1, |x64|) for each of the four
terms (FP32 outputs of FP64 sums whose magnitude grows with the number of
modalities)."""
import numpy as np
import pytest
import torch

from oracle import numpy_ref as R
from pymdp_amd import utils
from pymdp_amd.agent import Agent
from pymdp_amd.ffi import fpi_vfe_ffi, infer_states_ffi

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _np(t):
    return t.detach().double().cpu().numpy()


def _check(v, qs, A, obs, prior, deps, distr, batch):
    worst = 0.0
    for b in range(batch):
        A_b = [_np(a[b]) if a.dim() == len(d) + 2 else _np(a) for a, d in zip(A, deps)]
        ob = [_np(o[b]) for o in obs] if distr else [int(o[b]) for o in obs]
        q = [_np(x[b]) for x in qs]
        p = [_np(x[b]) for x in prior]
        ref = np.array(R.calc_vfe_one_step(q, p, ob, A_b, deps, distr_obs=distr, dt=np.float64))
        err = np.abs(_np(v[b]) - ref) / np.maximum(1.0, np.abs(ref))
        worst = max(worst, float(err.max()))
    assert worst <= TOL, worst
    return worst


@pytest.mark.parametrize("S,O,deps,distr,shared", [
    ([4, 3], [5, 6], [[0, 1], [0, 1]], False, False),
    ([16, 16, 8, 8], [32, 32], [[0, 1, 2, 3], [0, 1, 2, 3]], False, False),
    ([6, 5, 4], [7, 3, 9], [[0, 2], [1], [0, 1, 2]], True, False),
    ([8, 3], [4, 4], [[0], [0, 1]], False, True),
])
def test_fpi_vfe_parity(dev, S, O, deps, distr, shared):
    print(f"vfe S={S} O={O} deps={deps} distr={distr} shared={shared}", flush=True)
    g = utils._gen(7, dev)
    batch = 16
    A = utils.random_A_array(O, S, deps, batch=None if shared else batch, alpha=0.5, generator=g, device=dev)
    prior = utils.random_factorized_categorical(S, batch=batch, alpha=1.0, generator=g, device=dev)
    idx = [torch.randint(0, o, (batch,), generator=g, device=dev) for o in O]
    obs = [torch.nn.functional.one_hot(i, o).float() for i, o in zip(idx, O)] if distr else idx
    qs = infer_states_ffi(A, obs, prior, deps, 16, distr_obs=distr, batch=batch)
    v = fpi_vfe_ffi(A, obs, prior, torch.cat(qs, -1), deps, distr_obs=distr, batch=batch)
    torch.cuda.synchronize()
    worst = _check(v, qs, A, obs, prior, deps, distr, batch)
    print(f"vfe max rel err {worst:.3e}")


def test_agent_infer_states_return_info(dev):
    g = utils._gen(3, dev)
    batch, S, O = 8, [5, 4], [6, 3]
    deps = [[0, 1], [1]]
    A = utils.random_A_array(O, S, deps, batch=batch, alpha=0.5, generator=g, device=dev)
    B = utils.random_B_array(S, [2, 2], [[0], [1]], batch=batch, alpha=1.0, generator=g, device=dev)
    ag = Agent(A, B, A_dependencies=deps, batch_size=batch, device=dev)
    obs = [torch.randint(0, o, (batch,), generator=g, device=dev) for o in O]
    qs, info = ag.infer_states(obs, ag.D, return_info=True)
    qs_plain = ag.infer_states(obs, ag.D)
    torch.cuda.synchronize()
    for a, b in zip(qs, qs_plain):
        assert torch.equal(a, b)
    assert set(info) == {"vfe_t", "vfe", "vfe_components"}
    comp = info["vfe_components"]
    for k in ("neg_entropy_t", "prior_cross_entropy_t", "transition_cross_entropy_t", "accuracy_t", "state_vfe",
              "parameter_kl_A", "parameter_kl_B", "parameter_kl"):
        assert tuple(comp[k].shape) == (batch,)
    v = torch.stack([comp["neg_entropy_t"], comp["prior_cross_entropy_t"], comp["accuracy_t"], info["vfe"]], -1)
    _check(v, [q[:, -1] for q in qs], A, obs, ag.D, deps, False, batch)
    assert torch.equal(info["vfe_t"], info["vfe"]) and not comp["parameter_kl"].any()
