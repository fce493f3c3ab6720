"""GPU parity of the neg-EFE pipeline against the oracle.

Bar (per agent): |G - G_fp64| <= 1e-5 and |q_pi - q_pi_fp64| <= 1e-5, and
against the FP32 restatement of the reference path |x - x_fp32| <= 1e-5 + r
where r = |x_fp32 - x_fp64| is that path's own FP32 error (G sums 2 T M
entropy / ambiguity terms of opposite sign; two FP32 summation orders
legitimately differ by ~1e-5 at 32768 policies).  Argmax policies must be
identical.  The full 32768-policy workload is checked on sampled agents
with the history-dedup oracle (oracle/dedup_ref.py).
"""
import numpy as np
import pytest
import torch

from oracle import dedup_ref as D
from oracle import numpy_ref as R
from pymdp_amd import control, utils, workloads
from pymdp_amd.ffi import _efe

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _np(t):
    return t.detach().float().cpu().numpy()


def _bar(x, x32, x64, tol=TOL):
    """Asserts the bar of this module; returns (e32, e64)."""
    x = np.asarray(x, dtype=np.float64)
    e32 = float(np.abs(x - x32).max())
    e64 = float(np.abs(x - x64).max())
    r = float(np.abs(np.asarray(x32, dtype=np.float64) - x64).max())
    assert e64 <= tol, e64
    assert e32 <= tol + r, (e32, r)
    return (e32, e64)


def _agent(xs, b, core):
    return [_np(x[b]) if x.dim() == c + 1 else _np(x) for x, c in zip(xs, core)]


CASES = [
    ([3], [3], [4], [[0]], [[0]], 1),
    ([3, 2], [2, 3], [3, 4], [[0, 1], [1]], [[0], [1]], 2),
    ([3, 2, 4], [2, 3, 1], [3, 4], [[0, 1, 2], [2, 0]], [[0], [1, 0], [2]], 2),
    ([2, 3], [2, 2], [2, 3, 2], [[1], [0, 1], [1, 0]], [[0, 1], [1]], 3),
    ([8, 5, 3], [4, 3, 2], [6, 5], [[0, 1, 2], [1, 2]], [[0], [1], [2]], 2),
]


def _random_case(dev, case, batch=6, shared=False, timeC=True):
    S, U, O, Ad, Bd, T = CASES[case]
    g = torch.Generator(device=dev).manual_seed(50 + case)
    bt = None if shared else batch
    A = utils.random_A_array(O, S, Ad, batch=bt, alpha=0.5, generator=g, device=dev)
    B = utils.random_B_array(S, U, Bd, batch=bt, alpha=0.5, generator=g, device=dev)
    C = []
    for m, o in enumerate(O):
        shp = (T, o) if (timeC and m % 2 == 0) else (o,)
        C.append(torch.randn((batch,) + shp, generator=g, device=dev))
    qs = utils.random_factorized_categorical(S, batch=batch, generator=g, device=dev)
    pm = R.construct_policies(S, U, T)
    return S, U, O, Ad, Bd, T, A, B, C, qs, pm


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("shared", [False, True])
def test_neg_efe_random_parity(dev, case, shared):
    S, U, O, Ad, Bd, T, A, B, C, qs, pm = _random_case(dev, case, shared=shared)
    print("case", case, "shared", shared, S, U, O, Ad, Bd, T, flush=True)
    batch = qs[0].shape[0]
    gamma = 16.0
    q_pi, G = _efe.update_posterior_policies_ffi(pm, qs, A, B, C, None, Ad, Bd, gamma=gamma)
    torch.cuda.synchronize()
    worst = np.zeros(4)
    for b in range(batch):
        Ab = _agent(A, b, [len(d) + 1 for d in Ad])
        Bb = _agent(B, b, [len(d) + 2 for d in Bd])
        Cb = [_np(c[b]) for c in C]
        qb = [_np(q[b]) for q in qs]
        E = np.full(pm.shape[0], 1.0 / pm.shape[0])
        q32, G32 = R.update_posterior_policies(pm, qb, Ab, Bb, Cb, E, Ad, Bd, gamma=gamma, dt=np.float32)
        q64, G64 = R.update_posterior_policies(pm, qb, Ab, Bb, Cb, E, Ad, Bd, gamma=gamma, dt=np.float64)
        worst = np.maximum(worst, np.array(_bar(_np(G[b]), G32, G64) + _bar(_np(q_pi[b]), q32, q64)))
        assert int(torch.argmax(q_pi[b])) == int(np.argmax(q64))
    print(f"max|G-G32|={worst[0]:.2e} max|G-G64|={worst[1]:.2e} max|q-q32|={worst[2]:.2e} max|q-q64|={worst[3]:.2e}")


def test_term_flags_and_reference_packed_abi(dev):
    S, U, O, Ad, Bd, T, A, B, C, qs, pm = _random_case(dev, 2, timeC=False)
    batch = qs[0].shape[0]
    for use_u, use_ig in [(True, False), (False, True)]:
        _, G = _efe.update_posterior_policies_ffi(pm, qs, A, B, C, None, Ad, Bd, use_utility=use_u,
                                                  use_states_info_gain=use_ig, return_q_pi=False)
        for b in range(batch):
            args = (pm, [_np(q[b]) for q in qs], [_np(a[b]) for a in A], [_np(x[b]) for x in B],
                    [_np(c[b]) for c in C], Ad, Bd)
            kw = dict(use_utility=use_u, use_states_info_gain=use_ig)
            _bar(_np(G[b]), R.neg_efe_all_policies(*args, dt=np.float32, **kw),
                 R.neg_efe_all_policies(*args, dt=np.float64, **kw))
    # the reference's packed flat-buffer ABI gives the same G as the pointer entry
    from pymdp_amd.ffi import flatten_dep_lists, flatten_ragged_f32
    qs_flat, qs_off = flatten_ragged_f32(qs, batch)
    A_flat, A_off = flatten_ragged_f32(A, batch)
    B_flat, B_off = flatten_ragged_f32(B, batch)
    C_T = [c[:, None, :].expand(batch, T, c.shape[-1]) for c in C]
    C_flat, C_off = flatten_ragged_f32(C_T, batch)
    I = [torch.zeros((batch, 1, s), device=dev) for s in S]
    I_flat, I_off = flatten_ragged_f32(I, batch)
    Adf, Ado = flatten_dep_lists(Ad)
    Bdf, Bdo = flatten_dep_lists(Bd)
    flags = _efe.FLAG_USE_UTILITY | _efe.FLAG_USE_STATES_INFO_GAIN
    Gp = _efe.neg_efe_all_policies_packed(
        pm, qs_flat, A_flat, B_flat, C_flat, I_flat, torch.zeros(1, device=dev), S=S, O=O, U=U,
        qs_offsets=qs_off, A_offsets=A_off, B_offsets=B_off, C_offsets=C_off, I_offsets=I_off,
        I_depths=[1] * len(S), A_dep_flat=Adf, A_dep_offsets=Ado, B_dep_flat=Bdf, B_dep_offsets=Bdo, flags=flags)
    _, G = _efe.update_posterior_policies_ffi(pm, qs, A, B, C, None, Ad, Bd, return_q_pi=False)
    assert torch.allclose(Gp, G, atol=1e-6)


def _counts(x, g, zero_frac=0.15):
    """Dirichlet-count-like tensor shaped like x (some exact zeros: the
    (p > 0) mask of pymdp/control.py:546 / :594)."""
    p = x * (1.0 + 19.0 * torch.rand((), generator=g, device=x.device)) + 2.0 * torch.rand(x.shape, generator=g,
                                                                                            device=x.device)
    return torch.where(torch.rand(x.shape, generator=g, device=x.device) < zero_frac, torch.zeros_like(p), p)


@pytest.mark.parametrize("case", [1, 2, 4])
@pytest.mark.parametrize("which", ["pA", "pB", "both"])
def test_param_info_gain_parity(dev, case, which):
    """use_param_info_gain with pA and / or pB (pymdp/control.py:511-602,
    :677-682) against the oracle."""
    S, U, O, Ad, Bd, T, A, B, C, qs, pm = _random_case(dev, case)
    batch = qs[0].shape[0]
    g = torch.Generator(device=dev).manual_seed(90 + case)
    pA = [_counts(a, g) for a in A] if which in ("pA", "both") else None
    pB = [_counts(b, g) for b in B] if which in ("pB", "both") else None
    print("param info gain", case, which, flush=True)
    q_pi, G = _efe.update_posterior_policies_ffi(pm, qs, A, B, C, None, Ad, Bd, gamma=16.0, pA=pA, pB=pB,
                                                 use_param_info_gain=True)
    torch.cuda.synchronize()
    for b in range(batch):
        Ab = _agent(A, b, [len(d) + 1 for d in Ad])
        Bb = _agent(B, b, [len(d) + 2 for d in Bd])
        kw = dict(use_param_info_gain=True, pA=[_np(x[b]) for x in pA] if pA else None,
                  pB=[_np(x[b]) for x in pB] if pB else None)
        E = np.full(pm.shape[0], 1.0 / pm.shape[0])
        qb, Cb = [_np(q[b]) for q in qs], [_np(c[b]) for c in C]
        q32, G32 = R.update_posterior_policies(pm, qb, Ab, Bb, Cb, E, Ad, Bd, gamma=16.0, dt=np.float32, **kw)
        q64, G64 = R.update_posterior_policies(pm, qb, Ab, Bb, Cb, E, Ad, Bd, gamma=16.0, dt=np.float64, **kw)
        _bar(_np(G[b]), G32, G64)
        _bar(_np(q_pi[b]), q32, q64)
        assert int(torch.argmax(q_pi[b])) == int(np.argmax(q64))


def test_param_info_gain_packed_abi_and_errors(dev):
    S, U, O, Ad, Bd, T, A, B, C, qs, pm = _random_case(dev, 2, timeC=False)
    batch = qs[0].shape[0]
    g = torch.Generator(device=dev).manual_seed(5)
    pA = [_counts(a, g) for a in A]
    pB = [_counts(b, g) for b in B]
    from pymdp_amd.ffi import flatten_dep_lists, flatten_ragged_f32
    qs_flat, qs_off = flatten_ragged_f32(qs, batch)
    A_flat, A_off = flatten_ragged_f32(A, batch)
    B_flat, B_off = flatten_ragged_f32(B, batch)
    pA_flat, _ = flatten_ragged_f32(pA, batch)
    pB_flat, _ = flatten_ragged_f32(pB, batch)
    C_T = [c[:, None, :].expand(batch, T, c.shape[-1]) for c in C]
    C_flat, C_off = flatten_ragged_f32(C_T, batch)
    I = [torch.zeros((batch, 1, s), device=dev) for s in S]
    I_flat, I_off = flatten_ragged_f32(I, batch)
    Adf, Ado = flatten_dep_lists(Ad)
    Bdf, Bdo = flatten_dep_lists(Bd)
    flags = _efe.FLAG_USE_UTILITY | _efe.FLAG_USE_STATES_INFO_GAIN | _efe.FLAG_USE_PARAM_INFO_GAIN
    Gp = _efe.neg_efe_all_policies_packed(
        pm, qs_flat, A_flat, B_flat, C_flat, I_flat, torch.zeros(1, device=dev), S=S, O=O, U=U,
        qs_offsets=qs_off, A_offsets=A_off, B_offsets=B_off, C_offsets=C_off, I_offsets=I_off,
        I_depths=[1] * len(S), A_dep_flat=Adf, A_dep_offsets=Ado, B_dep_flat=Bdf, B_dep_offsets=Bdo, flags=flags,
        pA_flat=pA_flat, pB_flat=pB_flat)
    _, G = _efe.update_posterior_policies_ffi(pm, qs, A, B, C, None, Ad, Bd, return_q_pi=False, pA=pA, pB=pB,
                                              use_param_info_gain=True)
    assert torch.allclose(Gp, G, atol=1e-6)
    # the flag with neither pA nor pB is an error (pymdp/control.py:657-660)
    with pytest.raises(ValueError, match="pA or pB"):
        _efe.update_posterior_policies_ffi(pm, qs, A, B, C, None, Ad, Bd, use_param_info_gain=True)
    with pytest.raises(ValueError, match="efe_ffi: .*pA or pB"):
        _efe.neg_efe_all_policies_packed(
            pm, qs_flat, A_flat, B_flat, C_flat, I_flat, torch.zeros(1, device=dev), S=S, O=O, U=U,
            qs_offsets=qs_off, A_offsets=A_off, B_offsets=B_off, C_offsets=C_off, I_offsets=I_off,
            I_depths=[1] * len(S), A_dep_flat=Adf, A_dep_offsets=Ado, B_dep_flat=Bdf, B_dep_offsets=Bdo,
            flags=flags)


def test_inductive_parity(dev):
    S, U, O, Ad, Bd, T, A, B, C, qs, pm = _random_case(dev, 1, timeC=False)
    batch = qs[0].shape[0]
    g = torch.Generator(device=dev).manual_seed(7)
    I = [(torch.rand((batch, 3, s), generator=g, device=dev) > 0.5).float() for s in S]
    eps = torch.full((batch,), 1e-3, device=dev)
    q_pi, G = _efe.update_posterior_policies_ffi(pm, qs, A, B, C, None, Ad, Bd, gamma=4.0, I=I,
                                                 inductive_epsilon=eps, use_inductive=True)
    torch.cuda.synchronize()
    for b in range(batch):
        args = (pm, [_np(q[b]) for q in qs], [_np(a[b]) for a in A], [_np(x[b]) for x in B],
                [_np(c[b]) for c in C], Ad, Bd)
        kw = dict(I=[_np(i[b]) for i in I], inductive_epsilon=1e-3, use_inductive=True)
        _bar(_np(G[b]), R.neg_efe_all_policies(*args, dt=np.float32, **kw),
             R.neg_efe_all_policies(*args, dt=np.float64, **kw))


def test_plain_efe_full_size(dev):
    w = workloads.plain_efe(seed=0, device=dev)
    print("plain_efe", w["S"], w["U"], w["O"], w["T"], w["batch"], w["num_policies"], flush=True)
    pm = R.construct_policies(w["S"], w["U"], w["T"])
    assert pm.shape[0] == w["num_policies"] == 32768
    q_pi, G = _efe.update_posterior_policies_ffi(pm, w["qs"], w["A"], w["B"], w["C"], None, w["A_dependencies"],
                                                 w["B_dependencies"], gamma=w["gamma"])
    torch.cuda.synchronize()
    assert torch.isfinite(G).all() and torch.isfinite(q_pi).all()
    assert torch.allclose(q_pi.sum(-1), torch.ones(w["batch"], device=dev), atol=1e-4)
    worst = np.zeros(4)
    for b in (0, 63):
        args = (pm, [_np(q[b]) for q in w["qs"]], [_np(a[b]) for a in w["A"]], [_np(x[b]) for x in w["B"]],
                [_np(c[b]) for c in w["C"]], w["A_dependencies"], w["B_dependencies"])
        G64 = D.neg_efe_dedup(*args, dt=np.float64)
        G32 = D.neg_efe_dedup(*args, dt=np.float32)
        q64 = D.q_pi_from_G(G64, w["gamma"])
        q32 = D.q_pi_from_G(G32, w["gamma"], dt=np.float32)
        worst = np.maximum(worst, np.array(_bar(_np(G[b]), G32, G64) + _bar(_np(q_pi[b]), q32, q64)))
        assert int(torch.argmax(q_pi[b])) == int(np.argmax(q64))
    print(f"plain_efe max|G-G32|={worst[0]:.2e} max|G-G64|={worst[1]:.2e} max|q-q32|={worst[2]:.2e} "
          f"max|q-q64|={worst[3]:.2e}")


def test_agent_step_matches_oracle(dev):
    from pymdp_amd.agent import Agent
    w = workloads.plain_efe(seed=3, device=dev, batch=4, S=(4, 3), U=(3, 2), T=2, O=(5, 4))
    agent = Agent(A=w["A"], B=w["B"], C=w["C"], A_dependencies=w["A_dependencies"], B_dependencies=w["B_dependencies"],
                  policy_len=2, gamma=16.0, batch_size=4, num_iter=16)
    obs = utils.sample_observations(w["A"], w["qs"], w["A_dependencies"], torch.Generator(device=dev).manual_seed(1))
    qs = agent.infer_states(obs, agent.D)
    q_pi, G = agent.infer_policies(qs)
    act = agent.sample_action(q_pi)
    torch.cuda.synchronize()
    pm = agent.policies.policy_arr
    for b in range(4):
        Ab = [_np(a[b]) for a in w["A"]]
        Bb = [_np(x[b]) for x in w["B"]]
        o_b = [np.eye(a.shape[0], dtype=np.float32)[int(o[b])] for a, o in zip(Ab, obs)]
        D_b = [_np(d[b]) for d in agent.D]
        qr = R.run_factorized_fpi(Ab, o_b, D_b, w["A_dependencies"], num_iter=16, dt=np.float64)
        for f in range(2):
            assert np.abs(_np(qs[f][b, 0]) - qr[f]).max() <= TOL
        # policy inference from the GPU posterior (the step's own input)
        qg = [_np(q[b, 0]) for q in qs]
        q64, G64 = R.update_posterior_policies(pm, qg, Ab, Bb, [_np(c[b]) for c in w["C"]],
                                               np.full(pm.shape[0], 1.0 / pm.shape[0]), w["A_dependencies"],
                                               w["B_dependencies"], gamma=16.0, dt=np.float64)
        assert np.abs(_np(G[b]) - G64).max() <= TOL
        assert np.abs(_np(q_pi[b]) - q64).max() <= TOL
        assert np.array_equal(_np(act[b]).astype(int), R.sample_policy(pm, q64))
    nxt = agent.update_empirical_prior(act, qs)
    for f in range(2):
        for b in range(4):
            exp = R.compute_expected_state([_np(q[b, -1]) for q in qs], [_np(x[b]) for x in w["B"]],
                                           _np(act[b]).astype(int), w["B_dependencies"])
            assert np.abs(_np(nxt[f][b]) - exp[f]).max() <= 1e-6
