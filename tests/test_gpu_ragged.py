"""GPU parity of the batched ragged step (pymdp_hip_ragged_*) against the
oracle, per agent: qs, G and q_pi to the bar |x - x_fp64| <= 1e-5 and
|x - x_fp32| <= 1e-5 + |x_fp32 - x_fp64|; actions equal to the oracle's
first-index argmax policy's first actions.  The full 4096-agent batch is
also checked through size-independent properties (normalisation,
finiteness, bitwise determinism, graph replay == eager step)."""
import numpy as np
import pytest
import torch

from oracle import dedup_ref as DR
from oracle import numpy_ref as R
from pymdp_amd import ragged, workloads
from tests import ragged_emu as EM
from tests.test_ragged_host import _oracle_rollout

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _np(t):
    return t.detach().float().cpu().numpy()


def _bar(x, x32, x64, tol=TOL):
    x = np.asarray(x, dtype=np.float64)
    e32 = float(np.abs(x - x32).max())
    e64 = float(np.abs(x - x64).max())
    r = float(np.abs(np.asarray(x32, dtype=np.float64) - x64).max())
    assert e64 <= tol, e64
    assert e32 <= tol + r, (e32, r)
    return e32, e64


def _check_agent(rb, i, use_dedup=False):
    a = rb.agents[i]
    qs, G, q_pi, act = rb.agent_outputs(i)
    A = [_np(x) for x in a["A"]]
    B = [_np(x) for x in a["B"]]
    C = [_np(x) for x in a["C"]]
    D = [_np(x) for x in a["D"]]
    ob = [np.eye(o, dtype=np.float32)[x] for o, x in zip(a["O"], a["obs"])]
    errs = []
    q32 = R.run_factorized_fpi(A, ob, D, a["A_dependencies"], num_iter=16, dt=np.float32)
    q64 = R.run_factorized_fpi(A, ob, D, a["A_dependencies"], num_iter=16, dt=np.float64)
    for f in range(len(a["S"])):
        errs.append(_bar(_np(qs[f]), q32[f], q64[f]))
    pm = a["policies"]
    # the policy stage starts from the kernel's posterior (as infer_policies does)
    qg = [_np(q).astype(np.float64) for q in qs]
    if use_dedup:
        G32 = DR.neg_efe_dedup(pm, qg, A, B, C, a["A_dependencies"], a["B_dependencies"], dt=np.float32)
        G64 = DR.neg_efe_dedup(pm, qg, A, B, C, a["A_dependencies"], a["B_dependencies"], dt=np.float64)
        p32 = DR.q_pi_from_G(G32, a["gamma"], dt=np.float32)
        p64 = DR.q_pi_from_G(G64, a["gamma"], dt=np.float64)
    else:
        P = pm.shape[0]
        E = np.full(P, 1.0 / P)
        p32, G32 = R.update_posterior_policies(pm, qg, A, B, C, E, a["A_dependencies"], a["B_dependencies"],
                                               gamma=a["gamma"], dt=np.float32)
        p64, G64 = R.update_posterior_policies(pm, qg, A, B, C, E, a["A_dependencies"], a["B_dependencies"],
                                               gamma=a["gamma"], dt=np.float64)
    errs.append(_bar(_np(G), G32, G64))
    errs.append(_bar(_np(q_pi), p32, p64))
    np.testing.assert_array_equal(act.cpu().numpy(), R.sample_policy(pm, _np(q_pi)))
    np.testing.assert_array_equal(act.cpu().numpy(), R.sample_policy(pm, p64))
    return max(e[0] for e in errs), max(e[1] for e in errs)


def test_ragged_small_parity(dev):
    w = workloads.ragged(seed=1, device=dev, n_agents=64, state_range=(2, 9), obs_range=(2, 9), max_policies=256,
                         max_A_bytes=1 << 16)
    print("ragged small: 64 agents", flush=True)
    rb = ragged.RaggedBatch(w["agents"], num_iter=16, device=dev)
    rb.step()
    torch.cuda.synchronize()
    worst = [0.0, 0.0]
    for i in range(64):
        e32, e64 = _check_agent(rb, i)
        worst = [max(worst[0], e32), max(worst[1], e64)]
    print(f"ragged small max|x-x32|={worst[0]:.3e} max|x-x64|={worst[1]:.3e}")


def test_ragged_wide_shapes_parity(dev):
    """Full-range shapes (rank-3 dependencies, 64-state factors, workspace
    paths) with the policy count capped so the oracle stays quick."""
    w = workloads.ragged(seed=5, device=dev, n_agents=24, max_policies=64)
    print("ragged wide: 24 agents", flush=True)
    rb = ragged.RaggedBatch(w["agents"], num_iter=16, device=dev)
    rb.step()
    torch.cuda.synchronize()
    for i in range(24):
        _check_agent(rb, i, use_dedup=True)


def test_ragged_team_agents_parity(dev):
    """Agents whose log-likelihood spans several FPI units: the units of one
    agent exchange partial marginals every sweep (team barrier)."""
    w = workloads.ragged(seed=0, device=dev, n_agents=5, state_range=(20, 40), obs_range=(2, 6), max_policies=64,
                         max_A_bytes=1 << 22, max_modalities=3)
    print("ragged teams: 5 agents", flush=True)
    rb = ragged.RaggedBatch(w["agents"], num_iter=16, device=dev)
    assert rb.plan.info()["team_max"] > 2
    for _ in range(2):
        rb.step()
    torch.cuda.synchronize()
    assert rb.status() == 0
    for i in range(5):
        _check_agent(rb, i, use_dedup=True)


def test_ragged_team_exchange_across_tag_wrap(dev):
    """The team exchange tags each partial with the sweep sequence number
    (step * num_iter + sweep) in 10 mantissa bits, which wrap every 64 steps
    at 16 sweeps: 80 graph replays of a team workload stay bitwise identical
    to the first step, with no exchange timeout."""
    w = workloads.ragged(seed=4, device=dev, n_agents=8, state_range=(20, 40), obs_range=(2, 6), max_policies=64,
                         max_A_bytes=1 << 22, max_modalities=3)
    print("ragged team tags: 8 agents, 80 steps", flush=True)
    rb = ragged.RaggedBatch(w["agents"], num_iter=16, device=dev)
    assert rb.plan.info()["team_max"] > 2
    rb.step()
    torch.cuda.synchronize()
    out1, act1 = rb.out.clone(), rb.act.clone()
    g = rb.capture()
    for _ in range(80):
        g.replay()
    torch.cuda.synchronize()
    assert rb.status() == 0
    assert torch.equal(rb.out, out1) and torch.equal(rb.act, act1)


def test_ragged_full_batch(dev):
    w = workloads.ragged(seed=0, device=dev, n_agents=4096)
    print("ragged full: 4096 agents", flush=True)
    rb = ragged.RaggedBatch(w["agents"], num_iter=16, device=dev)
    rb.step()
    torch.cuda.synchronize()
    out1 = rb.out.clone()
    act1 = rb.act.clone()
    assert rb.status() == 0
    assert rb.plan.info()["team_max"] > 1
    assert torch.isfinite(out1).all()
    for i in range(0, 4096, 7):
        qs, G, q_pi, act = rb.agent_outputs(i)
        for q in qs:
            assert abs(float(q.sum()) - 1.0) < 1e-5
        assert abs(float(q_pi.sum()) - 1.0) < 1e-4
    # a second step is bitwise identical
    rb.step()
    torch.cuda.synchronize()
    assert torch.equal(out1, rb.out) and torch.equal(act1, rb.act)
    # oracle on a sample: the most expensive FPI agents and a spread of others
    cost = [sum(a_.numel() for a_ in a["A"]) * np.asarray(a["policies"]).shape[0] for a in w["agents"]]
    order = np.argsort(cost)
    sample = list(order[:: len(order) // 10][:10]) + [int(np.argmax([sum(a_.numel() for a_ in a["A"])
                                                                       for a in w["agents"]]))]
    for i in sample:
        _check_agent(rb, int(i), use_dedup=True)
    # every agent whose FPI units hold the log-likelihood as floats (those
    # too large for one unit of doubles): qs to the same bar
    _, units, _, _ = EM.tables(rb.plan)
    x32 = sorted({int(u["agent"]) for u in units if int(u["x32"])})
    assert len(x32) > 100
    worst = [0.0, 0.0]
    for i in x32:
        a = w["agents"][i]
        A = [_np(x) for x in a["A"]]
        D = [_np(x) for x in a["D"]]
        ob = [np.eye(o, dtype=np.float32)[x] for o, x in zip(a["O"], a["obs"])]
        q32 = R.run_factorized_fpi(A, ob, D, a["A_dependencies"], num_iter=16, dt=np.float32)
        q64 = R.run_factorized_fpi(A, ob, D, a["A_dependencies"], num_iter=16, dt=np.float64)
        for q, y32, y64 in zip(rb.agent_outputs(i)[0], q32, q64):
            e = _bar(_np(q), y32, y64)
            worst = [max(worst[0], e[0]), max(worst[1], e[1])]
    print(f"ragged full, {len(x32)} float-slice agents: qs max|x-x32|={worst[0]:.3e} max|x-x64|={worst[1]:.3e}")


def test_ragged_graph_replay_and_autotune(dev):
    w = workloads.ragged(seed=3, device=dev, n_agents=512)
    print("ragged graph: 512 agents", flush=True)
    rb = ragged.RaggedBatch(w["agents"], num_iter=16, device=dev)
    rb.step()
    torch.cuda.synchronize()
    ref_out, ref_act = rb.out.clone(), rb.act.clone()
    rb.autotune()
    cfg = rb.plan.tile_config()
    print("autotuned tiles:", cfg, flush=True)
    assert cfg["autotuned"] == 1
    rb.step()
    torch.cuda.synchronize()
    assert torch.equal(ref_out, rb.out) and torch.equal(ref_act, rb.act)
    g = rb.capture()
    torch.cuda.synchronize()
    for _ in range(3):
        rb.out.zero_()
        rb.act.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(ref_out, rb.out) and torch.equal(ref_act, rb.act)
        rb.step()  # eager work between replays
        torch.cuda.synchronize()
    # new observations through the same graph
    rng = np.random.default_rng(11)
    new_obs = [[int(rng.integers(0, o)) for o in a["O"]] for a in w["agents"]]
    for a, o in zip(w["agents"], new_obs):
        a["obs"] = o
    rb.set_obs(new_obs)
    g.replay()
    torch.cuda.synchronize()
    for i in (0, 100, 511):
        _check_agent(rb, i, use_dedup=True)


def test_ragged_rollout_carries_prior(dev):
    """Carry mode: three steps with fresh observations, each FPI starting from
    the previous step's predicted prior B . q of the chosen action (the first
    from D via reset_prior), eagerly and through a captured graph."""
    w = workloads.ragged(seed=3, device=dev, n_agents=48, state_range=(2, 12), obs_range=(2, 9), max_policies=128,
                         max_A_bytes=1 << 16)
    ags = w["agents"]
    print("ragged rollout: 48 agents, 3 steps", flush=True)
    rb = ragged.RaggedBatch(ags, num_iter=16, device=dev)
    rb.set_prior_source(True)
    rng = np.random.default_rng(7)
    obs_seq = [[[int(rng.integers(0, o)) for o in a["O"]] for a in ags] for _ in range(3)]
    rb.reset_prior()
    eager = []
    for t in range(3):
        rb.set_obs(obs_seq[t])
        rb.step()
        torch.cuda.synchronize()
        eager.append((rb.out.clone(), rb.act.clone()))
    assert rb.status() == 0
    worst = 0.0
    for i, a in enumerate(ags):
        ref = _oracle_rollout(a, [obs_seq[t][i] for t in range(3)])
        for t, (q64, act, pr) in enumerate(ref):
            rb.out.copy_(eager[t][0])
            rb.act.copy_(eager[t][1])
            qs, _, _, ac = rb.agent_outputs(i)
            for x, y in zip(qs, q64):
                worst = max(worst, float(np.abs(_np(x) - y).max()))
            np.testing.assert_array_equal(ac.cpu().numpy(), act)
            for x, y in zip(rb.agent_prior(i), pr):
                worst = max(worst, float(np.abs(_np(x) - y).max()))
    assert worst <= TOL, worst
    print(f"ragged rollout max|x-x64|={worst:.3e}")
    # the same rollout through one captured graph, replayed per step
    g = rb.capture()
    torch.cuda.synchronize()
    rb.reset_prior()
    for t in range(3):
        rb.set_obs(obs_seq[t])
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(eager[t][0], rb.out) and torch.equal(eager[t][1], rb.act), t


def test_ragged_tilings_bitwise_equal(dev):
    """Every EFE tiling in the autotuner's range gives bitwise the same
    step (a tile only partitions rows and history columns; each sum keeps its
    order)."""
    w = workloads.ragged(seed=4, device=dev, n_agents=256)
    print("ragged tilings: 256 agents", flush=True)
    rb = ragged.RaggedBatch(w["agents"], num_iter=16, device=dev)
    rb.step()
    torch.cuda.synchronize()
    ref_out, ref_act = rb.out.clone(), rb.act.clone()
    for cfg in [(128, 128), (256, 256), (128, 256), (256, 128)]:
        rb.plan.set_tiles(*cfg)
        rb.out.zero_()
        rb.act.zero_()
        rb.step()
        torch.cuda.synchronize()
        assert torch.equal(ref_out, rb.out) and torch.equal(ref_act, rb.act), cfg
    with pytest.raises(ValueError, match="tile sizes out of range"):
        rb.plan.set_tiles(256, 200)
