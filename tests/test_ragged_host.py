"""Host-side checks of the batched ragged plan (no GPU): the plan's tables
replayed by the kernel emulation in tests/ragged_emu.py reproduce the
oracle per agent, every access of the full 4096-agent plan stays in bounds,
and invalid agents are rejected with a "ragged:" ValueError."""
import numpy as np
import pytest

from oracle import dedup_ref as DR
from oracle import numpy_ref as R
from pymdp_amd import ragged, workloads
from tests import ragged_emu as E


def _emulate(agents):
    plan = ragged.RaggedPlan(agents, 16)
    params = ragged.pack_params(plan, agents, "cpu").numpy()
    obs = ragged.pack_obs(plan, [a["obs"] for a in agents], "cpu").numpy()
    return plan, E.Emu(plan, params, obs).run()


@pytest.mark.parametrize("kw", [
    dict(seed=1, n_agents=24, state_range=(2, 9), obs_range=(2, 9), max_policies=256, max_A_bytes=1 << 16),
    dict(seed=5, n_agents=10, max_policies=64),
])
def test_emulated_step_matches_oracle(kw):
    w = workloads.ragged(**kw)
    ags = w["agents"]
    plan, em = _emulate(ags)
    lay = plan.layout
    for i, a in enumerate(ags):
        A = [x.numpy() for x in a["A"]]
        B = [x.numpy() for x in a["B"]]
        C = [x.numpy() for x in a["C"]]
        D = [x.numpy() for x in a["D"]]
        ob = [np.eye(o, dtype=np.float32)[x] for o, x in zip(a["O"], a["obs"])]
        q64 = R.run_factorized_fpi(A, ob, D, a["A_dependencies"], num_iter=16, dt=np.float64)
        S = a["S"]
        qs = np.split(em.out[lay["qs_off"][i]:lay["qs_off"][i] + sum(S)], np.cumsum(S)[:-1])
        for x, y in zip(qs, q64):
            assert np.abs(x - y).max() <= 1e-6
        pm = a["policies"]
        P = pm.shape[0]
        G64 = DR.neg_efe_dedup(pm, [x.astype(np.float64) for x in qs], A, B, C, a["A_dependencies"],
                               a["B_dependencies"], dt=np.float64)
        p64 = DR.q_pi_from_G(G64, a["gamma"])
        G = em.out[lay["G_off"][i]:lay["G_off"][i] + P]
        q = em.out[lay["qpi_off"][i]:lay["qpi_off"][i] + P]
        assert np.abs(G - G64).max() <= 2e-6
        assert np.abs(q - p64).max() <= 2e-6
        act = em.act[lay["act_off"][i]:lay["act_off"][i] + len(S)]
        np.testing.assert_array_equal(act, R.sample_policy(pm, p64))


def test_emulation_detects_planted_index_error():
    """The emulation is a real check: swapping the two score strides of a
    rank-2 modality (policies read the wrong history combination) breaks
    agreement with the oracle."""
    w = workloads.ragged(seed=5, n_agents=10, max_policies=64)
    ags = w["agents"]
    plan = ragged.RaggedPlan(ags, 16)
    params = ragged.pack_params(plan, ags, "cpu").numpy()
    obs = ragged.pack_obs(plan, [a["obs"] for a in ags], "cpu").numpy()
    em = E.Emu(plan, params, obs)
    planted = 0
    for a in em.ag:
        for t in range(int(a["T"])):
            for m in range(int(a["M"])):
                g = int(a["mod_group"][m])
                st = a["score_stride"][t][m]
                if a["gR"][g] == 2 and st[0] != st[1]:
                    st[0], st[1] = int(st[1]), int(st[0])
                    planted += 1
    assert planted > 0
    worst = 0.0
    try:
        em.run()
    except AssertionError:
        return  # an access left its buffer: detected
    lay = plan.layout
    for i, a in enumerate(ags):
        qs = np.split(em.out[lay["qs_off"][i]:lay["qs_off"][i] + sum(a["S"])], np.cumsum(a["S"])[:-1])
        G64 = DR.neg_efe_dedup(a["policies"], [x.astype(np.float64) for x in qs], [x.numpy() for x in a["A"]],
                               [x.numpy() for x in a["B"]], [x.numpy() for x in a["C"]], a["A_dependencies"],
                               a["B_dependencies"], dt=np.float64)
        P = a["policies"].shape[0]
        worst = max(worst, float(np.abs(em.out[lay["G_off"][i]:lay["G_off"][i] + P] - G64).max()))
    assert worst > 1e-3


def test_full_plan_bounds():
    w = workloads.ragged(seed=0, n_agents=4096, tensors=False)
    plan = ragged.RaggedPlan(w["agents"], 16)
    info = E.check_bounds_only(plan)
    assert info["agents"] == 4096 and info["units"] >= 4096 and info["team_max"] > 1
    ag, units, (t1, t2, t3), iarena = E.tables(plan)
    # FPI units: teams first (largest first), every agent present; the EFE
    # stage tiles cover every modality of every agent at every level
    teams = [int(u["agent"]) for u in units if ag[int(u["agent"])]["team"] > 1]
    assert teams == [int(u["agent"]) for u in units[:len(teams)]]
    assert sorted(set(int(u["agent"]) for u in units)) == list(range(4096))
    per = np.zeros((4096, 8), np.int64)
    for tl in t1:
        per[tl["agent"], tl["m"]] += 1
    for i, a in enumerate(w["agents"]):
        assert (per[i, :len(a["O"])] >= 1).all() and (per[i, len(a["O"]):] == 0).all()
    lv = np.zeros((4096, 8, 4), np.int64)
    for tl in t3:
        lv[tl["agent"], tl["m"], tl["t"]] += 1
    for i, a in enumerate(w["agents"]):
        assert (lv[i, :len(a["O"]), :a["T"]] >= 1).all()


def _one_agent(**over):
    a = dict(S=[3, 2], U=[2, 1], O=[4], T=1, A=None, B=None, C=None, D=None, E=None, obs=[0],
             A_dependencies=[[0, 1]], B_dependencies=[[0], [1]],
             policies=np.array([[[0, 0]], [[1, 0]]], np.int32), gamma=1.0)
    a.update(over)
    return a


@pytest.mark.parametrize("over,msg", [
    (dict(S=[3, 65], U=[2, 1]), "state counts"),
    (dict(O=[65]), "outcome counts"),
    (dict(A_dependencies=[[0, 0]]), "duplicate factor"),
    (dict(A_dependencies=[[0, 2]]), "out of range"),
    (dict(policies=np.array([[[2, 0]]], np.int32)), "out of bounds"),
])
def test_invalid_agents_rejected(over, msg):
    with pytest.raises(ValueError, match="ragged: agent 0: .*" + msg):
        ragged.RaggedPlan([_one_agent(**over)], 16)


def test_sweep_count_limit():
    """The sweep number tags the partial marginals a team exchanges (10
    mantissa bits, one value kept for a fresh buffer): num_iter <= 1022."""
    ragged.RaggedPlan([_one_agent()], 1022)
    with pytest.raises(ValueError, match="num_iter must be <= 1022"):
        ragged.RaggedPlan([_one_agent()], 1023)


def test_plan_layout_is_disjoint():
    w = workloads.ragged(seed=2, n_agents=50, tensors=False)
    plan = ragged.RaggedPlan(w["agents"], 16)
    lay = plan.layout
    spans = []
    mi = fi = 0
    for a in w["agents"]:
        for m, o in enumerate(a["O"]):
            K = int(np.prod([a["S"][d] for d in a["A_dependencies"][m]]))
            spans.append((int(lay["A_off"][mi]), o * K))
            spans.append((int(lay["C_off"][mi]), o))
            mi += 1
        for f, s in enumerate(a["S"]):
            spans.append((int(lay["B_off"][fi]), s * s * a["U"][f]))
            spans.append((int(lay["D_off"][fi]), s))
            fi += 1
    spans.sort()
    for (o0, n0), (o1, _) in zip(spans, spans[1:]):
        assert o0 + n0 <= o1
    assert spans[-1][0] + spans[-1][1] <= int(plan.sizes[0])


def test_emulated_teams_match_oracle():
    """Agents whose log-likelihood spans several FPI units (a team exchanging
    partial marginals every sweep, the log-likelihood held as floats), with
    rank-1, -2 and -3 modalities."""
    w = workloads.ragged(seed=0, n_agents=5, state_range=(20, 40), obs_range=(2, 6), max_policies=64,
                         max_A_bytes=1 << 22, max_modalities=3)
    ags = w["agents"]
    plan, em = _emulate(ags)
    teams = [int(a["team"]) for a in em.ag]
    assert max(teams) > 2, teams
    assert {int(u["x32"]) for u in em.units} == {0, 1}
    ranks = {len(d) for a in ags for d in a["A_dependencies"]}
    assert {1, 2, 3} <= ranks, ranks
    lay = plan.layout
    for i, a in enumerate(ags):
        A = [x.numpy() for x in a["A"]]
        ob = [np.eye(o, dtype=np.float32)[x] for o, x in zip(a["O"], a["obs"])]
        q64 = R.run_factorized_fpi(A, ob, [x.numpy() for x in a["D"]], a["A_dependencies"], num_iter=16,
                                   dt=np.float64)
        S = a["S"]
        qs = np.split(em.out[lay["qs_off"][i]:lay["qs_off"][i] + sum(S)], np.cumsum(S)[:-1])
        for x, y in zip(qs, q64):
            assert np.abs(x - y).max() <= 1e-6
        pm = a["policies"]
        G64 = DR.neg_efe_dedup(pm, [x.astype(np.float64) for x in qs], A, [x.numpy() for x in a["B"]],
                               [x.numpy() for x in a["C"]], a["A_dependencies"], a["B_dependencies"], dt=np.float64)
        G = em.out[lay["G_off"][i]:lay["G_off"][i] + pm.shape[0]]
        assert np.abs(G - G64).max() <= 2e-6


def _oracle_rollout(a, obs_seq):
    """FP64 oracle of a carry-mode rollout (pymdp/envs/rollout.py:65-82):
    prior_0 = D, qs_t = FPI(obs_t, prior_t), action_t = first action of the
    argmax policy, prior_{t+1} = B . qs_t for action_t."""
    def npy(x):
        return x.detach().cpu().numpy()
    A = [npy(x) for x in a["A"]]
    B = [npy(x).astype(np.float64) for x in a["B"]]
    C = [npy(x) for x in a["C"]]
    prior = [npy(x) for x in a["D"]]
    out = []
    for obs in obs_seq:
        ob = [np.eye(o, dtype=np.float32)[x] for o, x in zip(a["O"], obs)]
        q64 = R.run_factorized_fpi(A, ob, prior, a["A_dependencies"], num_iter=16, dt=np.float64)
        G64 = DR.neg_efe_dedup(a["policies"], q64, A, B, C, a["A_dependencies"], a["B_dependencies"],
                               dt=np.float64)
        act = R.sample_policy(a["policies"], DR.q_pi_from_G(G64, a["gamma"]))
        prior = R.compute_expected_state(q64, B, act, a["B_dependencies"])
        out.append((q64, act, prior))
    return out


def test_emulated_rollout_carries_prior():
    """Three carry-mode steps with fresh observations: each step's FPI starts
    from the previous step's predicted prior B . q (the first from D, copied
    into the prior slots as RaggedBatch.reset_prior does), and the step writes
    the next one."""
    w = workloads.ragged(seed=3, n_agents=12, state_range=(2, 12), obs_range=(2, 9), max_policies=128,
                         max_A_bytes=1 << 16)
    ags = w["agents"]
    plan = ragged.RaggedPlan(ags, 16)
    lay = plan.layout
    params = ragged.pack_params(plan, ags, "cpu").numpy()
    rng = np.random.default_rng(7)
    obs_seq = [[[int(rng.integers(0, o)) for o in a["O"]] for a in ags] for _ in range(3)]
    out = np.zeros(int(plan.sizes[2]), np.float32)
    fi = 0
    for i, a in enumerate(ags):
        po = int(lay["pred_off"][i])
        for f, s in enumerate(a["S"]):
            d = int(lay["D_off"][fi])
            out[po:po + s] = params[d:d + s]
            po += s
            fi += 1
    ems = []
    for t in range(3):
        obs = ragged.pack_obs(plan, obs_seq[t], "cpu").numpy()
        em = E.Emu(plan, params, obs, carry=True, out=out).run()
        out = em.out
        ems.append(em)
    for i, a in enumerate(ags):
        ref = _oracle_rollout(a, [obs_seq[t][i] for t in range(3)])
        S = a["S"]
        for t, (q64, act, pr) in enumerate(ref):
            em = ems[t]
            qo = int(lay["qs_off"][i])
            for x, y in zip(np.split(em.out[qo:qo + sum(S)], np.cumsum(S)[:-1]), q64):
                assert np.abs(x - y).max() <= 1e-6, (i, t)
            np.testing.assert_array_equal(em.act[lay["act_off"][i]:lay["act_off"][i] + len(S)], act)
            po = int(lay["pred_off"][i])
            for x, y in zip(np.split(em.out[po:po + sum(S)], np.cumsum(S)[:-1]), pr):
                assert np.abs(x - y).max() <= 1e-6, (i, t)


@pytest.mark.parametrize("cfg", [(128, 128), (256, 256), (128, 256)])
def test_tilings_bounds_and_emulation(cfg):
    """Every tiling of the autotuner's range keeps every access of the full
    4096-agent plan in bounds, and the emulated step is bitwise the same as
    with the default tiles."""
    w = workloads.ragged(seed=0, n_agents=4096, tensors=False)
    plan = ragged.RaggedPlan(w["agents"], 16)
    n0 = plan.info()["tiles1"]
    plan.set_tiles(*cfg)
    assert plan.tile_config()["e1_rows"] == cfg[0]
    E.check_bounds_only(plan)
    if cfg[0] < 256:
        assert plan.info()["tiles1"] > n0
    wk = dict(seed=1, n_agents=24, state_range=(2, 9), obs_range=(2, 9), max_policies=256, max_A_bytes=1 << 16)
    ags = workloads.ragged(**wk)["agents"]
    ref = _emulate(ags)[1]
    plan2 = ragged.RaggedPlan(ags, 16)
    plan2.set_tiles(*cfg)
    params = ragged.pack_params(plan2, ags, "cpu").numpy()
    obs = ragged.pack_obs(plan2, [a["obs"] for a in ags], "cpu").numpy()
    em = E.Emu(plan2, params, obs).run()
    np.testing.assert_array_equal(em.out, ref.out)
    np.testing.assert_array_equal(em.act, ref.act)
