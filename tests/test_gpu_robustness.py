"""GPU tests of the numeric edge cases and of the boundary's robustness.

* Many modalities on one dependency tuple with deterministic (0/1) A and
  contradictory observations: every joint state has 20 or more clipped
  (MINVAL) likelihood terms, so a likelihood product formed in one piece
  would underflow FP64 (MINVAL^20 = 2^-1040).  Checked against the FP64
  oracle at 1e-5.
* Exact zeros and sub-MINVAL entries planted in A (FPI, EFE) and B (EFE):
  the clip-then-log of pymdp/maths.py:64-79 and the xlogx clip of
  :17-30 on every path, against the FP64 oracle at 1e-5.
* Host validation of E, the priors and the observations (ValueError before
  any device work).
* Scratch ownership: a graph captured from the plain entry points replays
  correctly after a larger call has grown the scratch; two streams run the
  entry points concurrently; PYMDP_HIP_FLAG_MODEL_PARAMS_STATIC reuses the
  A-derived tables only while the same tensors, unmodified, are passed.
"""
import numpy as np
import pytest
import torch

from oracle import numpy_ref as R
from pymdp_amd import utils
from pymdp_amd.ffi import _efe, infer_states_ffi

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _np(t):
    return t.detach().float().cpu().numpy()


def _contradictory_model(M, per_state, dev):
    """S = [4, 3] (12 joint states), M binary modalities on deps [0, 1].
    A_m[1, s] = 1 iff joint state s belongs to modality m's set, else
    A_m[0, s] = 1; every observation is 1.  State s belongs to the first
    per_state[s] sets, so it contradicts M - per_state[s] observations."""
    S = [4, 3]
    deps = [[0, 1] for _ in range(M)]
    A = []
    for m in range(M):
        a = torch.zeros((2, 12), dtype=torch.float32)
        for s in range(12):
            a[1 if m < per_state[s] else 0, s] = 1.0
        A.append(a.reshape(2, 4, 3).to(dev))
    obs = [torch.ones(1, dtype=torch.int64, device=dev) for _ in range(M)]
    D = [torch.full((1, s), 1.0 / s, device=dev) for s in S]
    return S, deps, A, obs, D


@pytest.mark.parametrize("M,step", [(24, 1), (32, 3)])
def test_many_modalities_clipped_likelihood(dev, M, step):
    per_state = [(s % 5) * step for s in range(12)]
    assert M - max(per_state) >= 20  # every state has >= 20 clipped terms
    S, deps, A, obs, D = _contradictory_model(M, per_state, dev)
    print("modalities", M, "members per state", per_state, flush=True)
    for num_iter in (1, 16):
        qs = infer_states_ffi(A, obs, D, deps, num_iter, distr_obs=False, batch=1)
        torch.cuda.synchronize()
        A_np = [_np(a) for a in A]
        ob = [np.array([0.0, 1.0], dtype=np.float32) for _ in range(M)]
        D_np = [_np(d[0]) for d in D]
        q64 = R.run_factorized_fpi(A_np, ob, D_np, deps, num_iter=num_iter, dt=np.float64)
        for f in range(2):
            err = float(np.abs(_np(qs[f][0]).astype(np.float64) - q64[f]).max())
            assert err <= TOL, (M, num_iter, f, err, _np(qs[f][0]), q64[f])


def _plant_clips(x, dim, g, zero_frac=0.2, tiny_frac=0.1):
    """Sets ~zero_frac of the entries to exactly 0 and ~tiny_frac to 1e-20
    (below MINVAL = 2.2e-16), then renormalises along `dim` (zeros stay
    zero, tiny entries stay tiny)."""
    u = torch.rand(x.shape, generator=g, device=x.device)
    x = torch.where(u < zero_frac, torch.zeros_like(x), x)
    x = torch.where((u >= zero_frac) & (u < zero_frac + tiny_frac), torch.full_like(x, 1e-20), x)
    s = x.sum(dim=dim, keepdim=True)
    x = torch.where(s > 0, x / s.clamp_min(1e-30), torch.full_like(x, 1.0 / x.shape[dim]))
    return x.contiguous()


def test_fpi_planted_zeros_and_sub_minval(dev):
    g = torch.Generator(device=dev).manual_seed(7)
    S, O, deps = [5, 4, 3], [6, 5, 4], [[0, 1, 2], [1, 2], [0]]
    batch = 24
    A = utils.random_A_array(O, S, deps, batch=batch, alpha=0.5, generator=g, device=dev)
    A = [_plant_clips(a, 1, g) for a in A]
    n_zero = sum(int((a == 0).sum()) for a in A)
    n_tiny = sum(int(((a > 0) & (a < 2.220446e-16)).sum()) for a in A)
    assert n_zero > 0 and n_tiny > 0
    D = utils.random_factorized_categorical(S, batch=batch, alpha=1.0, generator=g, device=dev)
    D = [_plant_clips(d, 1, g, 0.2, 0.1) for d in D]
    obs_idx = [torch.randint(0, o, (batch,), generator=g, device=dev) for o in O]
    obs_dist = [_plant_clips(utils.dirichlet((batch, o), 1.0, 1, g, dev), 1, g) for o in O]
    for distr, obs in ((False, obs_idx), (True, obs_dist)):
        qs = infer_states_ffi(A, obs, D, deps, 16, distr_obs=distr)
        torch.cuda.synchronize()
        worst = 0.0
        for b in range(batch):
            Ab = [_np(a[b]) for a in A]
            if distr:
                ob = [_np(o[b]) for o in obs]
            else:
                ob = [np.eye(o, dtype=np.float32)[int(x[b])] for o, x in zip(O, obs)]
            q64 = R.run_factorized_fpi(Ab, ob, [_np(d[b]) for d in D], deps, num_iter=16, dt=np.float64)
            for f in range(3):
                worst = max(worst, float(np.abs(_np(qs[f][b]).astype(np.float64) - q64[f]).max()))
        print("distributional" if distr else "index", "max |q - q_fp64| =", worst)
        assert worst <= TOL


def test_efe_planted_zeros_and_sub_minval(dev):
    g = torch.Generator(device=dev).manual_seed(8)
    S, U, O = [4, 3, 2], [3, 2, 2], [5, 4]
    Ad, Bd, T = [[0, 1, 2], [1, 2]], [[0], [1, 0], [2]], 2
    batch = 8
    A = [_plant_clips(a, 1, g) for a in utils.random_A_array(O, S, Ad, batch=batch, alpha=0.5, generator=g,
                                                              device=dev)]
    B = [_plant_clips(b, 1, g) for b in utils.random_B_array(S, U, Bd, batch=batch, alpha=0.5, generator=g,
                                                              device=dev)]
    C = [torch.randn((batch, o), generator=g, device=dev) for o in O]
    qs = [_plant_clips(q, 1, g, 0.3, 0.1) for q in utils.random_factorized_categorical(S, batch=batch,
                                                                                      generator=g, device=dev)]
    pm = R.construct_policies(S, U, T)
    q_pi, G = _efe.update_posterior_policies_ffi(pm, qs, A, B, C, None, Ad, Bd, gamma=16.0)
    torch.cuda.synchronize()
    wg = wq = 0.0
    for b in range(batch):
        E = np.full(pm.shape[0], 1.0 / pm.shape[0])
        q64, G64 = R.update_posterior_policies(pm, [_np(q[b]) for q in qs], [_np(a[b]) for a in A],
                                               [_np(x[b]) for x in B], [_np(c[b]) for c in C], E, Ad, Bd,
                                               gamma=16.0, dt=np.float64)
        wg = max(wg, float(np.abs(_np(G[b]).astype(np.float64) - G64).max()))
        wq = max(wq, float(np.abs(_np(q_pi[b]).astype(np.float64) - q64).max()))
        assert int(torch.argmax(q_pi[b])) == int(np.argmax(q64))
    print(f"max |G - G_fp64| = {wg:.2e}, max |q_pi - q_pi_fp64| = {wq:.2e}")
    assert wg <= TOL and wq <= TOL


# --------------------------------------------------------------------------
# host validation
# --------------------------------------------------------------------------

def _efe_case(dev, batch=4):
    g = torch.Generator(device=dev).manual_seed(3)
    S, U, O, Ad, Bd, T = [3, 2], [2, 2], [3, 2], [[0, 1], [1]], [[0], [1]], 1
    A = utils.random_A_array(O, S, Ad, batch=batch, alpha=1.0, generator=g, device=dev)
    B = utils.random_B_array(S, U, Bd, batch=batch, alpha=1.0, generator=g, device=dev)
    C = [torch.randn((batch, o), generator=g, device=dev) for o in O]
    qs = utils.random_factorized_categorical(S, batch=batch, generator=g, device=dev)
    pm = R.construct_policies(S, U, T)
    return pm, qs, A, B, C, Ad, Bd


def test_efe_rejects_bad_E(dev):
    pm, qs, A, B, C, Ad, Bd = _efe_case(dev)
    P = pm.shape[0]
    for E in (torch.full((2, P), 1.0 / P, device=dev),       # leading dim neither 1 nor batch
              torch.full((4, P + 1), 1.0 / P, device=dev),   # wrong policy count
              torch.full((4, P), 1.0 / P)):                  # host tensor
        with pytest.raises(ValueError):
            _efe.update_posterior_policies_ffi(pm, qs, A, B, C, E, Ad, Bd)
    # [P] and [1, P] are one E for every agent and give the same result as [batch, P]
    E1 = torch.softmax(torch.randn(P, device=dev), 0)
    outs = [_efe.update_posterior_policies_ffi(pm, qs, A, B, C, e, Ad, Bd)[0]
            for e in (E1, E1[None], E1[None].expand(4, P).contiguous())]
    torch.cuda.synchronize()
    assert torch.equal(outs[0], outs[1]) and torch.equal(outs[0], outs[2])


def test_efe_rejects_mismatched_qs(dev):
    pm, qs, A, B, C, Ad, Bd = _efe_case(dev)
    with pytest.raises(ValueError):
        _efe.update_posterior_policies_ffi(pm, [qs[0], qs[1][:2]], A, B, C, None, Ad, Bd)
    with pytest.raises(ValueError):
        _efe.update_posterior_policies_ffi(pm, [qs[0], qs[1].cpu()], A, B, C, None, Ad, Bd)


def test_fpi_rejects_bad_prior_and_obs(dev):
    g = torch.Generator(device=dev).manual_seed(4)
    S, O, deps, batch = [3, 2], [3, 4], [[0, 1], [1]], 4
    A = utils.random_A_array(O, S, deps, batch=batch, alpha=1.0, generator=g, device=dev)
    D = utils.random_factorized_categorical(S, batch=batch, generator=g, device=dev)
    obs = [torch.randint(0, o, (batch,), generator=g, device=dev) for o in O]
    with pytest.raises(ValueError):   # prior batch 2 != 4
        infer_states_ffi(A, obs, [D[0], D[1][:2]], deps, 4, distr_obs=False)
    with pytest.raises(ValueError):   # host observations
        infer_states_ffi(A, [obs[0].cpu(), obs[1]], D, deps, 4, distr_obs=False)
    dist = [torch.softmax(torch.randn(batch, o, device=dev), -1) for o in O]
    with pytest.raises(ValueError):   # distribution width != O_m
        infer_states_ffi(A, [dist[0][:, :2], dist[1]], D, deps, 4, distr_obs=True)
    # a [1, S_f] prior is one prior for every agent
    q1 = infer_states_ffi(A, obs, [D[0][:1], D[1][:1]], deps, 4, distr_obs=False)
    qb = infer_states_ffi(A, obs, [D[0][:1].expand(batch, -1).contiguous(), D[1][:1].expand(batch, -1).contiguous()],
                          deps, 4, distr_obs=False)
    torch.cuda.synchronize()
    for a, b in zip(q1, qb):
        assert torch.equal(a, b)


# --------------------------------------------------------------------------
# scratch ownership, streams, capture, static parameters
# --------------------------------------------------------------------------

def _fpi_big_case(dev, batch, seed):
    """Rank-8 dependency tuple (32768 joint states, more than 32 FP64 values
    per lane of a 512-thread workgroup): the streaming FPI variant, which
    keeps its per-agent log-likelihood in the library's scratch."""
    g = torch.Generator(device=dev).manual_seed(seed)
    S, O, deps = [4, 4, 4, 4, 4, 4, 4, 2], [3, 2], [[0, 1, 2, 3, 4, 5, 6, 7], [7, 0]]
    A = utils.random_A_array(O, S, deps, batch=batch, alpha=0.5, generator=g, device=dev)
    D = utils.random_factorized_categorical(S, batch=batch, generator=g, device=dev)
    obs = [torch.randint(0, o, (batch,), generator=g, device=dev) for o in O]
    return A, D, obs, deps


def test_graph_capture_survives_scratch_growth(dev):
    from pymdp_amd import _native
    A, D, obs, deps = _fpi_big_case(dev, 8, 11)
    pm, qs, Ae, Be, Ce, Ad, Bd = _efe_case(dev)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        # eager reference results, and the first call of each shape on the
        # capture stream outside capture (allocates its scratch)
        ref_q = [q.clone() for q in infer_states_ffi(A, obs, D, deps, 8, distr_obs=False)]
        ref_pi, ref_G = _efe.update_posterior_policies_ffi(pm, qs, Ae, Be, Ce, None, Ad, Bd)
        ref_pi, ref_G = ref_pi.clone(), ref_G.clone()
    torch.cuda.synchronize()
    out_q = torch.empty((8, sum(int(d.shape[1]) for d in D)), device=dev)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=s):
        qg = infer_states_ffi(A, obs, D, deps, 8, distr_obs=False, out=out_q)
        pig, Gg = _efe.update_posterior_policies_ffi(pm, qs, Ae, Be, Ce, None, Ad, Bd)
    _, retired0 = _scratch_state(_native)
    # larger calls on the same stream grow (and retire) both scratch buffers
    with torch.cuda.stream(s):
        A2, D2, obs2, _ = _fpi_big_case(dev, 64, 12)
        infer_states_ffi(A2, obs2, D2, deps, 8, distr_obs=False)
        pm2, qs2, Ae2, Be2, Ce2, _, _ = _efe_case(dev, batch=256)
        _efe.update_posterior_policies_ffi(pm2, qs2, Ae2, Be2, Ce2, None, Ad, Bd)
    torch.cuda.synchronize()
    _, retired1 = _scratch_state(_native)
    assert retired1 > retired0, "the larger calls did not grow the scratch"
    for rep in range(3):
        out_q.zero_()
        Gg.zero_()
        graph.replay()
        torch.cuda.synchronize()
        for f, q in enumerate(qg):
            assert torch.equal(q, ref_q[f]), (rep, f)
        assert torch.equal(Gg, ref_G) and torch.equal(pig, ref_pi), rep
        # eager work between replays
        with torch.cuda.stream(s):
            infer_states_ffi(A2, obs2, D2, deps, 8, distr_obs=False)
        torch.cuda.synchronize()


def _scratch_state(native):
    import ctypes
    n = ctypes.c_int64(0)
    live = native.lib().pymdp_hip_scratch_bytes(ctypes.byref(n))
    return live, n.value


def test_two_streams_concurrently(dev):
    cases = [_fpi_big_case(dev, 32, 21), _fpi_big_case(dev, 32, 22)]
    efes = [_efe_case(dev, batch=64), _efe_case(dev, batch=64)]
    ref = []
    for (A, D, obs, deps), (pm, qs, Ae, Be, Ce, Ad, Bd) in zip(cases, efes):
        q = [x.clone() for x in infer_states_ffi(A, obs, D, deps, 8, distr_obs=False)]
        _, G = _efe.update_posterior_policies_ffi(pm, qs, Ae, Be, Ce, None, Ad, Bd)
        ref.append((q, G.clone()))
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    outs = [None, None]
    for rep in range(4):
        for i, st in enumerate(streams):
            with torch.cuda.stream(st):
                A, D, obs, deps = cases[i]
                pm, qs, Ae, Be, Ce, Ad, Bd = efes[i]
                q = infer_states_ffi(A, obs, D, deps, 8, distr_obs=False)
                _, G = _efe.update_posterior_policies_ffi(pm, qs, Ae, Be, Ce, None, Ad, Bd)
                outs[i] = (q, G)
        torch.cuda.synchronize()
        for i in range(2):
            for a, b in zip(outs[i][0], ref[i][0]):
                assert torch.equal(a, b), (rep, i)
            assert torch.equal(outs[i][1], ref[i][1]), (rep, i)


def test_static_params_cache(dev):
    pm, qs, A, B, C, Ad, Bd = _efe_case(dev, batch=16)
    _, G0 = _efe.update_posterior_policies_ffi(pm, qs, A, B, C, None, Ad, Bd)
    G0 = G0.clone()
    for _ in range(2):  # fill, then hit
        _, Gs = _efe.update_posterior_policies_ffi(pm, qs, A, B, C, None, Ad, Bd, ffi_cache_params_static=True)
        torch.cuda.synchronize()
        assert torch.equal(Gs, G0)
    # an in-place change of A (its version counter moves) refills the tables
    A[0].mul_(0.5).add_(0.5 / A[0].shape[1])
    _, G1 = _efe.update_posterior_policies_ffi(pm, qs, A, B, C, None, Ad, Bd)
    _, G1s = _efe.update_posterior_policies_ffi(pm, qs, A, B, C, None, Ad, Bd, ffi_cache_params_static=True)
    torch.cuda.synchronize()
    assert not torch.equal(G1, G0)
    assert torch.equal(G1s, G1)
    # new tensors (possibly at a recycled address) refill as well
    A2 = [a.clone() for a in A]
    A2[1].copy_(torch.softmax(torch.randn_like(A2[1]), 1))
    del A
    _, G2 = _efe.update_posterior_policies_ffi(pm, qs, A2, B, C, None, Ad, Bd)
    _, G2s = _efe.update_posterior_policies_ffi(pm, qs, A2, B, C, None, Ad, Bd, ffi_cache_params_static=True)
    torch.cuda.synchronize()
    assert torch.equal(G2s, G2)
