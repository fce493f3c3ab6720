"""GPU parity of the fused FPI kernel against the oracle.

Bar (per agent, see _check): |q - q_fp64| <= 1e-5 and
|q - q_fp32| <= 1e-5 + |q_fp32 - q_fp64|, where q_fp32 is the FP32
restatement of the reference JAX path and q_fp64 the FP64 oracle.
Full-size runs are additionally checked through size-independent properties
(normalisation, finiteness, determinism).
"""
import numpy as np
import pytest
import torch

from oracle import numpy_ref as R
from pymdp_amd import algos, workloads
from pymdp_amd.ffi import fpi_ffi, infer_states_ffi

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _np(t):
    return t.detach().float().cpu().numpy()


def _oracle_agent(A, obs_idx, D, deps, b, num_iter, dt, obs_distr=None):
    A_b = [_np(a[b]) if a.dim() == len(d) + 2 else _np(a) for a, d in zip(A, deps)]
    if obs_distr is None:
        obs = [np.eye(a.shape[0], dtype=np.float32)[int(o[b])] for a, o in zip(A_b, obs_idx)]
    else:
        obs = [_np(o[b]) for o in obs_distr]
    prior = [_np(d[b]) for d in D]
    return R.run_factorized_fpi(A_b, obs, prior, deps, num_iter=num_iter, dt=dt)


def _check(qs, A, obs, D, deps, agents, num_iter, obs_distr=None, tol=TOL):
    """Per agent: e32 = |gpu - fp32 oracle|, e64 = |gpu - fp64 oracle|,
    r = |fp32 oracle - fp64 oracle| (the FP32 reference path's own error).

    The mean-field sweep oscillates for some agents (period-2 cycles of the
    Jacobi update, e.g. q after 15 and 16 sweeps differ by O(0.1)); there
    FP32 rounding is amplified sweep by sweep and two FP32 implementations
    legitimately disagree by more than 1e-5.  The kernel accumulates the
    sweeps in FP64, so the bar is: e64 <= 1e-5 for every agent, and
    e32 <= 1e-5 + r (within 1e-5 of the FP32 reference wherever the
    reference itself is within r of the FP64 result)."""
    worst32 = worst64 = worst_r = 0.0
    for b in agents:
        o32 = _oracle_agent(A, obs, D, deps, b, num_iter, np.float32, obs_distr)
        o64 = _oracle_agent(A, obs, D, deps, b, num_iter, np.float64, obs_distr)
        e32 = e64 = r = 0.0
        for f, q in enumerate(qs):
            g = _np(q[b]).astype(np.float64)
            e32 = max(e32, float(np.abs(g - o32[f]).max()))
            e64 = max(e64, float(np.abs(g - o64[f]).max()))
            r = max(r, float(np.abs(o32[f] - o64[f]).max()))
        assert e64 <= tol, (b, e64)
        assert e32 <= tol + r, (b, e32, r)
        worst32, worst64, worst_r = max(worst32, e32), max(worst64, e64), max(worst_r, r)
    print(f"max|q-q_fp32|={worst32:.3e} max|q-q_fp64|={worst64:.3e} max|q_fp32-q_fp64|={worst_r:.3e}")
    return worst32, worst64


def test_plain_fpi_small_parity(dev):
    w = workloads.plain_fpi_small(seed=0, device=dev)
    print("plain_fpi_small", w["S"], w["O"], w["batch"], flush=True)
    qs = infer_states_ffi(w["A"], w["obs"], w["D"], w["A_dependencies"], w["num_iter"], distr_obs=False)
    torch.cuda.synchronize()
    _check(qs, w["A"], w["obs"], w["D"], w["A_dependencies"], range(w["batch"]), w["num_iter"])


def test_plain_fpi_large_parity(dev):
    w = workloads.plain_fpi_large(seed=0, device=dev)
    print("plain_fpi_large", w["S"], w["O"], w["batch"], flush=True)
    qs = infer_states_ffi(w["A"], w["obs"], w["D"], w["A_dependencies"], w["num_iter"], distr_obs=False)
    torch.cuda.synchronize()
    for q in qs:
        assert torch.isfinite(q).all()
        assert torch.allclose(q.sum(-1), torch.ones_like(q[:, 0]), atol=1e-5)
    _check(qs, w["A"], w["obs"], w["D"], w["A_dependencies"], [0, 1, 511, 1023], w["num_iter"])
    # the FP32 spread of the reference path on agents 0, 512 and 1023: a
    # second FP32 restatement with another summation order
    from tests.fp32_alt import run_factorized_fpi_alt32
    for b in (0, 512, 1023):
        A_b = [_np(a[b]) for a in w["A"]]
        ob = [np.eye(a.shape[0], dtype=np.float32)[int(o[b])] for a, o in zip(A_b, w["obs"])]
        pb = [_np(d[b]) for d in w["D"]]
        deps = w["A_dependencies"]
        o32 = R.run_factorized_fpi(A_b, ob, pb, deps, num_iter=16, dt=np.float32)
        o64 = R.run_factorized_fpi(A_b, ob, pb, deps, num_iter=16, dt=np.float64)
        alt = run_factorized_fpi_alt32(A_b, ob, pb, deps, 16)
        rnd = run_factorized_fpi_alt32(A_b, ob, pb, deps, 16, ll_rounding="rounded64")
        d = lambda x, y: max(float(np.abs(np.asarray(u, np.float64) - np.asarray(v, np.float64)).max())
                             for u, v in zip(x, y))
        gq = [_np(q[b]) for q in qs]
        print(f"agent {b}: |gpu-fp64|={d(gq, o64):.2e} |gpu-fp32|={d(gq, o32):.2e} "
              f"|gpu-fp32alt|={d(gq, alt):.2e} |fp32-fp32alt|={d(o32, alt):.2e} |fp32-fp64|={d(o32, o64):.2e} "
              f"|fp32-fp32(ln A rounded from fp64)|={d(o32, rnd):.2e}", flush=True)
        assert d(gq, o64) <= TOL
    # determinism: a second launch gives bitwise the same result
    qs2 = infer_states_ffi(w["A"], w["obs"], w["D"], w["A_dependencies"], w["num_iter"], distr_obs=False)
    for a, b in zip(qs, qs2):
        assert torch.equal(a, b)


RAND_CASES = [
    ([3], [3, 2], [[0], [0]]),
    ([4, 3], [5, 6], [[0, 1], [0]]),
    ([3, 2, 4], [3, 4, 2], [[0, 2], [1], [2, 1, 0]]),
    ([5, 7, 3], [4, 4, 9], [[0, 1, 2], [1, 2], [2]]),
    ([3, 2, 2, 2], [3, 2], [[0, 1, 2, 3], [3, 0]]),
    ([64, 33], [17], [[0, 1]]),
    ([2, 3, 2, 3, 2], [3, 3], [[0, 1, 2, 3, 4], [4, 2]]),
    ([7, 9, 11], [5, 6], [[2, 0, 1], [1]]),
]


@pytest.mark.parametrize("case", range(len(RAND_CASES)))
def test_random_dependency_parity(dev, case):
    S, O, deps = RAND_CASES[case]
    print("case", case, S, O, deps, flush=True)
    g = torch.Generator(device=dev).manual_seed(10 + case)
    from pymdp_amd import utils
    batch = 16
    A = utils.random_A_array(O, S, deps, batch=batch, alpha=0.5, generator=g, device=dev)
    D = utils.random_factorized_categorical(S, batch=batch, alpha=1.0, generator=g, device=dev)
    obs = utils.sample_observations(A, D, deps, g)
    for num_iter in (1, 16):
        qs = infer_states_ffi(A, obs, D, deps, num_iter, distr_obs=False)
        torch.cuda.synchronize()
        _check(qs, A, obs, D, deps, range(batch), num_iter)


def test_distributional_obs_parity(dev):
    from pymdp_amd import utils
    S, O, deps = [4, 5, 3], [6, 3], [[0, 1, 2], [1, 2]]
    g = torch.Generator(device=dev).manual_seed(3)
    batch = 32
    A = utils.random_A_array(O, S, deps, batch=batch, alpha=0.5, generator=g, device=dev)
    D = utils.random_factorized_categorical(S, batch=batch, generator=g, device=dev)
    obs = utils.random_factorized_categorical(O, batch=batch, generator=g, device=dev)
    qs = infer_states_ffi(A, obs, D, deps, 16, distr_obs=True)
    torch.cuda.synchronize()
    _check(qs, A, None, D, deps, range(batch), 16, obs_distr=obs)


def test_index_equals_onehot_and_shared_A(dev):
    w = workloads.plain_fpi_small(seed=1, device=dev)
    A, D, deps, obs = w["A"], w["D"], w["A_dependencies"], w["obs"]
    q_idx = infer_states_ffi(A, obs, D, deps, 16, distr_obs=False)
    onehot = [torch.nn.functional.one_hot(o, a.shape[1]).float() for o, a in zip(obs, A)]
    q_oh = infer_states_ffi(A, onehot, D, deps, 16, distr_obs=True)
    for a, b in zip(q_idx, q_oh):
        assert torch.equal(a, b)
    # an A shared by all agents (batch stride 0) == the broadcast copy
    A0 = [a[0].contiguous() for a in A]
    Ab = [a0[None].expand_as(a).contiguous() for a0, a in zip(A0, A)]
    q_sh = infer_states_ffi(A0, obs, D, deps, 16, distr_obs=False)
    q_bc = infer_states_ffi(Ab, obs, D, deps, 16, distr_obs=False)
    for a, b in zip(q_sh, q_bc):
        assert torch.equal(a, b)


def test_fpi_ffi_from_loglikelihood(dev):
    """The reference-ABI entry (pymdp_hip_fpi): FPI from given ll / lp."""
    rng = np.random.default_rng(0)
    S, O, deps = [4, 3, 5], [5, 6, 2], [[0, 1], [2, 0], [1, 2, 0]]
    batch = 8
    ll = [rng.normal(size=(batch,) + tuple(S[d] for d in dl)).astype(np.float32) * 3 for dl in deps]
    lp = [np.log(rng.dirichlet(np.ones(s), size=batch)).astype(np.float32) for s in S]
    q = fpi_ffi([torch.tensor(x, device=dev) for x in ll], [torch.tensor(x, device=dev) for x in lp], deps, 16)
    torch.cuda.synchronize()
    for b in range(batch):
        o32 = R.fpi_from_ll([x[b] for x in ll], [x[b] for x in lp], deps, 16, np.float32)
        o64 = R.fpi_from_ll([x[b] for x in ll], [x[b] for x in lp], deps, 16, np.float64)
        for f in range(len(S)):
            r = np.abs(o32[f] - o64[f]).max()
            assert np.abs(_np(q[f][b]) - o64[f]).max() <= TOL
            assert np.abs(_np(q[f][b]) - o32[f]).max() <= TOL + r


def test_iteration_count_semantics(dev):
    w = workloads.plain_fpi_small(seed=2, device=dev, batch=8)
    A, D, deps, obs = w["A"], w["D"], w["A_dependencies"], w["obs"]
    q0 = infer_states_ffi(A, obs, D, deps, 0, distr_obs=False)
    for q in q0:
        assert torch.allclose(q, torch.full_like(q, 1.0 / q.shape[-1]))
    qe = infer_states_ffi(A, obs, D, deps, 16, distr_obs=False, method_exact=True)
    q1 = infer_states_ffi(A, obs, D, deps, 1, distr_obs=False)
    for a, b in zip(qe, q1):
        assert torch.equal(a, b)


def test_algos_unbatched_entry(dev):
    w = workloads.plain_fpi_small(seed=4, device=dev, batch=2)
    A1 = [a[0].contiguous() for a in w["A"]]
    obs1 = [torch.nn.functional.one_hot(o[0], a.shape[0]).float() for o, a in zip(w["obs"], A1)]
    D1 = [d[0].contiguous() for d in w["D"]]
    q = algos.run_factorized_fpi(A1, obs1, D1, w["A_dependencies"], num_iter=16)
    ref = R.run_factorized_fpi([_np(a) for a in A1], [_np(o) for o in obs1], [_np(d) for d in D1],
                               w["A_dependencies"], num_iter=16, dt=np.float32)
    for a, b in zip(q, ref):
        assert np.abs(_np(a) - b).max() <= TOL
    qv = algos.run_vanilla_fpi(A1, obs1, D1, num_iter=16)
    for a, b in zip(qv, q):
        assert torch.allclose(a, b, atol=1e-6)


def test_fast_kernel_timing_trace(dev):
    """The fast kernel's diagnostic trace leaves the posterior bitwise
    unchanged and records, per agent, non-decreasing stamps: start,
    log-likelihood rows formed, the end of each sweep, end."""
    import ctypes

    from pymdp_amd import _native
    w = workloads.plain_fpi_large(seed=2, device=dev, batch=64)
    B, it = w["batch"], w["num_iter"]
    obs = torch.stack([o.to(torch.int32) for o in w["obs"]], -1).contiguous()
    ref = [q.clone() for q in infer_states_ffi(w["A"], obs, w["D"], w["A_dependencies"], it, distr_obs=False)]
    tr = torch.zeros(B * (it + 3), dtype=torch.int64, device=dev)
    L = _native.lib()
    _native.check(L.pymdp_hip_fpi_set_trace(ctypes.c_void_p(tr.data_ptr())))
    try:
        got = infer_states_ffi(w["A"], obs, w["D"], w["A_dependencies"], it, distr_obs=False)
        torch.cuda.synchronize()
    finally:
        _native.check(L.pymdp_hip_fpi_set_trace(None))
    assert "fpi_fast_kernel" in L.pymdp_hip_last_launch().decode()
    for a, b in zip(ref, got):
        assert torch.equal(a, b)
    t = tr.cpu().numpy().reshape(B, it + 3)
    assert (t > 0).all() and (np.diff(t, axis=1) >= 0).all()
