"""GPU checks of the hand-written HIP adjoints (pymdp_hip_fpi_adjoint,
pymdp_hip_neg_efe_adjoint) through the trajectory loss of the
model-fitting workload: gradients w.r.t. the logits of A, B, C and D
against (a) FP64 central differences of the FP64 oracle (h = 1e-6) on
sampled coordinates and (b) torch autograd of the pure-torch reference
forward in FP64 on the host, for every coordinate."""
import numpy as np
import pytest
import torch

from oracle import grad_ref as GR
from oracle import autograd_ref as AB
from pymdp_amd import differentiable as DF
from pymdp_amd import workloads

pytestmark = pytest.mark.gpu


def _grads_hip(w, dev, mode):
    params = [[x.to(dev).clone().requires_grad_() for x in w[k]] for k in ("A_logits", "B_logits", "C_logits",
                                                                          "D_logits")]
    loss = DF.trajectory_loss(*params, w["obs"].to(dev), w["actions"].to(dev), w["policies"], w["A_dependencies"],
                              w["B_dependencies"], num_iter=w["num_iter"], gamma=w["gamma"], mode=mode)
    loss.backward()
    torch.cuda.synchronize()
    return float(loss), [[x.grad.detach().double().cpu() for x in grp] for grp in params]


def _grads_ref64(w):
    params = [[x.detach().double().cpu().clone().requires_grad_() for x in w[k]]
              for k in ("A_logits", "B_logits", "C_logits", "D_logits")]
    loss = AB.trajectory_loss(*params, w["obs"].cpu(), w["actions"].cpu(), w["policies"], w["A_dependencies"],
                              num_iter=w["num_iter"], gamma=w["gamma"])
    loss.backward()
    return float(loss.detach()), [[x.grad.detach() for x in grp] for grp in params]


def _rel(g, r):
    num = max(float((a - b).abs().max()) for ga, gb in zip(g, r) for a, b in zip(ga, gb))
    den = max(float(b.abs().max()) for gb in r for b in gb)
    return num, den


@pytest.mark.parametrize("mode", ["unrolled", "auto"])
def test_adjoint_matches_fp64_autograd(dev, mode):
    w = workloads.grad_fpi_efe(seed=3, device="cpu", batch=8, T=4)
    print("adjoint vs fp64 autograd", mode, flush=True)
    L, g = _grads_hip(w, dev, mode)
    L64, g64 = _grads_ref64(w)
    err, scale = _rel(g, g64)
    print(f"loss {L:.6f} vs {L64:.6f}; max|grad - grad64| = {err:.3e} (max|grad64| = {scale:.3e})")
    assert abs(L - L64) <= 1e-4 * abs(L64)
    assert err <= 1e-4 * scale


def test_adjoint_matches_central_differences(dev):
    w = workloads.grad_fpi_efe(seed=4, device="cpu", batch=2, T=3)
    print("adjoint vs central differences", flush=True)
    _, g = _grads_hip(w, dev, "unrolled")
    b = 1
    keys = ("A_logits", "B_logits", "C_logits", "D_logits")
    params = [x[b].double().numpy() for k in keys for x in w[k]]
    sizes = [len(w[k]) for k in keys]
    flat_g = [x[b].numpy() for grp in g for x in grp]
    nA, nB, nC = sizes[0], sizes[1], sizes[2]

    def fn(p):
        return GR.trajectory_loss(p[:nA], p[nA:nA + nB], p[nA + nB:nA + nB + nC], p[nA + nB + nC:],
                                  w["obs"][:, b].numpy(), w["actions"][:, b].numpy(), w["policies"],
                                  w["A_dependencies"], w["B_dependencies"])

    rng = np.random.default_rng(0)
    coords = [(i, tuple(int(rng.integers(0, s)) for s in x.shape)) for i, x in enumerate(params) for _ in range(3)]
    fd = GR.central_difference(fn, params, coords, h=1e-6)
    an = np.array([flat_g[i][idx] for i, idx in coords])
    scale = max(1e-3, float(np.abs(fd).max()))
    print(f"max|adjoint - fd| = {np.abs(an - fd).max():.3e} (max|fd| = {scale:.3e})")
    assert np.abs(an - fd).max() <= 1e-4 * scale + 1e-7


def test_adjoint_full_size_matches_autograd_baseline(dev):
    """512 agents x 32 steps: HIP adjoints vs torch autograd through the
    pure-torch forward, both FP32 on the GPU."""
    w = workloads.grad_fpi_efe(seed=0, device=dev)
    print("adjoint full size", flush=True)
    L, g = _grads_hip(w, dev, "auto")
    params = [[x.clone().requires_grad_() for x in w[k]] for k in ("A_logits", "B_logits", "C_logits", "D_logits")]
    loss = AB.trajectory_loss(*params, w["obs"], w["actions"], w["policies"], w["A_dependencies"],
                              num_iter=w["num_iter"], gamma=w["gamma"])
    loss.backward()
    gb = [[x.grad.detach().double().cpu() for x in grp] for grp in params]
    err, scale = _rel(g, gb)
    print(f"loss {L:.4f} vs {float(loss):.4f}; max|hip - autograd| = {err:.3e} (max|grad| = {scale:.3e})")
    assert abs(L - float(loss)) <= 1e-4 * abs(float(loss))
    assert err <= 1e-3 * scale
    for grp in g:
        for x in grp:
            assert torch.isfinite(x).all()


def test_implicit_adjoint_on_converged_agents(dev):
    """With many sweeps the forward reaches its fixed point and the implicit
    adjoint is used; it agrees with the unrolled reverse sweep."""
    from pymdp_amd import utils
    g = torch.Generator(device=dev).manual_seed(7)
    S, O, deps = [5, 4], [6, 3], [[0, 1], [1]]
    batch = 32
    A = utils.random_A_array(O, S, deps, batch=batch, alpha=2.0, generator=g, device=dev)
    D = utils.random_factorized_categorical(S, batch=batch, alpha=1.0, generator=g, device=dev)
    obs = torch.stack([torch.randint(0, o, (batch,), generator=g, device=dev) for o in O], -1).to(torch.int32)
    qbar = torch.randn((batch, sum(S)), generator=g, device=dev)
    Ai, pi, used_i = DF.fpi_adjoint(A, obs, D, deps, 200, qbar, mode="auto", implicit_tol=1e-9)
    Au, pu, used_u = DF.fpi_adjoint(A, obs, D, deps, 200, qbar, mode="unrolled")
    torch.cuda.synchronize()
    n_imp = int(used_i.sum())
    print(f"implicit adjoint used for {n_imp}/{batch} agents")
    assert n_imp >= batch // 2 and int(used_u.sum()) == 0
    sel = used_i.bool()
    for a, b in zip(Ai, Au):
        assert torch.allclose(a[sel], b[sel], atol=1e-5, rtol=1e-4)
    for a, b in zip(pi, pu):
        assert torch.allclose(a[sel], b[sel], atol=1e-5, rtol=1e-4)
