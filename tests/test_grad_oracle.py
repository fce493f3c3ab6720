"""CPU pins of the gradient oracle and the autograd reference forward:
the pure-torch forward of oracle/autograd_ref.py equals the FP64
NumPy restatement in oracle/grad_ref.py, and its autograd gradient equals
FP64 central differences of that restatement (h = 1e-6)."""
import numpy as np


from oracle import grad_ref as GR
from oracle import autograd_ref as AB
from pymdp_amd import workloads


def test_autograd_reference_matches_oracle_and_central_differences():
    w = workloads.grad_fpi_efe(seed=0, batch=3, T=3)
    keys = ("A_logits", "B_logits", "C_logits", "D_logits")
    params = [[x.double().clone().requires_grad_() for x in w[k]] for k in keys]
    L = AB.trajectory_loss(*params, w["obs"], w["actions"], w["policies"], w["A_dependencies"])
    L.backward()
    nA, nB, nC = (len(w[k]) for k in keys[:3])
    tot = 0.0
    for b in range(3):
        p = [x[b].detach().numpy() for grp in params for x in grp]
        tot += GR.trajectory_loss(p[:nA], p[nA:nA + nB], p[nA + nB:nA + nB + nC], p[nA + nB + nC:],
                                  w["obs"][:, b].numpy(), w["actions"][:, b].numpy(), w["policies"],
                                  w["A_dependencies"], w["B_dependencies"])
    assert abs(float(L.detach()) - tot) <= 1e-9 * abs(tot)
    b = 0
    flat = [x[b].detach().numpy() for grp in params for x in grp]
    grads = [x.grad[b].numpy() for grp in params for x in grp]

    def fn(p):
        return GR.trajectory_loss(p[:nA], p[nA:nA + nB], p[nA + nB:nA + nB + nC], p[nA + nB + nC:],
                                  w["obs"][:, b].numpy(), w["actions"][:, b].numpy(), w["policies"],
                                  w["A_dependencies"], w["B_dependencies"])

    rng = np.random.default_rng(1)
    coords = [(i, tuple(int(rng.integers(0, s)) for s in x.shape)) for i, x in enumerate(flat) for _ in range(2)]
    fd = GR.central_difference(fn, flat, coords)
    an = np.array([grads[i][idx] for i, idx in coords])
    assert np.abs(fd - an).max() <= 1e-7


def test_central_difference_detects_planted_sign_error():
    """A planted mistake in the oracle (x ln x terms of the entropy and the
    ambiguity doubled) moves the loss, so the oracle pins those terms."""
    w = workloads.grad_fpi_efe(seed=0, batch=1, T=2)
    p = [x[0].numpy() for k in ("A_logits", "B_logits", "C_logits", "D_logits") for x in w[k]]
    base = GR.trajectory_loss(p[:3], p[3:5], p[5:8], p[8:], w["obs"][:, 0].numpy(), w["actions"][:, 0].numpy(),
                              w["policies"], w["A_dependencies"], w["B_dependencies"])
    from oracle import numpy_ref as R
    orig = R.stable_xlogx
    try:
        R.stable_xlogx = lambda x, dt=np.float32: 2 * orig(x, dt)
        bad = GR.trajectory_loss(p[:3], p[3:5], p[5:8], p[8:], w["obs"][:, 0].numpy(), w["actions"][:, 0].numpy(),
                                 w["policies"], w["A_dependencies"], w["B_dependencies"])
    finally:
        R.stable_xlogx = orig
    assert abs(bad - base) > 1e-3

