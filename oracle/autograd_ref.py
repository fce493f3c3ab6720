"""Pure-torch forward of the trajectory loss of pymdp_amd.differentiable,
differentiated by torch autograd: the comparison baseline for the HIP
adjoints (the counterpart of the reference's with_jax_grad backward, which
differentiates the JAX reference path, pymdp/ffi/_core.py:282-310), timed
on the GPU and on the host cores.  Test and benchmark infrastructure only:
no product entry point calls it.

Same arithmetic as the reference JAX path: FPI of pymdp/algos.py:88-155
(Jacobi sweeps from log_q = 0), neg-EFE of pymdp/control.py:604-690 at
horizon 1, q_pi = softmax(gamma G + ln E) with E uniform.
"""
from __future__ import annotations

import string

import numpy as np
import torch

MINVAL = 2.220446049250313e-16


def _log_stable(x):
    return torch.log(torch.clamp(x, min=MINVAL))


def fpi(A, obs, prior, A_dependencies, num_iter):
    """A[m] [b, O_m, S_deps...], obs int [b, M], prior[f] [b, S_f]."""
    F = len(prior)
    batch = prior[0].shape[0]
    ar = torch.arange(batch, device=prior[0].device)
    ll = [_log_stable(a[ar, obs[:, m].long()]) for m, a in enumerate(A)]
    lp = [_log_stable(p) for p in prior]
    if F == 1:
        num_iter = 1
    log_q = [torch.zeros_like(p) for p in prior]
    for _ in range(num_iter):
        q = [torch.softmax(x, -1) for x in log_q]
        new = [x.clone() for x in lp]
        for m, deps in enumerate(A_dependencies):
            letters = string.ascii_lowercase[1:1 + len(deps)]
            for a, f in enumerate(deps):
                ops = [ll[m]] + [q[d] for j, d in enumerate(deps) if j != a]
                spec = "z" + letters + "," + ",".join("z" + letters[j] for j in range(len(deps)) if j != a)
                spec = spec.rstrip(",") + "->z" + letters[a]
                new[f] = new[f] + torch.einsum(spec, *ops)
        log_q = new
    return [torch.softmax(x, -1) for x in log_q]


def neg_efe(qs, A, B, C, policies, A_dependencies):
    """G [b, P] at horizon 1 with factor-local B[f] [b, S, S, U]."""
    pm = torch.as_tensor(np.asarray(policies)[:, 0, :], device=qs[0].device, dtype=torch.long)
    q1 = [torch.einsum("bsju,bj->bus", B[f], qs[f])[:, pm[:, f], :] for f in range(len(qs))]  # [b, P, S]
    G = 0.0
    for m, deps in enumerate(A_dependencies):
        letters = string.ascii_lowercase[1:1 + len(deps)]
        w = torch.einsum(",".join("zp" + letters[j] for j in range(len(deps))) + "->zp" + letters,
                         *[q1[d] for d in deps])
        qo = torch.einsum("zo" + letters + ",zp" + letters + "->zpo", A[m], w)
        H = -(qo * _log_stable(qo)).sum(-1)
        xl = (A[m] * _log_stable(A[m])).sum(1)
        amb = torch.einsum("z" + letters + ",zp" + letters + "->zp", xl, w)
        util = (qo * C[m][:, None, :]).sum(-1)
        G = G + H + amb + util
    return G


def trajectory_loss(A_logits, B_logits, C_logits, D_logits, obs, actions, policies, A_dependencies,
                    num_iter=16, gamma=1.0):
    A = [torch.softmax(x, dim=1) for x in A_logits]
    B = [torch.softmax(x, dim=1) for x in B_logits]
    C = [torch.log_softmax(x, dim=-1) for x in C_logits]
    prior = [torch.softmax(x, dim=-1) for x in D_logits]
    pm = np.asarray(policies)
    act_f = torch.as_tensor(pm[:, 0, :], device=A[0].device, dtype=torch.long)
    batch = A[0].shape[0]
    ar = torch.arange(batch, device=A[0].device)
    loss = torch.zeros((), device=A[0].device, dtype=A[0].dtype)
    for t in range(obs.shape[0]):
        qs = fpi(A, obs[t], prior, A_dependencies, num_iter)
        G = neg_efe(qs, A, B, C, pm, A_dependencies)
        logq = torch.log_softmax(gamma * G, dim=-1)
        loss = loss + logq.gather(1, actions[t].view(-1, 1)).sum()
        u = act_f[actions[t]]
        prior = [torch.bmm(B[f][ar, :, :, u[:, f]], qs[f].unsqueeze(-1)).squeeze(-1) for f in range(len(qs))]
    return loss
