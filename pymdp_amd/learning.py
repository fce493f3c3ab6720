"""Dirichlet-parameter learning (mirror of pymdp/learning.py), batched over a
leading agent axis the way Agent.infer_parameters vmaps the reference.

Every update is one launch of the fused HIP kernel of libpymdp_hip
(`pymdp_hip_update_pA` / `pymdp_hip_update_pB`): the count tensor is read once
and the posterior counts and their expected value are written once.  There is
no other implementation: the functions raise if the library is missing.

Shapes (batch = number of agents):
    pA[m]  [batch, O_m, S_deps...] or [O_m, S_deps...] (shared by the batch)
    pB[f]  [batch, S_f, S_deps..., U_f] or without the batch axis
    obs[m] [batch, T] int indices, or [batch, T, O_m] when categorical_obs
    qs[f]  [batch, T, S_f]
    actions [batch, T - 1, F] int (negative = no valid action)
    lr     float or [batch] tensor
"""
from __future__ import annotations

from typing import Sequence

import torch

from . import _native


def _lr(lr, batch, dev):
    if isinstance(lr, torch.Tensor):
        t = lr.to(device=dev, dtype=torch.float32).reshape(-1)
        if t.numel() == 1:
            return None, float(t.item()), None
        if t.numel() != batch:
            raise ValueError(f"lr must be a scalar or have one entry per agent ({batch}); got {t.numel()}")
        t = t.contiguous()
        return t.data_ptr(), 0.0, t
    return None, float(lr), None


def _counts(p, core: int, name: str):
    """(tensor, batch stride) of a count tensor with `core` non-batch axes."""
    p = p.to(torch.float32)
    if p.dim() == core:
        return p.contiguous(), 0
    if p.dim() != core + 1:
        raise ValueError(f"{name} must have {core} or {core + 1} dimensions; got shape {tuple(p.shape)}")
    if p.shape[0] == 1:
        return p[0].contiguous(), 0
    p = p.contiguous()
    return p, int(p[0].numel())


def _seq(x, batch, name):
    x = x.to(torch.float32)
    if x.dim() == 2:  # [batch, S]: one time step
        x = x[:, None, :]
    if x.dim() != 3 or x.shape[0] != batch:
        raise ValueError(f"{name} must have shape (batch, T, n); got {tuple(x.shape)}")
    return x.contiguous()


def update_obs_likelihood_dirichlet_m(pA_m: torch.Tensor, obs_m: torch.Tensor, qs: Sequence[torch.Tensor],
                                      dependencies_m: Sequence[int], lr=1.0):
    """pymdp/learning.py:11-55: obs_m [batch, T, O] distributional; returns
    (qA_m, E[qA_m]), each [batch, O, S_deps...]."""
    deps = list(dependencies_m)
    dev = obs_m.device
    batch = int(obs_m.shape[0])
    obs_m = _seq(obs_m, batch, "obs")
    T, O = int(obs_m.shape[1]), int(obs_m.shape[2])
    q_deps = [_seq(qs[f], batch, f"qs[{f}]") for f in deps]
    for f, q in zip(deps, q_deps):
        if int(q.shape[1]) != T:
            raise ValueError(f"qs[{f}] has {int(q.shape[1])} time steps, observations have {T}")
    sizes = [int(q.shape[2]) for q in q_deps]
    p, p_bs = _counts(pA_m, 1 + len(deps), "pA")
    if tuple(p.shape[-1 - len(deps):]) != tuple([O] + sizes):
        raise ValueError(f"pA has shape {tuple(pA_m.shape)}, expected (..., {O}, {', '.join(map(str, sizes))})")
    qA = torch.empty((batch, O, *sizes), dtype=torch.float32, device=dev)
    EA = torch.empty_like(qA)
    lr_ptr, lr_sc, _keep = _lr(lr, batch, dev)
    rc = _native.lib().pymdp_hip_update_pA(
        p.data_ptr(), p_bs, qA.data_ptr(), EA.data_ptr(), batch, T, O, len(deps), _native.I64(sizes or [0]).ptr,
        obs_m.data_ptr(), T * O, _native.ptr_array([q.data_ptr() for q in q_deps]),
        _native.I64([T * s for s in sizes] or [0]).ptr, lr_ptr, lr_sc, _native.stream_handle(dev))
    _native.check(rc)
    return qA, EA


def update_obs_likelihood_dirichlet(pA, A, obs, qs, *, A_dependencies, categorical_obs: bool,
                                    num_obs: Sequence[int], lr):
    """pymdp/learning.py:57-117: (qA, E_qA); a modality whose pA is None keeps
    its A."""
    qA, E_qA = [], []
    for m, pA_m in enumerate(pA):
        if pA_m is None:
            qA.append(None)
            E_qA.append(A[m])
            continue
        o = torch.as_tensor(obs[m], device=qs[0].device)
        if not categorical_obs:
            o = torch.nn.functional.one_hot(o.long().reshape(o.shape[0], -1), int(num_obs[m])).float()
        a, e = update_obs_likelihood_dirichlet_m(pA_m, o, qs, A_dependencies[m], lr=lr)
        qA.append(a)
        E_qA.append(e)
    return qA, E_qA


def _update_B_f(pB_f, joint_qs_f: Sequence[torch.Tensor], actions: torch.Tensor, factor: int, lr,
                mask_all: bool):
    nxt = joint_qs_f[0]
    batch = int(nxt.shape[0])
    dev = nxt.device
    nxt = _seq(nxt, batch, "joint beliefs[0]")
    T1, S = int(nxt.shape[1]), int(nxt.shape[2])
    deps = [_seq(x, batch, f"joint beliefs[{j + 1}]") for j, x in enumerate(joint_qs_f[1:])]
    sizes = [int(x.shape[2]) for x in deps]
    core = 2 + len(deps)
    p, p_bs = _counts(pB_f, core, "pB")
    U = int(p.shape[-1])
    if tuple(p.shape[-core:]) != tuple([S] + sizes + [U]):
        raise ValueError(f"pB has shape {tuple(pB_f.shape)}, expected (..., {S}, "
                         f"{', '.join(map(str, sizes + [U]))})")
    act = torch.as_tensor(actions, device=dev).to(torch.int32)
    if act.dim() != 3 or act.shape[0] != batch or act.shape[1] != T1:
        raise ValueError(f"actions must have shape (batch, {T1}, F); got {tuple(act.shape)}")
    act = act.contiguous()
    F = int(act.shape[2])
    qB = torch.empty((batch, S, *sizes, U), dtype=torch.float32, device=dev)
    EB = torch.empty_like(qB)
    lr_ptr, lr_sc, _keep = _lr(lr, batch, dev)
    rc = _native.lib().pymdp_hip_update_pB(
        p.data_ptr(), p_bs, qB.data_ptr(), EB.data_ptr(), batch, T1, S, len(deps), _native.I64(sizes).ptr, U,
        nxt.data_ptr(), T1 * S, S, _native.ptr_array([x.data_ptr() for x in deps]),
        _native.I64([T1 * s for s in sizes]).ptr, _native.I64(sizes).ptr,
        act.data_ptr(), T1 * F, F, int(factor), 1 if mask_all else 0, lr_ptr, lr_sc, _native.stream_handle(dev))
    _native.check(rc)
    return qB, EB


def update_state_transition_dirichlet(pB, B, joint_beliefs, actions, *, num_controls: Sequence[int], lr,
                                      factors_to_update="all", mask_invalid: bool = False):
    """pymdp/learning.py:159-234: joint_beliefs[f] = [qs_f at t + 1, qs_dep at
    t, ...] each [batch, T - 1, .]; actions [batch, T - 1, F] int.  A factor
    not in factors_to_update gets (pB_f, E[pB_f]).  mask_invalid drops every
    t at which any factor's action is negative (what Agent.infer_parameters
    does by masking the joint beliefs)."""
    F = len(pB)
    upd = set(range(F)) if factors_to_update == "all" else set(factors_to_update)
    qB, E_qB = [], []
    for f in range(F):
        if f in upd:
            if int(pB[f].shape[-1]) != int(num_controls[f]):
                raise ValueError(f"pB[{f}] has {int(pB[f].shape[-1])} controls, num_controls says {num_controls[f]}")
            b, e = _update_B_f(pB[f], joint_beliefs[f], actions, f, lr, mask_invalid)
        else:
            jb = joint_beliefs[f]
            empty = [x[:, :0] for x in jb]
            act0 = torch.zeros((int(jb[0].shape[0]), 0, F), dtype=torch.int32, device=jb[0].device)
            b, e = _update_B_f(pB[f], empty, act0, f, 0.0, False)
        qB.append(b)
        E_qB.append(e)
    return qB, E_qB


__all__ = ["update_obs_likelihood_dirichlet_m", "update_obs_likelihood_dirichlet",
           "update_state_transition_dirichlet"]
