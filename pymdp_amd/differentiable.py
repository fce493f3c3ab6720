"""Differentiable state and policy inference with hand-written HIP adjoints.

The reference differentiates through its FFI calls with `with_jax_grad`
(pymdp/ffi/_core.py:282-310): the forward runs the native kernel and the
backward re-runs jax.vjp of the JAX reference.  Here the forward runs the
HIP kernels of pymdp_amd.ffi and the backward runs the HIP adjoint kernels
(pymdp_hip_fpi_adjoint, pymdp_hip_neg_efe_adjoint), exposed as
torch.autograd.Function objects (the torch counterpart of jax.custom_vjp):

  infer_states_diff(A, obs, prior, A_dependencies, num_iter)  -> qs
  neg_efe_diff(policies, qs, A, B, C, A_dependencies, B_dependencies) -> G
  trajectory_loss(A_logits, B_logits, C_logits, D_logits, obs, actions, ...)

The FPI adjoint is the implicit one (fixed-point differentiation) where
the forward sweep has converged and the unrolled reverse sweep elsewhere
(mode="auto"); mode="unrolled" forces the exact derivative of the N-sweep
forward.
"""
from __future__ import annotations

from typing import Sequence

import numpy as np
import torch

from . import _native
from .ffi._core import flatten_dep_lists
from .ffi._efe import update_posterior_policies_ffi
from .ffi._fpi import infer_states_ffi

FLAG_USE_UTILITY = 1 << 0
FLAG_USE_STATES_INFO_GAIN = 1 << 1


def _ptrs(ts):
    return _native.ptr_array([t.data_ptr() for t in ts])


class _FPI(torch.autograd.Function):
    @staticmethod
    def forward(ctx, meta, obs_idx, *tensors):
        deps, num_iter, mode, tol, M = meta
        A, prior = list(tensors[:M]), [p.contiguous() for p in tensors[M:]]
        batch = int(prior[0].shape[0])
        totS = sum(int(p.shape[-1]) for p in prior)
        out = torch.empty((batch, totS), dtype=torch.float32, device=prior[0].device)
        infer_states_ffi(A, obs_idx, prior, deps, num_iter, distr_obs=False, out=out)
        ctx.meta = meta
        ctx.save_for_backward(obs_idx, *A, *prior)
        return out

    @staticmethod
    def backward(ctx, qbar):
        deps, num_iter, mode, tol, M = ctx.meta
        saved = ctx.saved_tensors
        obs_idx, A, prior = saved[0], list(saved[1:1 + M]), list(saved[1 + M:])
        A_bar, pb, _ = fpi_adjoint(A, obs_idx, prior, deps, num_iter, qbar, mode=mode, implicit_tol=tol)
        return (None, None, *A_bar, *pb)


def fpi_adjoint(A, obs_idx, prior, A_dependencies, num_iter, q_bar, *, mode="auto", implicit_tol=1e-6):
    """dL/dA (observed rows), dL/dprior and the per-agent choice (1 =
    implicit, 0 = unrolled) for dL/dq = q_bar [batch, sum S]."""
    M = len(A)
    batch = int(prior[0].shape[0])
    S = np.array([int(p.shape[-1]) for p in prior], np.int64)
    O = np.array([int(a.shape[1]) for a in A], np.int64)
    A = [a.contiguous() for a in A]
    prior = [p.contiguous() for p in prior]
    A_bar = [torch.zeros_like(a) for a in A]
    prior_bar = torch.empty((batch, int(S.sum())), dtype=torch.float32, device=prior[0].device)
    used = torch.empty(batch, dtype=torch.int32, device=prior[0].device)
    df, do = flatten_dep_lists(A_dependencies)
    k = {n: _native.I64(v) for n, v in dict(As=[a[0].numel() for a in A], Ps=[p.stride(0) for p in prior], S=S,
                                             O=O, df=df, do=do).items()}
    _native.check(_native.lib().pymdp_hip_fpi_adjoint(
        _ptrs(A), k["As"].ptr, obs_idx.to(torch.int32).contiguous().data_ptr(), _ptrs(prior), k["Ps"].ptr,
        q_bar.contiguous().data_ptr(), _ptrs(A_bar), prior_bar.data_ptr(), None, used.data_ptr(), batch,
        k["S"].ptr, len(S), k["O"].ptr, M, k["df"].ptr, k["do"].ptr, int(num_iter), 1 if mode == "auto" else 0,
        float(implicit_tol), _native.stream_handle(prior[0].device)))
    return A_bar, list(torch.split(prior_bar, [int(s) for s in S], dim=1)), used


class _EFE(torch.autograd.Function):
    @staticmethod
    def forward(ctx, meta, policy_dev, qs_packed, *tensors):
        pm, A_deps, B_deps, S, M, F, flags = meta
        A, B, C = list(tensors[:M]), list(tensors[M:M + F]), list(tensors[M + F:])
        qs = list(torch.split(qs_packed, S, dim=1))
        _, G = update_posterior_policies_ffi(pm, qs, A, B, C, None, A_deps, B_deps, gamma=1.0,
                                             use_utility=bool(flags & FLAG_USE_UTILITY),
                                             use_states_info_gain=bool(flags & FLAG_USE_STATES_INFO_GAIN),
                                             return_q_pi=False)
        ctx.meta = meta
        ctx.save_for_backward(policy_dev, qs_packed, *A, *B, *C)
        return G

    @staticmethod
    def backward(ctx, Gbar):
        pm, A_deps, B_deps, S, M, F, flags = ctx.meta
        saved = ctx.saved_tensors
        policy_dev, qs_packed = saved[0], saved[1].contiguous()
        A, B, C = list(saved[2:2 + M]), list(saved[2 + M:2 + M + F]), list(saved[2 + M + F:])
        batch = int(qs_packed.shape[0])
        P = int(pm.shape[0])
        qs_bar = torch.empty_like(qs_packed)
        A_bar = [torch.empty_like(a) for a in A]
        B_bar = [torch.empty_like(b) for b in B]
        C_bar = [torch.empty_like(c) for c in C]
        Adf, Ado = flatten_dep_lists(A_deps)
        Bdf, Bdo = flatten_dep_lists(B_deps)
        k = {n: _native.I64(v) for n, v in dict(
            As=[a[0].numel() for a in A], Bs=[b[0].numel() for b in B], Cs=[c[0].numel() for c in C],
            S=S, O=[int(a.shape[1]) for a in A], U=[int(b.shape[-1]) for b in B], Adf=Adf, Ado=Ado, Bdf=Bdf,
            Bdo=Bdo).items()}
        _native.check(_native.lib().pymdp_hip_neg_efe_adjoint(
            policy_dev.data_ptr(), P, int(pm.shape[1]), qs_packed.data_ptr(), _ptrs(A), k["As"].ptr, _ptrs(B),
            k["Bs"].ptr, _ptrs(C), k["Cs"].ptr, Gbar.contiguous().data_ptr(), qs_bar.data_ptr(), _ptrs(A_bar),
            _ptrs(B_bar), _ptrs(C_bar), batch, k["S"].ptr, F, k["O"].ptr, M, k["U"].ptr, k["Adf"].ptr,
            k["Ado"].ptr, k["Bdf"].ptr, k["Bdo"].ptr, int(flags), _native.stream_handle(qs_packed.device)))
        return (None, None, qs_bar, *A_bar, *B_bar, *C_bar)


def infer_states_diff(A: Sequence[torch.Tensor], obs_idx: torch.Tensor, prior: Sequence[torch.Tensor],
                      A_dependencies, num_iter: int = 16, *, mode: str = "auto",
                      implicit_tol: float = 1e-6) -> torch.Tensor:
    """Packed posterior [batch, sum S] of the fused FPI, differentiable w.r.t.
    A (per agent, [batch, O_m, S_deps...]) and prior ([batch, S_f]).
    obs_idx: int32 [batch, M] observation indices."""
    if mode not in ("auto", "unrolled"):
        raise ValueError("mode must be 'auto' or 'unrolled'")
    meta = (A_dependencies, int(num_iter), mode, float(implicit_tol), len(A))
    return _FPI.apply(meta, obs_idx.to(torch.int32).contiguous(), *[a.contiguous() for a in A], *prior)


def neg_efe_diff(policies: np.ndarray, policy_dev: torch.Tensor, qs_packed: torch.Tensor, S: Sequence[int],
                 A, B, C, A_dependencies, B_dependencies, *, use_utility: bool = True,
                 use_states_info_gain: bool = True) -> torch.Tensor:
    """G [batch, P] of every policy (horizon 1), differentiable w.r.t. the
    packed posterior, A, B and C (all per agent)."""
    pm = np.ascontiguousarray(np.asarray(policies, np.int32))
    flags = (FLAG_USE_UTILITY if use_utility else 0) | (FLAG_USE_STATES_INFO_GAIN if use_states_info_gain else 0)
    meta = (pm, A_dependencies, B_dependencies, [int(s) for s in S], len(A), len(B), flags)
    return _EFE.apply(meta, policy_dev, qs_packed, *[a.contiguous() for a in A], *[b.contiguous() for b in B],
                      *[c.contiguous() for c in C])


def trajectory_loss(A_logits, B_logits, C_logits, D_logits, obs: torch.Tensor, actions: torch.Tensor,
                    policies: np.ndarray, A_dependencies, B_dependencies, *, num_iter: int = 16,
                    gamma: float = 1.0, mode: str = "auto") -> torch.Tensor:
    """sum_t sum_b ln q_pi[b, actions[t, b]] of an agent batch that observes
    obs[t] (int [T, batch, M]) and takes policy actions[t] (long [T, batch])
    at each step; A, B, D = softmax of their logits over the outcome / next
    state axis, C = log_softmax (log preferences).  The prior of step t+1 is
    B[:, :, u] q_t for the policy's action u (Agent.update_empirical_prior,
    pymdp/agent.py:825-851); q_pi = softmax(gamma G + ln E) with E uniform
    (pymdp/control.py:963)."""
    A = [torch.softmax(x, dim=1) for x in A_logits]
    B = [torch.softmax(x, dim=1) for x in B_logits]
    C = [torch.log_softmax(x, dim=-1) for x in C_logits]
    prior = [torch.softmax(x, dim=-1) for x in D_logits]
    S = [int(x.shape[-1]) for x in D_logits]
    dev = A[0].device
    pm = np.ascontiguousarray(np.asarray(policies, np.int32))
    policy_dev = torch.as_tensor(pm, device=dev)
    act_f = torch.as_tensor(pm[:, 0, :], device=dev, dtype=torch.long)
    batch = int(A[0].shape[0])
    ar = torch.arange(batch, device=dev)
    loss = torch.zeros((), device=dev)
    for t in range(int(obs.shape[0])):
        qs_packed = infer_states_diff(A, obs[t], prior, A_dependencies, num_iter, mode=mode)
        G = neg_efe_diff(pm, policy_dev, qs_packed, S, A, B, C, A_dependencies, B_dependencies)
        logq = torch.log_softmax(gamma * G, dim=-1)
        loss = loss + logq.gather(1, actions[t].view(-1, 1)).sum()
        u = act_f[actions[t]]  # [batch, F]
        qs = torch.split(qs_packed, S, dim=1)
        prior = [torch.bmm(B[f][ar, :, :, u[:, f]], qs[f].unsqueeze(-1)).squeeze(-1) for f in range(len(S))]
    return loss
