"""Sequence inference over a window (marginal / variational message passing)
on the gfx950 kernel `pymdp_hip_sequence_infer`.

Mirrors the sequence branch of pymdp/inference.py:194-271 (run_mmp /
run_vmp of pymdp/algos.py:311-500 after _condition_transitions_on_actions and
_build_sequence_validity_masks), batched over agents.  The reference has no
native layer for it (pure JAX); errors use the prefix "seq_ffi: ".
"""
from __future__ import annotations

from typing import Sequence

import numpy as np
import torch

from .. import _native
from ._core import flatten_dep_lists

METHODS = {"mmp": 0, "vmp": 1}


def sequence_infer_ffi(method: str, A: Sequence[torch.Tensor], B: Sequence[torch.Tensor], obs, prior,
                       A_dependencies, B_dependencies=None, *, past_actions=None, valid_steps=None,
                       num_iter: int = 16, tau: float = 1.0, distr_obs: bool = True,
                       batch: int | None = None) -> list[torch.Tensor]:
    """Posterior q[f] [batch, T, S_f] of the window.

    A[m]   [batch?, O_m, S_deps...]          B[f] [batch?, S_f, S_f, U_f] (B_dependencies[f] == [f])
    obs[m] [batch, T] indices, or [batch, T, O_m] distributions (distr_obs)
    prior[f] [batch, S_f] or [S_f]           past_actions [batch, T-1, F] int or None
    valid_steps None, an int or [batch] ints (pymdp/inference.py:100-126)
    """
    if method not in METHODS:
        raise ValueError(f"seq_ffi: method must be 'mmp' or 'vmp', got {method!r}")
    F, M = len(prior), len(A)
    if B_dependencies is not None and any(list(d) != [f] for f, d in enumerate(B_dependencies)):
        raise NotImplementedError("seq_ffi: sequence inference supports factor-local transitions only "
                                  "(B_dependencies[f] == [f])")
    dev = prior[0].device
    if batch is None:
        batch = int(prior[0].shape[0]) if prior[0].dim() == 2 else int(obs[0].shape[0])
    S = [int(p.shape[-1]) for p in prior]
    U = [int(b.shape[-1]) for b in B]
    T = int(obs[0].shape[1])
    O = []
    A_ptrs, A_bs = [], []
    for m, (a, deps) in enumerate(zip(A, A_dependencies)):
        if a.dtype != torch.float32 or not a.is_contiguous() or a.device != dev:
            raise ValueError(f"seq_ffi: A[{m}] must be a contiguous float32 tensor on {dev}")
        shared = a.dim() == len(deps) + 1
        O.append(int(a.shape[0] if shared else a.shape[1]))
        A_ptrs.append(a.data_ptr())
        A_bs.append(0 if shared else int(a[0].numel()))
    B_ptrs, B_bs = [], []
    for f, b in enumerate(B):
        if b.dim() not in (3, 4) or tuple(b.shape[-3:-1]) != (S[f], S[f]):
            raise ValueError(f"seq_ffi: B[{f}] must have shape (batch?, {S[f]}, {S[f]}, U)")
        if b.dtype != torch.float32 or not b.is_contiguous() or b.device != dev:
            raise ValueError(f"seq_ffi: B[{f}] must be a contiguous float32 tensor on {dev}")
        B_ptrs.append(b.data_ptr())
        B_bs.append(0 if b.dim() == 3 else int(b[0].numel()))
    P_ptrs, P_bs = [], []
    prior = [p.to(torch.float32).contiguous() for p in prior]
    for p in prior:
        P_ptrs.append(p.data_ptr())
        P_bs.append(int(p.shape[-1]) if p.dim() == 2 else 0)
    if distr_obs:
        obs_t = torch.cat([o.to(torch.float32).reshape(batch, T, -1) for o in obs], -1).contiguous()
    else:
        obs_t = torch.stack([o.reshape(batch, T).to(torch.int32) for o in obs], -1).contiguous()
    obs_off = np.zeros(M + 1, np.int64)
    obs_off[1:] = np.cumsum(O)
    act = None
    if past_actions is not None:
        act = torch.as_tensor(past_actions, device=dev).to(torch.int32).reshape(batch, max(T - 1, 0), F).contiguous()
    vs = None
    if valid_steps is not None:
        vs = torch.as_tensor(valid_steps, device=dev).to(torch.int32).reshape(-1)
        vs = (vs.expand(batch) if vs.numel() == 1 else vs).contiguous()
        if vs.numel() != batch:
            raise ValueError("seq_ffi: valid_steps must be a scalar or have one entry per agent")
    out = torch.empty((batch, T, sum(S)), dtype=torch.float32, device=dev)
    df, do = flatten_dep_lists(A_dependencies)
    i64 = {k: _native.I64(v) for k, v in dict(As=A_bs, Bs=B_bs, Ps=P_bs, S=S, U=U, O=O, oo=obs_off, df=df,
                                                 do=do).items()}
    keep = (A, B, prior, obs_t, act, vs)  # alive until the launch is enqueued
    rc = _native.lib().pymdp_hip_sequence_infer(
        METHODS[method], _native.ptr_array(A_ptrs), i64["As"].ptr, _native.ptr_array(B_ptrs), i64["Bs"].ptr,
        obs_t.data_ptr(), 0 if distr_obs else 1, _native.ptr_array(P_ptrs), i64["Ps"].ptr,
        act.data_ptr() if act is not None else None, vs.data_ptr() if vs is not None else None, out.data_ptr(),
        batch, T, i64["S"].ptr, i64["U"].ptr, F, i64["O"].ptr, M, i64["oo"].ptr, i64["df"].ptr, i64["do"].ptr,
        int(num_iter), float(tau), _native.stream_handle(dev))
    del keep
    _native.check(rc)
    off = np.concatenate([[0], np.cumsum(S)])
    return [out[:, :, int(off[f]):int(off[f + 1])] for f in range(F)]


__all__ = ["sequence_infer_ffi", "METHODS"]
