"""FPI boundary (mirror of pymdp/ffi/_fpi.py).

`can_handle_fpi` and `fpi_ffi` keep the reference's names, argument meaning
and error behaviour (pymdp/ffi/_fpi.py:22-65, :95-140, :143-270); the work is
done by the gfx950 kernel behind `pymdp_hip_fpi`.  `infer_states_ffi` is the
fused log-likelihood + FPI entry (`pymdp_hip_infer_states`) used by
`pymdp_amd.algos.run_factorized_fpi`.
"""
from __future__ import annotations

from typing import Sequence

import numpy as np
import torch

from .. import _native
from ._core import dep_indices_in_range, dep_list_rank_ok, flatten_dep_lists, is_available

# Limits of the gfx950 kernel (pymdp_amd/csrc/fpi.hip).
MAX_FACTORS = 16
MAX_MODALITIES = 32
MAX_DEP_TUPLES = 8
MAX_DEP_STATE_DIM = 64


def can_handle_fpi(prior: Sequence[torch.Tensor], A_dependencies: Sequence[Sequence[int]], num_iter: int) -> bool:
    """True iff the HIP FPI kernel can run this problem (pymdp/ffi/_fpi.py:22-65).

    Restrictions (reference): num_iter > 0, F >= 1, M >= 1, every modality's
    dependency rank in [1, 8], indices in range and distinct, library
    loadable.  Kernel limits on top: F <= 16, M <= 32, at most 8 distinct
    dependency tuples, every dependency's state dimension <= 64.
    """
    if not is_available():
        return False
    if num_iter <= 0:
        return False
    F = len(prior)
    M = len(A_dependencies)
    if F < 1 or M < 1 or F > MAX_FACTORS or M > MAX_MODALITIES:
        return False
    tuples = set()
    for deps in A_dependencies:
        if not dep_list_rank_ok(deps) or not dep_indices_in_range(deps, F):
            return False
        if len(set(deps)) != len(deps):
            return False
        for d in deps:
            if int(prior[d].shape[-1]) > MAX_DEP_STATE_DIM:
                return False
        tuples.add(tuple(deps))
    return len(tuples) <= MAX_DEP_TUPLES


def _validate_fpi_inputs(log_likelihoods, log_prior, A_dependencies, num_iter) -> None:
    """pymdp/ffi/_fpi.py:95-140."""
    F = len(log_prior)
    M = len(log_likelihoods)
    if F < 1 or M < 1:
        raise ValueError("FPI FFI requires at least one factor and one modality")
    if not can_handle_fpi(log_prior, A_dependencies, num_iter):
        raise ValueError(
            "FPI FFI cannot handle this problem (rank > 8, num_iter <= 0, or library unavailable)"
        )
    if len(A_dependencies) != M:
        raise ValueError(
            f"A_dependencies length {len(A_dependencies)} does not match len(log_likelihoods) = {M}"
        )
    S = [int(lp.shape[-1]) for lp in log_prior]
    for m, (ll, deps) in enumerate(zip(log_likelihoods, A_dependencies)):
        rank = len(deps)
        if ll.dim() < rank:
            raise ValueError(f"log_likelihoods[{m}] rank {ll.dim()} < |A_deps[{m}]| = {rank}")
        trailing = tuple(ll.shape[ll.dim() - rank:])
        expected = tuple(S[d] for d in deps)
        if trailing != expected:
            raise ValueError(
                f"log_likelihoods[{m}] trailing shape {trailing} does not match expected {expected} "
                f"from A_dependencies[{m}] = {tuple(deps)}"
            )


def fpi_ffi(log_likelihoods: Sequence[torch.Tensor], log_prior: Sequence[torch.Tensor],
            A_dependencies: Sequence[Sequence[int]], num_iter: int) -> list[torch.Tensor]:
    """Generalised fused FPI (pymdp/ffi/_fpi.py:143-270).

    log_likelihoods[m]: [batch?, S[deps[0]], ...]; log_prior[f]: [batch?, S_f].
    Returns qs[f] with the same leading shape as log_prior[f].  Runs all
    num_iter sweeps (no early stop), like the reference JAX path.
    """
    _validate_fpi_inputs(log_likelihoods, log_prior, A_dependencies, num_iter)
    F = len(log_prior)
    M = len(log_likelihoods)
    S = np.array([int(p.shape[-1]) for p in log_prior], dtype=np.int64)
    lead = tuple(log_prior[0].shape[:-1])
    batch = int(np.prod(lead)) if lead else 1
    ll_offsets = np.zeros(M + 1, dtype=np.int64)
    for m, deps in enumerate(A_dependencies):
        ll_offsets[m + 1] = ll_offsets[m] + int(np.prod([S[d] for d in deps]))
    lp_offsets = np.zeros(F + 1, dtype=np.int64)
    lp_offsets[1:] = np.cumsum(S)
    dev = log_prior[0].device
    for name, ts in (("log_likelihoods", log_likelihoods), ("log_prior", log_prior)):
        for i, t in enumerate(ts):
            if not t.is_cuda or t.device != dev:
                raise ValueError(f"fpi_ffi: {name}[{i}] must be a HIP tensor on {dev}")
    ll_flat = torch.cat([ll.to(torch.float32).reshape(batch, -1) for ll in log_likelihoods], -1).contiguous()
    lp_flat = torch.cat([lp.to(torch.float32).reshape(batch, -1) for lp in log_prior], -1).contiguous()
    q_flat = torch.empty_like(lp_flat)
    A_dep_flat, A_dep_offsets = flatten_dep_lists(A_dependencies)
    a = [_native.I64(x) for x in (S, ll_offsets, lp_offsets, A_dep_flat, A_dep_offsets)]
    rc = _native.lib().pymdp_hip_fpi(
        ll_flat.data_ptr(), lp_flat.data_ptr(), q_flat.data_ptr(),
        ll_flat.numel(), lp_flat.numel(), q_flat.numel(),
        a[0].ptr, F, a[1].ptr, M, a[2].ptr, a[3].ptr, a[4].ptr,
        int(num_iter), _native.stream_handle(dev))
    _native.check(rc)
    return [q_flat[:, int(lp_offsets[f]):int(lp_offsets[f + 1])].reshape(lead + (int(S[f]),))
            for f in range(F)]


def infer_states_ffi(A: Sequence[torch.Tensor], obs: Sequence[torch.Tensor], prior: Sequence[torch.Tensor],
                     A_dependencies: Sequence[Sequence[int]], num_iter: int, *, distr_obs: bool = True,
                     method_exact: bool = False, batch: int | None = None,
                     out: torch.Tensor | None = None) -> list[torch.Tensor]:
    """Fused log-likelihood + FPI for a batch of agents (`pymdp_hip_infer_states`).

    A[m]     [batch, O_m, S_deps...] per-agent, or [O_m, S_deps...] shared
    obs[m]   distr_obs: [batch, O_m] distributions; else [batch] int indices
    prior[f] [batch, S_f] probabilities (any row stride; last axis contiguous)
    Returns qs[f] [batch, S_f] (views into one packed [batch, sum S] buffer).
    """
    k = _pack_infer_args(A, obs, prior, A_dependencies, num_iter, distr_obs, batch)
    batch, F, lp_offsets = k["batch"], k["F"], k["lp_offsets"]
    totS = int(lp_offsets[-1])
    if out is None:
        out = torch.empty((batch, totS), dtype=torch.float32, device=prior[0].device)
    a = k["i64"]
    rc = _native.lib().pymdp_hip_infer_states(
        k["Ap"], a["As"].ptr, k["obs_t"].data_ptr(), k["is_index"], k["Pp"], a["Ps"].ptr, out.data_ptr(), batch,
        a["S"].ptr, F, a["O"].ptr, k["M"], a["oo"].ptr, a["lo"].ptr, a["df"].ptr, a["do"].ptr,
        int(num_iter), int(bool(method_exact)), _native.stream_handle(prior[0].device))
    _native.check(rc)
    return [out[:, int(lp_offsets[f]):int(lp_offsets[f + 1])] for f in range(F)]


def fpi_vfe_ffi(A: Sequence[torch.Tensor], obs: Sequence[torch.Tensor], prior: Sequence[torch.Tensor],
                qs: torch.Tensor, A_dependencies: Sequence[Sequence[int]], *, distr_obs: bool = True,
                batch: int | None = None) -> torch.Tensor:
    """Variational free energy of a one-step posterior (`pymdp_hip_fpi_vfe`),
    the return_info diagnostics of pymdp/inference.py:456-483 for "fpi" /
    "exact": [batch, 4] = (neg_entropy, prior_cross_entropy, accuracy, vfe).
    A, obs, prior as for infer_states_ffi; qs the packed [batch, sum S]
    posterior."""
    k = _pack_infer_args(A, obs, prior, A_dependencies, 1, distr_obs, batch)
    batch = k["batch"]
    totS = int(k["lp_offsets"][-1])
    if not qs.is_cuda or qs.device != prior[0].device or qs.numel() != batch * totS:
        raise ValueError(f"qs must be a HIP tensor on {prior[0].device} holding [{batch}, {totS}] values")
    qs = qs.to(torch.float32).reshape(batch, totS).contiguous()
    out = torch.empty((batch, 4), dtype=torch.float32, device=qs.device)
    a = k["i64"]
    rc = _native.lib().pymdp_hip_fpi_vfe(
        k["Ap"], a["As"].ptr, k["obs_t"].data_ptr(), k["is_index"], k["Pp"], a["Ps"].ptr, qs.data_ptr(),
        out.data_ptr(), batch, a["S"].ptr, k["F"], a["O"].ptr, k["M"], a["oo"].ptr, a["df"].ptr, a["do"].ptr,
        _native.stream_handle(qs.device))
    _native.check(rc)
    return out


def _pack_infer_args(A, obs, prior, A_dependencies, num_iter, distr_obs, batch):
    """Device pointers, strides and int64 attribute arrays of the fused
    infer_states ABI (shared with the VFE entry)."""
    F = len(prior)
    M = len(A)
    if len(A_dependencies) != M or (not isinstance(obs, torch.Tensor) and len(obs) != M):
        raise ValueError("A, obs and A_dependencies must have one entry per modality")
    if batch is None:
        # the largest leading axis of the batched operands; [S_f] / [1, S_f]
        # priors and shared A broadcast over it
        lead = [int(p.shape[0]) for p in prior if p.dim() == 2]
        lead += [int(a.shape[0]) for a, d in zip(A, A_dependencies) if a.dim() == len(d) + 2]
        batch = max(lead) if lead else 1
    S = np.array([int(p.shape[-1]) for p in prior], dtype=np.int64)
    O = np.array([int(a.shape[0] if a.dim() == len(d) + 1 else a.shape[1])
                  for a, d in zip(A, A_dependencies)], dtype=np.int64)
    if not can_handle_fpi(prior, A_dependencies, max(1, num_iter)):
        raise ValueError("fpi_ffi: problem outside the HIP kernel's limits "
                         f"(F<={MAX_FACTORS}, M<={MAX_MODALITIES}, <= {MAX_DEP_TUPLES} dependency tuples, "
                         f"dependency state dims <= {MAX_DEP_STATE_DIM}, rank in [1, 8])")
    A_ptrs, A_stride = [], []
    for m, (a, deps) in enumerate(zip(A, A_dependencies)):
        if a.dtype != torch.float32 or not a.is_cuda:
            raise ValueError(f"A[{m}] must be a float32 HIP tensor")
        K = int(np.prod([S[d] for d in deps]))
        if a.dim() == len(deps) + 1:
            shared = True
        elif a.dim() == len(deps) + 2:
            shared = False
            if int(a.shape[0]) != batch:
                raise ValueError(f"A[{m}] batch {a.shape[0]} != {batch}")
        else:
            raise ValueError(f"A[{m}] has rank {a.dim()}, expected {len(deps) + 1} or {len(deps) + 2}")
        tail = tuple(a.shape[-len(deps):])
        if tail != tuple(int(S[d]) for d in deps):
            raise ValueError(f"A[{m}] trailing shape {tail} does not match S[A_dependencies[{m}]]")
        if not a.is_contiguous():
            raise ValueError(f"A[{m}] must be contiguous")
        A_ptrs.append(a.data_ptr())
        A_stride.append(0 if shared else int(O[m]) * K)
    dev = prior[0].device
    prior_ptrs, prior_stride = [], []
    for f, p in enumerate(prior):
        if p.dtype != torch.float32 or not p.is_cuda or p.device != dev:
            raise ValueError(f"prior[{f}] must be a float32 HIP tensor on {dev}")
        if p.stride(-1) != 1:
            raise ValueError(f"prior[{f}] last axis must be contiguous")
        if p.dim() not in (1, 2) or (p.dim() == 2 and int(p.shape[0]) not in (1, batch)):
            raise ValueError(f"prior[{f}] has shape {tuple(p.shape)}; expected (S_{f},), (1, S_{f}) "
                             f"or ({batch}, S_{f})")
        prior_ptrs.append(p.data_ptr())
        # [S_f] and [1, S_f] are one prior for every agent
        prior_stride.append(int(p.stride(0)) if p.dim() == 2 and int(p.shape[0]) == batch and batch > 1 else 0)
    for m, a in enumerate(A):
        if a.device != dev:
            raise ValueError(f"A[{m}] must be on {dev}, the device of the prior")
    obs_list = [obs] if isinstance(obs, torch.Tensor) else list(obs)
    for m, o in enumerate(obs_list):
        if not isinstance(o, torch.Tensor) or not o.is_cuda or o.device != dev:
            raise ValueError(f"obs[{m}] must be a HIP tensor on {dev}")
    if distr_obs:
        obs_t = torch.cat([o.to(torch.float32).reshape(batch, -1) for o in obs], -1).contiguous()
        if int(obs_t.shape[1]) != int(O.sum()):
            raise ValueError(f"distributional observations have {int(obs_t.shape[1])} entries per agent; "
                             f"expected sum(O) = {int(O.sum())}")
        for m, o in enumerate(obs):
            if int(o.shape[-1]) != int(O[m]):
                raise ValueError(f"obs[{m}] has {int(o.shape[-1])} outcomes; A[{m}] has {int(O[m])}")
        is_index = 0
    elif isinstance(obs, torch.Tensor):
        # pre-packed int32 [batch, M] index observations (no packing kernels)
        if obs.dtype != torch.int32 or tuple(obs.shape) != (batch, M) or not obs.is_contiguous():
            raise ValueError("packed index observations must be a contiguous int32 tensor [batch, M]")
        obs_t = obs
        is_index = 1
    else:
        obs_t = torch.stack([o.reshape(batch).to(torch.int32) for o in obs], -1).contiguous()
        is_index = 1
    obs_offsets = np.zeros(M + 1, dtype=np.int64)
    obs_offsets[1:] = np.cumsum(O)
    lp_offsets = np.zeros(F + 1, dtype=np.int64)
    lp_offsets[1:] = np.cumsum(S)
    A_dep_flat, A_dep_offsets = flatten_dep_lists(A_dependencies)
    i64 = {k: _native.I64(v) for k, v in dict(As=A_stride, Ps=prior_stride, S=S, O=O, oo=obs_offsets,
                                                 lo=lp_offsets, df=A_dep_flat, do=A_dep_offsets).items()}
    return dict(batch=batch, F=F, M=M, lp_offsets=lp_offsets, obs_t=obs_t, is_index=is_index, i64=i64,
                Ap=_native.ptr_array(A_ptrs), Pp=_native.ptr_array(prior_ptrs))


__all__ = ["can_handle_fpi", "fpi_ffi", "infer_states_ffi", "fpi_vfe_ffi"]
