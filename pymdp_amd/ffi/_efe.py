"""Neg-EFE boundary (mirror of pymdp/ffi/_efe.py).

`can_handle`, `neg_efe_all_policies_ffi` and the FLAG_* bits keep the
reference's names, argument meaning and error behaviour
(pymdp/ffi/_efe.py:26-37, :296-358, :361-488, :491-711).  The kernels behind
them are the gfx950 action-history pipeline of pymdp_amd/csrc/efe.hip.
`update_posterior_policies_ffi` additionally fuses q_pi = softmax(gamma G +
ln E) into the same launch sequence.
"""
from __future__ import annotations

from typing import Sequence

import numpy as np
import torch

from .. import _native
from ._core import dep_indices_in_range, dep_list_rank_ok, flatten_dep_lists, is_available

FLAG_USE_UTILITY = 1 << 0
FLAG_USE_STATES_INFO_GAIN = 1 << 1
FLAG_USE_INDUCTIVE = 1 << 2
FLAG_USE_PARAM_INFO_GAIN = 1 << 3
FLAG_MODEL_PARAMS_STATIC = 1 << 4

MAX_FACTORS = 16
MAX_MODALITIES = 32
MAX_HORIZON = 8


def can_handle(A_dependencies, B_dependencies, use_param_info_gain: bool, *, num_factors=None,
               num_modalities=None, ffi_beneficial=None) -> bool:
    """pymdp/ffi/_efe.py:296-358 (the HIP pipeline also handles the
    parameter information-gain term)."""
    if not is_available():
        return False
    if ffi_beneficial is False:
        return False
    if num_modalities is not None and len(A_dependencies) != num_modalities:
        return False
    if num_factors is not None and len(B_dependencies) != num_factors:
        return False
    if len(A_dependencies) > MAX_MODALITIES or len(B_dependencies) > MAX_FACTORS:
        return False
    if len({tuple(d) for d in A_dependencies}) > 8:
        return False
    for deps in A_dependencies:
        if not dep_list_rank_ok(deps):
            return False
        if num_factors is not None and not dep_indices_in_range(deps, num_factors):
            return False
    for deps in B_dependencies:
        if not dep_list_rank_ok(deps):
            return False
        if num_factors is not None and not dep_indices_in_range(deps, num_factors):
            return False
    return True


def _policy_host(policy_matrix) -> np.ndarray:
    if isinstance(policy_matrix, torch.Tensor):
        policy_matrix = policy_matrix.detach().cpu().numpy()
    pm = np.ascontiguousarray(np.asarray(policy_matrix, dtype=np.int32))
    if pm.ndim != 3:
        raise ValueError(f"policy_matrix must have shape (P, T, F); got {pm.shape}")
    return pm


def _flags(use_utility, use_states_info_gain, use_param_info_gain, use_inductive, static):
    flags = 0
    if use_utility:
        flags |= FLAG_USE_UTILITY
    if use_states_info_gain:
        flags |= FLAG_USE_STATES_INFO_GAIN
    if use_inductive:
        flags |= FLAG_USE_INDUCTIVE
    if use_param_info_gain:
        flags |= FLAG_USE_PARAM_INFO_GAIN
    if static:
        flags |= FLAG_MODEL_PARAMS_STATIC
    return flags


# Identity of the parameter tensors whose derived tables the library holds
# under FLAG_MODEL_PARAMS_STATIC.  The library's warm test is device-pointer
# identity; torch may hand a freed tensor's address to a new tensor, so the
# flag is passed only after checking that the same live tensor objects, with
# unchanged in-place version counters, fill the pointers; otherwise the cache
# is dropped first.
_static_owner: list = []


def _static_params_guard(tensors) -> None:
    import weakref
    global _static_owner
    cur = [(t.data_ptr(), t._version) for t in tensors]
    same = len(_static_owner) == len(tensors) and all(
        ref() is t and key == k for (ref, key), t, k in zip(_static_owner, tensors, cur))
    if not same:
        _native.check(_native.lib().pymdp_hip_param_cache_clear())
        _static_owner = [(weakref.ref(t), k) for t, k in zip(tensors, cur)]


def update_posterior_policies_ffi(policy_matrix, qs_init: Sequence[torch.Tensor], A: Sequence[torch.Tensor],
                                  B: Sequence[torch.Tensor], C: Sequence[torch.Tensor], E: torch.Tensor | None,
                                  A_dependencies, B_dependencies, *, gamma=16.0, I=None,
                                  inductive_epsilon=1e-3, use_utility=True, use_states_info_gain=True,
                                  use_param_info_gain=False, use_inductive=False, ffi_cache_params_static=False,
                                  return_q_pi=True, C_time: Sequence[bool] | None = None, pA=None, pB=None):
    """Batched (q_pi, neg_efe) for every agent; shapes [batch, P].

    qs_init[f] [batch, S_f]; A[m] [batch?, O_m, S_deps...]; B[f]
    [batch?, S_f, S_Bdeps..., U_f]; C[m] [batch?, O_m] or [batch?, T, O_m];
    E [batch, P] or None (uniform); gamma float or [batch] tensor;
    I[f] [batch?, depth_f, S_f] (inductive only); inductive_epsilon float or
    [batch] tensor; pA / pB Dirichlet counts shaped like A / B (read only
    with use_param_info_gain, pymdp/control.py:511-602).
    """
    pm = _policy_host(policy_matrix)
    P, T, Fp = pm.shape
    F = len(qs_init)
    M = len(A)
    if F < 1 or M < 1:
        raise ValueError("neg-EFE FFI requires at least one factor and one modality")
    if len(B) != F:
        raise ValueError("B must have one entry per hidden-state factor")
    if len(C) != M:
        raise ValueError("C must have one entry per observation modality")
    if Fp != F:
        raise ValueError(f"policy_matrix shape {pm.shape} is inconsistent with {F} factors")
    if use_param_info_gain and pA is None and pB is None:
        raise ValueError("use_param_info_gain=True requires at least one of pA or pB.")
    if not can_handle(A_dependencies, B_dependencies, False, num_factors=F, num_modalities=M):
        raise ValueError("unsupported or inconsistent A/B dependency metadata")
    if T > MAX_HORIZON:
        raise ValueError(f"efe_ffi: policy horizon {T} above the supported {MAX_HORIZON}")
    batch = int(qs_init[0].shape[0])
    dev = qs_init[0].device
    for f, q in enumerate(qs_init):
        if not q.is_cuda or q.device != dev:
            raise ValueError(f"qs_init[{f}] must be a HIP tensor on {dev}")
        if q.dim() != 2 or int(q.shape[0]) != batch:
            raise ValueError(f"qs_init[{f}] has shape {tuple(q.shape)}; expected ({batch}, S_{f})")
    S = np.array([int(q.shape[-1]) for q in qs_init], dtype=np.int64)
    O = np.array([int(a.shape[-len(d) - 1]) for a, d in zip(A, A_dependencies)], dtype=np.int64)
    U = np.array([int(b.shape[-1]) for b in B], dtype=np.int64)
    qs_flat = torch.cat([q.to(torch.float32).reshape(batch, -1) for q in qs_init], -1).contiguous()

    def _ptr_stride(t: torch.Tensor, core_ndim: int, name: str):
        if t.dtype != torch.float32 or not t.is_cuda or t.device != dev:
            raise ValueError(f"{name} must be a float32 HIP tensor on {dev}")
        if not t.is_contiguous():
            raise ValueError(f"{name} must be contiguous")
        if t.dim() == core_ndim:
            return t.data_ptr(), 0
        if t.dim() == core_ndim + 1 and int(t.shape[0]) == batch:
            return t.data_ptr(), int(t[0].numel())
        raise ValueError(f"{name} has shape {tuple(t.shape)}; expected {core_ndim} dims (+ batch {batch})")

    A_p, A_s = zip(*[_ptr_stride(a, len(d) + 1, f"A[{m}]") for m, (a, d) in enumerate(zip(A, A_dependencies))])
    for m, (a, d) in enumerate(zip(A, A_dependencies)):
        if tuple(a.shape[-len(d):]) != tuple(int(S[x]) for x in d):
            raise ValueError(f"A[{m}] shape {tuple(a.shape)} does not match S[A_dependencies[{m}]]")
    B_p, B_s = zip(*[_ptr_stride(b, len(d) + 2, f"B[{f}]") for f, (b, d) in enumerate(zip(B, B_dependencies))])
    for f, (b, d) in enumerate(zip(B, B_dependencies)):
        exp = (int(S[f]),) + tuple(int(S[x]) for x in d)
        if tuple(b.shape[-len(d) - 2:-1]) != exp:
            raise ValueError(f"B[{f}] shape {tuple(b.shape)} does not match expected {exp} + (U,)")
    pA_p = pA_s = pB_p = pB_s = None
    if use_param_info_gain and pA is not None:
        if len(pA) != M:
            raise ValueError("pA must have one entry per observation modality")
        pA = [x.to(torch.float32).contiguous() for x in pA]
        for m, (x, a, d) in enumerate(zip(pA, A, A_dependencies)):
            if tuple(x.shape[-len(d) - 1:]) != tuple(a.shape[-len(d) - 1:]):
                raise ValueError(f"pA[{m}] must have the shape of A[{m}]")
        pA_p, pA_s = zip(*[_ptr_stride(x, len(d) + 1, f"pA[{m}]") for m, (x, d) in enumerate(zip(pA, A_dependencies))])
    if use_param_info_gain and pB is not None:
        if len(pB) != F:
            raise ValueError("pB must have one entry per hidden-state factor")
        pB = [x.to(torch.float32).contiguous() for x in pB]
        for f, (x, b, d) in enumerate(zip(pB, B, B_dependencies)):
            if tuple(x.shape[-len(d) - 2:]) != tuple(b.shape[-len(d) - 2:]):
                raise ValueError(f"pB[{f}] must have the shape of B[{f}]")
        pB_p, pB_s = zip(*[_ptr_stride(x, len(d) + 2, f"pB[{f}]") for f, (x, d) in enumerate(zip(pB, B_dependencies))])
    C_p, C_s, C_t = [], [], []
    for m, c in enumerate(C):
        if C_time is not None:
            timed = bool(C_time[m])
        else:
            timed = c.dim() == 3 or (c.dim() == 2 and not (int(c.shape[0]) == batch and int(c.shape[1]) == O[m]))
        p, s = _ptr_stride(c, 2 if timed else 1, f"C[{m}]")
        if int(c.shape[-1]) != O[m] or (timed and int(c.shape[-2]) != T):
            raise ValueError(f"C[{m}] must have shape (O_m,) or (T, O_m); got {tuple(c.shape)}")
        C_p.append(p)
        C_s.append(s)
        C_t.append(1 if timed else 0)
    if E is not None:
        # [P] or [1, P]: one E for every agent (stride 0); [batch, P]: per agent
        if E.dtype != torch.float32 or not E.is_cuda or E.device != dev:
            raise ValueError(f"E must be a float32 HIP tensor on {dev}")
        if E.dim() not in (1, 2) or int(E.shape[-1]) != P or (E.dim() == 2 and int(E.shape[0]) not in (1, batch)):
            raise ValueError(f"E has shape {tuple(E.shape)}; expected ({P},), (1, {P}) or ({batch}, {P})")
        E = E.contiguous()
        E_ptr, E_stride = E.data_ptr(), (P if E.dim() == 2 and int(E.shape[0]) == batch and batch > 1 else 0)
    else:
        E_ptr, E_stride = None, 0
    if isinstance(gamma, torch.Tensor):
        g_t = gamma.to(device=dev, dtype=torch.float32).reshape(-1)
        if g_t.numel() not in (1, batch):
            raise ValueError(f"gamma must be a scalar or have one entry per agent ({batch}); got {g_t.numel()}")
        if g_t.numel() != batch:
            g_t = g_t.expand(batch)
        g_t = g_t.contiguous()
        g_ptr, g_sc = g_t.data_ptr(), 1.0
    else:
        g_ptr, g_sc = None, float(gamma)
    I_p = [0] * F
    I_s = [0] * F
    I_d = [1] * F
    eps_ptr, eps_cnt, eps_sc = None, 0, 0.0
    if use_inductive:
        if I is None or len(I) != F:
            raise ValueError("B and I must have one entry per hidden-state factor")
        I_f = []
        for f, i_f in enumerate(I):
            i_f = i_f.to(torch.float32).contiguous()
            I_f.append(i_f)
            if int(i_f.shape[-1]) != S[f]:
                raise ValueError(f"I[{f}] must have shape (depth_f, S_f); got {tuple(i_f.shape)}")
            p, s = _ptr_stride(i_f, 2, f"I[{f}]")
            I_p[f], I_s[f], I_d[f] = p, s, int(i_f.shape[-2])
        I = I_f
        if isinstance(inductive_epsilon, torch.Tensor):
            e_t = inductive_epsilon.to(device=dev, dtype=torch.float32).reshape(-1).contiguous()
            eps_ptr, eps_cnt = e_t.data_ptr(), int(e_t.numel())
        else:
            if float(inductive_epsilon) <= 0.0:
                raise ValueError("inductive_epsilon must be positive when use_inductive=True")
            eps_sc = float(inductive_epsilon)
    G = torch.empty((batch, P), dtype=torch.float32, device=dev)
    q_pi = torch.empty((batch, P), dtype=torch.float32, device=dev) if return_q_pi else None
    Adf, Ado = flatten_dep_lists(A_dependencies)
    Bdf, Bdo = flatten_dep_lists(B_dependencies)
    i64 = {k: _native.I64(v) for k, v in dict(As=A_s, Bs=B_s, Cs=C_s, Ct=C_t, Is=I_s, Id=I_d, S=S, O=O, U=U,
                                               Adf=Adf, Ado=Ado, Bdf=Bdf, Bdo=Bdo, pAs=pA_s or [0],
                                               pBs=pB_s or [0]).items()}
    flags = _flags(use_utility, use_states_info_gain, use_param_info_gain, use_inductive, ffi_cache_params_static)
    if ffi_cache_params_static:
        _static_params_guard(list(A) + list(pA or []) + list(pB or []))
    pA_arr = _native.ptr_array(pA_p) if pA_p is not None else None
    pB_arr = _native.ptr_array(pB_p) if pB_p is not None else None
    rc = _native.lib().pymdp_hip_neg_efe(
        pm.ctypes.data, P, T, qs_flat.data_ptr(),
        _native.ptr_array(A_p), i64["As"].ptr, _native.ptr_array(B_p), i64["Bs"].ptr,
        _native.ptr_array(C_p), i64["Cs"].ptr, i64["Ct"].ptr,
        pA_arr, i64["pAs"].ptr, pB_arr, i64["pBs"].ptr,
        E_ptr, E_stride, g_ptr, g_sc,
        _native.ptr_array(I_p), i64["Is"].ptr, i64["Id"].ptr,
        eps_ptr, eps_cnt, eps_sc,
        G.data_ptr(), q_pi.data_ptr() if q_pi is not None else None,
        batch, i64["S"].ptr, F, i64["O"].ptr, M, i64["U"].ptr,
        i64["Adf"].ptr, i64["Ado"].ptr, i64["Bdf"].ptr, i64["Bdo"].ptr,
        flags, _native.stream_handle(dev))
    _native.check(rc)
    return q_pi, G


def neg_efe_all_policies_ffi(policy_matrix, qs_init, A, B, C, I, A_dependencies, B_dependencies, *,
                             pA=None, pB=None, use_utility=True, use_states_info_gain=True,
                             use_param_info_gain=False, use_inductive=True, inductive_epsilon=1e-3,
                             ffi_cache_params_static=False):
    """pymdp/ffi/_efe.py:491-711: neg_efe of every policy.

    Unbatched inputs (qs_init[f] of shape (S_f,)) return shape (P,); batched
    inputs (leading agent axis on qs_init) return (batch, P)."""
    unbatched = qs_init[0].dim() == 1
    if unbatched:
        qs_init = [q[None] for q in qs_init]
    _, G = update_posterior_policies_ffi(
        policy_matrix, qs_init, A, B, C, None, A_dependencies, B_dependencies, gamma=1.0, I=I,
        inductive_epsilon=inductive_epsilon, use_utility=use_utility, use_states_info_gain=use_states_info_gain,
        use_param_info_gain=use_param_info_gain, use_inductive=use_inductive,
        ffi_cache_params_static=ffi_cache_params_static, return_q_pi=False, pA=pA, pB=pB)
    return G[0] if unbatched else G


def neg_efe_all_policies_packed(policy_matrix, qs_flat: torch.Tensor, A_flat: torch.Tensor, B_flat: torch.Tensor,
                                C_flat: torch.Tensor, I_flat: torch.Tensor, inductive_epsilon: torch.Tensor, *,
                                S, O, U, qs_offsets, A_offsets, B_offsets, C_offsets, I_offsets, I_depths,
                                A_dep_flat, A_dep_offsets, B_dep_flat, B_dep_offsets, flags,
                                pA_flat: torch.Tensor | None = None, pB_flat: torch.Tensor | None = None
                                ) -> torch.Tensor:
    """The reference's packed ABI verbatim (`pymdp_neg_efe_all_policies`,
    pymdp/ffi/registration/xla_register.cc:30-64): agent-major flat buffers
    (pA / pB share A's / B's offsets, absent = empty) plus the 13 int64
    attribute spans and the flags word."""
    pm = _policy_host(policy_matrix)
    P, T, F = pm.shape
    batch = qs_flat.shape[0] if qs_flat.dim() == 2 else 1
    out = torch.empty((batch, P), dtype=torch.float32, device=qs_flat.device)
    a = [_native.I64(x) for x in (S, O, U, qs_offsets, A_offsets, B_offsets, C_offsets, I_offsets, I_depths,
                                  A_dep_flat, A_dep_offsets, B_dep_flat, B_dep_offsets)]
    eps = inductive_epsilon.to(torch.float32).reshape(-1).contiguous()
    pa = pA_flat.to(torch.float32).contiguous() if pA_flat is not None else None
    pb = pB_flat.to(torch.float32).contiguous() if pB_flat is not None else None
    rc = _native.lib().pymdp_hip_neg_efe_all_policies(
        pm.ctypes.data, qs_flat.data_ptr(), A_flat.data_ptr(), B_flat.data_ptr(), C_flat.data_ptr(),
        I_flat.data_ptr(), pa.data_ptr() if pa is not None and pa.numel() else None,
        pa.numel() if pa is not None else 0, pb.data_ptr() if pb is not None and pb.numel() else None,
        pb.numel() if pb is not None else 0, eps.data_ptr(), eps.numel(), out.data_ptr(), batch, P, T,
        a[0].ptr, len(S), a[1].ptr, len(O), a[2].ptr, a[3].ptr, a[4].ptr, a[5].ptr, a[6].ptr, a[7].ptr,
        a[8].ptr, a[9].ptr, a[10].ptr, a[11].ptr, a[12].ptr, int(flags), _native.stream_handle(qs_flat.device))
    _native.check(rc)
    return out


__all__ = [
    "FLAG_USE_UTILITY", "FLAG_USE_STATES_INFO_GAIN", "FLAG_USE_INDUCTIVE", "FLAG_USE_PARAM_INFO_GAIN",
    "FLAG_MODEL_PARAMS_STATIC", "can_handle", "neg_efe_all_policies_ffi", "neg_efe_all_policies_packed",
    "update_posterior_policies_ffi",
]
