"""Shared helpers of the native boundary (mirror of pymdp/ffi/_core.py).

Same names and meaning as the reference: dependency-list packing, ragged
flattening, rank checks.  The XLA registration helpers have no counterpart:
the HIP library is called directly through ctypes (pymdp_amd._native) on the
current torch stream.
"""
from __future__ import annotations

from typing import Sequence

import numpy as np
import torch

from .. import _native

# pymdp/ffi/_core.py:37  (kMaxFfiDependencyRank)
MAX_FFI_DEP_RANK = 8


def is_available() -> bool:
    """True if the HIP library is loadable (pymdp/ffi/_core.py:224-227)."""
    return _native.is_available()


def dep_list_rank_ok(deps: Sequence[int]) -> bool:
    """pymdp/ffi/_core.py:79-82."""
    return 1 <= len(deps) <= MAX_FFI_DEP_RANK


def dep_indices_in_range(deps: Sequence[int], num_factors: int) -> bool:
    """pymdp/ffi/_core.py:85-87."""
    return all(0 <= d < num_factors for d in deps)


def flatten_dep_lists(dep_lists: Sequence[Sequence[int]]) -> tuple[np.ndarray, np.ndarray]:
    """pymdp/ffi/_core.py:264-279: (flat int64, offsets int64 of length N+1)."""
    offsets = np.zeros(len(dep_lists) + 1, dtype=np.int64)
    for i, deps in enumerate(dep_lists):
        offsets[i + 1] = offsets[i] + len(deps)
    flat = np.array([d for deps in dep_lists for d in deps], dtype=np.int64)
    return flat, offsets


def flatten_ragged_f32(arrays: Sequence[torch.Tensor], batch: int | None = None) -> tuple[torch.Tensor, np.ndarray]:
    """pymdp/ffi/_core.py:242-261, batched: with `batch` given every array's
    leading axis is the agent axis and the result is [batch, sum_i size_i]."""
    offsets = np.zeros(len(arrays) + 1, dtype=np.int64)
    chunks = []
    for i, arr in enumerate(arrays):
        a = arr.to(torch.float32)
        a = a.reshape(batch, -1) if batch is not None else a.reshape(-1)
        chunks.append(a)
        offsets[i + 1] = offsets[i] + a.shape[-1]
    flat = torch.cat(chunks, dim=-1).contiguous()
    return flat, offsets


def require_device_tensor(t: torch.Tensor, name: str) -> None:
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise ValueError(f"{name} must be a torch tensor on a HIP device")
