"""ctypes binding of libpymdp_hip.so (the C ABI declared in include/pymdp_hip.h).

The library is required: there is no fallback implementation.  Loading fails
with a RuntimeError that says how to build it.  Mirrors the role of the
reference's pymdp/ffi/_core.py:41-107 (library path, ctypes.CDLL, symbol
lookup), minus the XLA registration which has no counterpart without JAX.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path
from typing import Sequence

import numpy as np

_LIB: ctypes.CDLL | None = None
_LOAD_ERROR: Exception | None = None

PYMDP_HIP_OK = 0
PYMDP_HIP_EINVAL = 1
PYMDP_HIP_EDEVICE = 2

_i64p = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p

# Exported symbol -> (restype, argtypes).  Must match include/pymdp_hip.h.
SIGNATURES: dict[str, tuple] = {
    "pymdp_hip_last_error": (ctypes.c_char_p, []),
    "pymdp_hip_version": (ctypes.c_char_p, []),
    "pymdp_hip_last_launch": (ctypes.c_char_p, []),
    "pymdp_hip_release_scratch": (ctypes.c_int, []),
    "pymdp_hip_scratch_bytes": (ctypes.c_int64, [_i64p]),
    "pymdp_hip_param_cache_clear": (ctypes.c_int, []),
    "pymdp_hip_fpi": (ctypes.c_int, [
        _vp, _vp, _vp,                       # ll, lp, q
        ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
        _i64p, ctypes.c_int64,               # S, F
        _i64p, ctypes.c_int64,               # ll_offsets, M
        _i64p,                               # lp_offsets
        _i64p, _i64p,                        # A_dep_flat, A_dep_offsets
        ctypes.c_int32, _vp]),
    "pymdp_hip_infer_states": (ctypes.c_int, [
        _vp, _i64p,                          # A_ptrs, A_batch_stride
        _vp, ctypes.c_int32,                 # obs, obs_is_index
        _vp, _i64p,                          # prior_ptrs, prior_batch_stride
        _vp, ctypes.c_int64,                 # q_out, batch
        _i64p, ctypes.c_int64,               # S, F
        _i64p, ctypes.c_int64,               # O, M
        _i64p, _i64p,                        # obs_offsets, lp_offsets
        _i64p, _i64p,                        # A_dep_flat, A_dep_offsets
        ctypes.c_int32, ctypes.c_int32, _vp]),
    "pymdp_hip_infer_states_noalloc": (ctypes.c_int, [
        _vp, _i64p, _vp, ctypes.c_int32, _vp, _i64p, _vp, ctypes.c_int64,
        _i64p, ctypes.c_int64, _i64p, ctypes.c_int64, _i64p, _i64p, _i64p, _i64p,
        ctypes.c_int32, ctypes.c_int32, _vp]),
    "pymdp_hip_fpi_vfe": (ctypes.c_int, [
        _vp, _i64p,                          # A_ptrs, A_batch_stride
        _vp, ctypes.c_int32,                 # obs, obs_is_index
        _vp, _i64p,                          # prior_ptrs, prior_batch_stride
        _vp, _vp, ctypes.c_int64,            # qs, vfe_out, batch
        _i64p, ctypes.c_int64,               # S, F
        _i64p, ctypes.c_int64,               # O, M
        _i64p, _i64p, _i64p,                 # obs_offsets, A_dep_flat, A_dep_offsets
        _vp]),
    "pymdp_hip_sequence_infer": (ctypes.c_int, [
        ctypes.c_int32,                      # method
        _vp, _i64p, _vp, _i64p,              # A ptrs/stride, B ptrs/stride
        _vp, ctypes.c_int32,                 # obs, obs_is_index
        _vp, _i64p,                          # prior ptrs/stride
        _vp, _vp, _vp,                       # actions, valid_steps, q_out
        ctypes.c_int64, ctypes.c_int64,      # batch, T
        _i64p, _i64p, ctypes.c_int64,        # S, U, F
        _i64p, ctypes.c_int64, _i64p,        # O, M, obs_offsets
        _i64p, _i64p,                        # A_dep_flat, A_dep_offsets
        ctypes.c_int32, ctypes.c_float, _vp]),
    "pymdp_hip_neg_efe": (ctypes.c_int, [
        _vp, ctypes.c_int64, ctypes.c_int64,     # policy_host, P, T
        _vp,                                     # qs_flat
        _vp, _i64p, _vp, _i64p,                  # A ptrs/stride, B ptrs/stride
        _vp, _i64p, _i64p,                       # C ptrs/stride, C_has_time
        _vp, _i64p, _vp, _i64p,                  # pA ptrs/stride, pB ptrs/stride (or NULL)
        _vp, ctypes.c_int64,                     # E, E stride
        _vp, ctypes.c_float,                     # gamma, gamma_scalar
        _vp, _i64p, _i64p,                       # I ptrs/stride/depths
        _vp, ctypes.c_int64, ctypes.c_float,     # eps ptr, count, scalar
        _vp, _vp,                                # G_out, q_pi_out
        ctypes.c_int64,                          # batch
        _i64p, ctypes.c_int64, _i64p, ctypes.c_int64, _i64p,  # S, F, O, M, U
        _i64p, _i64p, _i64p, _i64p,              # A deps, B deps
        ctypes.c_int32, _vp]),
    "pymdp_hip_neg_efe_all_policies": (ctypes.c_int, [
        _vp, _vp, _vp, _vp, _vp, _vp,            # policy, qs, A, B, C, I
        _vp, ctypes.c_int64, _vp, ctypes.c_int64,  # pA, pA count, pB, pB count
        _vp, ctypes.c_int64,                     # eps, eps count
        _vp, ctypes.c_int64,                     # out, batch
        ctypes.c_int64, ctypes.c_int64,          # P, T
        _i64p, ctypes.c_int64, _i64p, ctypes.c_int64, _i64p,  # S, F, O, M, U
        _i64p, _i64p, _i64p, _i64p, _i64p, _i64p,            # qs/A/B/C/I offsets, I_depths
        _i64p, _i64p, _i64p, _i64p,              # A deps, B deps
        ctypes.c_int32, _vp]),
    "pymdp_hip_fpi_adjoint": (ctypes.c_int, [
        _vp, _i64p, _vp, _vp, _i64p,             # A ptrs/stride, obs, prior ptrs/stride
        _vp, _vp, _vp, _vp, _vp,                 # q_bar, A_bar ptrs, prior_bar, q_out, mode_out
        ctypes.c_int64, _i64p, ctypes.c_int64, _i64p, ctypes.c_int64,  # batch, S, F, O, M
        _i64p, _i64p,                            # A deps
        ctypes.c_int32, ctypes.c_int32, ctypes.c_float, _vp]),
    "pymdp_hip_neg_efe_adjoint": (ctypes.c_int, [
        _vp, ctypes.c_int64, ctypes.c_int64, _vp,  # policy (device), P, T, qs
        _vp, _i64p, _vp, _i64p, _vp, _i64p,      # A, B, C ptrs/strides
        _vp, _vp, _vp, _vp, _vp,                 # G_bar, qs_bar, A_bar, B_bar, C_bar
        ctypes.c_int64, _i64p, ctypes.c_int64, _i64p, ctypes.c_int64, _i64p,  # batch, S, F, O, M, U
        _i64p, _i64p, _i64p, _i64p,              # A deps, B deps
        ctypes.c_int32, _vp]),
    # Dirichlet learning (pymdp/learning.py)
    "pymdp_hip_update_pA": (ctypes.c_int, [
        _vp, ctypes.c_int64, _vp, _vp,           # pA, stride, qA, EA
        ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,  # batch, T, O
        ctypes.c_int32, _i64p,                   # n_deps, dep_sizes
        _vp, ctypes.c_int64,                     # obs, obs batch stride
        _vp, _i64p,                              # qs_deps ptrs, batch strides
        _vp, ctypes.c_float, _vp]),              # lr, lr_scalar, stream
    "pymdp_hip_update_pB": (ctypes.c_int, [
        _vp, ctypes.c_int64, _vp, _vp,           # pB, stride, qB, EB
        ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,  # batch, T1, S
        ctypes.c_int32, _i64p, ctypes.c_int64,   # n_deps, dep_sizes, U
        _vp, ctypes.c_int64, ctypes.c_int64,     # qs_next, batch / t strides
        _vp, _i64p, _i64p,                       # qs_deps ptrs, batch / t strides
        _vp, ctypes.c_int64, ctypes.c_int32,     # actions, batch stride, num_factors
        ctypes.c_int32, ctypes.c_int32,          # factor, mask_all
        _vp, ctypes.c_float, _vp]),              # lr, lr_scalar, stream
    # batched ragged agents (no counterpart in the reference)
    "pymdp_hip_ragged_plan_create": (ctypes.c_int, [
        ctypes.c_int64, _i64p,                   # n_agents, hdr [n, 4]
        _i64p, _i64p, _i64p,                     # S, U, O
        _i64p, _i64p, _i64p, _i64p,              # A dep rank/flat, B dep rank/flat
        _i64p, _vp, _vp, _vp,                    # C_time, E_present, gamma, policies
        ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]),
    "pymdp_hip_ragged_plan_upload": (ctypes.c_int, [_vp]),
    "pymdp_hip_ragged_layout": (ctypes.c_int, [_vp, _i64p, ctypes.c_int64]),
    "pymdp_hip_ragged_info": (ctypes.c_int, [_vp, _i64p]),
    "pymdp_hip_ragged_tables": (ctypes.c_int, [_vp, _i64p, _vp, _vp, _vp, _vp]),
    "pymdp_hip_ragged_e1_records": (ctypes.c_int, [_vp, _vp, _i64p, _i64p]),
    "pymdp_hip_ragged_e2_records": (ctypes.c_int, [_vp, _vp, _i64p, _i64p]),
    "pymdp_hip_ragged_e3_records": (ctypes.c_int, [_vp, _vp, _i64p, _i64p]),
    "pymdp_hip_ragged_roll_tasks": (ctypes.c_int, [_vp, _vp, _i64p, _i64p, _i64p]),
    "pymdp_hip_ragged_set_params_static": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "pymdp_hip_ragged_tile_config": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int32)]),
    "pymdp_hip_ragged_set_tiles": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int32)]),
    "pymdp_hip_ragged_step": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "pymdp_hip_ragged_autotune": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(ctypes.c_float)]),
    "pymdp_hip_ragged_status": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int32)]),
    "pymdp_hip_ragged_set_prior_source": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "pymdp_hip_ragged_set_trace": (ctypes.c_int, [_vp, _vp]),
    "pymdp_hip_fpi_set_trace": (ctypes.c_int, [_vp]),
    "pymdp_hip_ragged_plan_destroy": (ctypes.c_int, [_vp]),
}


def library_path() -> Path:
    override = os.environ.get("PYMDP_HIP_LIB")
    if override:
        return Path(override)
    return Path(__file__).resolve().parent / "_lib" / "libpymdp_hip.so"


def lib() -> ctypes.CDLL:
    """Load (once) and return the HIP library; raise if it is missing."""
    global _LIB, _LOAD_ERROR
    if _LIB is not None:
        return _LIB
    path = library_path()
    if not path.exists():
        raise RuntimeError(
            f"pymdp_amd: HIP library not found at {path}. Build it with `python -m pymdp_amd.build`."
        )
    try:
        handle = ctypes.CDLL(str(path))
    except OSError as exc:  # pragma: no cover - surfaced to the caller
        _LOAD_ERROR = exc
        raise RuntimeError(f"pymdp_amd: cannot load {path}: {exc}") from exc
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(handle, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = handle
    return _LIB


def is_available() -> bool:
    try:
        lib()
        return True
    except RuntimeError:
        return False


def check(rc: int) -> None:
    """Map a C-ABI status code to a Python exception (ValueError for
    kInvalidArgument, RuntimeError for device errors)."""
    if rc == PYMDP_HIP_OK:
        return
    msg = lib().pymdp_hip_last_error().decode(errors="replace")
    if rc == PYMDP_HIP_EINVAL:
        raise ValueError(msg)
    raise RuntimeError(msg)


class I64:
    """Host int64 array kept alive for the duration of a call."""

    __slots__ = ("arr",)

    def __init__(self, values: Sequence[int] | np.ndarray):
        self.arr = np.ascontiguousarray(np.asarray(values, dtype=np.int64).reshape(-1))

    @property
    def ptr(self):
        return self.arr.ctypes.data_as(_i64p)


def ptr_array(ptrs: Sequence[int]):
    arr = (ctypes.c_void_p * max(1, len(ptrs)))(*[int(p) for p in ptrs])
    return arr


def stream_handle(device=None) -> int:
    """The current torch HIP stream of `device` as a raw hipStream_t value."""
    import torch

    return int(torch.cuda.current_stream(device).cuda_stream)
