"""ctypes wrapper of oracle/_build/libpymdp_cpu_ref.so (the C/OpenMP
restatement in cpu_baseline.c).  Test infrastructure only."""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = _HERE / "_build" / "libpymdp_cpu_ref.so"
        if not path.exists():
            subprocess.run(["make", "-C", str(_HERE)], check=True, capture_output=True)
        _LIB = ctypes.CDLL(str(path))
        _LIB.cpu_ref_infer_states.restype = ctypes.c_int
        _LIB.cpu_ref_max_threads.restype = ctypes.c_int
    return _LIB


def _i64(x):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.int64))
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


def infer_states(A, obs_idx, prior, A_dependencies, num_iter=16, nthreads=0):
    """A[m] float32 [batch, O_m, S_deps...] (or shared without batch axis),
    obs_idx int [batch, M], prior[f] float32 [batch, S_f] -> q [batch, sum S]."""
    A = [np.ascontiguousarray(a, dtype=np.float32) for a in A]
    prior = [np.ascontiguousarray(p, dtype=np.float32) for p in prior]
    obs = np.ascontiguousarray(np.asarray(obs_idx, dtype=np.int32))
    batch = obs.shape[0]
    S = [p.shape[-1] for p in prior]
    O = [a.shape[-len(d) - 1] for a, d in zip(A, A_dependencies)]
    strides = [0 if a.ndim == len(d) + 1 else int(np.prod(a.shape[1:])) for a, d in zip(A, A_dependencies)]
    flat = [d for dl in A_dependencies for d in dl]
    offs = np.concatenate([[0], np.cumsum([len(d) for d in A_dependencies])])
    Ap = (ctypes.c_void_p * len(A))(*[a.ctypes.data for a in A])
    Pp = (ctypes.c_void_p * len(prior))(*[p.ctypes.data for p in prior])
    q = np.empty((batch, int(sum(S))), dtype=np.float32)
    pstride = [p.shape[-1] if p.ndim == 2 else 0 for p in prior]
    keep = [_i64(x) for x in (strides, pstride, S, O, flat, offs)]
    (_, a_str), (_, p_str), (_, s_p), (_, o_p), (_, f_p), (_, d_p) = keep
    rc = lib().cpu_ref_infer_states(
        Ap, a_str, obs.ctypes.data_as(ctypes.c_void_p), Pp, p_str, q.ctypes.data_as(ctypes.c_void_p),
        ctypes.c_int64(batch), s_p, ctypes.c_int64(len(S)), o_p, ctypes.c_int64(len(A)), f_p, d_p,
        ctypes.c_int32(num_iter), ctypes.c_int32(nthreads))
    if rc != 0:
        raise ValueError("cpu_ref_infer_states: unsupported shape")
    return q


def max_threads() -> int:
    return int(lib().cpu_ref_max_threads())
