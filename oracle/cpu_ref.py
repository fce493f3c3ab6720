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


def lib(path=None):
    """The restatement's library: the prebuilt x86-64-v3 build by default
    (built here when absent), or the library at `path`."""
    global _LIB
    if _LIB is None or path is not None:
        if path is None:
            path = _HERE / "_build" / "libpymdp_cpu_ref.so"
            if not path.exists():
                subprocess.run(["make", "-C", str(_HERE)], check=True, capture_output=True)
        _LIB = ctypes.CDLL(str(path))
        _LIB.cpu_ref_infer_states.restype = ctypes.c_int
        _LIB.cpu_ref_max_threads.restype = ctypes.c_int
        _LIB.cpu_ref_prepared_floats.restype = ctypes.c_int64
    return _LIB


def use_native() -> str:
    """Builds the restatement with -march=native for the host that runs it
    (`make native`, into a fresh temporary folder: a prebuilt file may come
    from another CPU) and makes it the library of every call that follows.
    Returns the gcc flags used."""
    import tempfile
    out = Path(tempfile.mkdtemp(prefix="pymdp_cpu_ref_")) / "libpymdp_cpu_ref_native.so"
    subprocess.run(["make", "-s", "-C", str(_HERE), "native", f"NATIVE_OUT={out}"], check=True, capture_output=True)
    lib(out)
    return "-O3 -march=native -fopenmp"


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


# ---------------------------------------------------------------------------
# Per-agent FPI + factor-history neg-EFE + q_pi + argmax (cpu_ref_agents)
# ---------------------------------------------------------------------------
_P = ctypes.c_void_p


class _Agent(ctypes.Structure):
    """Mirror of cpu_agent_t in cpu_baseline.c (same field order)."""
    _fields_ = [(n, ctypes.c_int32) for n in ("F", "M", "T", "P", "iters", "do_fpi")] + [
        ("gamma", ctypes.c_float), ("pad", ctypes.c_int32)] + [
        (n, _P) for n in ("S", "U", "O", "adep", "adep_off", "bdep", "bdep_off", "C_t", "A", "B", "C", "D", "E",
                          "obs", "hn", "hpar", "hact", "tn", "tup", "p2t", "qs", "G", "qpi", "best", "ha", "btr")]


def _i32(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.int32))


def _ptrs(arrs):
    return (ctypes.c_void_p * max(1, len(arrs)))(*[x.ctypes.data for x in arrs])


def _history_tables(pm, adeps, bdeps):
    """Distinct action histories of every (level, factor) and distinct
    dependency-history tuples of every (level, modality), with the policy ->
    tuple map (pymdp/ffi/neg_efe/factor_history_tree.h:57-126,
    factor_history_tables.h:187-242, via the oracle's dedup)."""
    from .dedup_ref import factor_histories
    P, T, F = pm.shape
    lv = factor_histories(pm, bdeps)
    hn = _i32([len(lv[t][f][1]) for t in range(T) for f in range(F)])
    hpar = [_i32(lv[t][f][0]) for t in range(T) for f in range(F)]
    hact = [_i32(lv[t][f][1]) for t in range(T) for f in range(F)]
    tn, tup, p2t = [], [], []
    for t in range(T):
        for m in range(len(adeps)):
            idx = np.stack([lv[t][d][2] for d in adeps[m]], 1)
            u, inv = np.unique(idx, axis=0, return_inverse=True)
            tn.append(len(u))
            tup.append(_i32(u))
            p2t.append(_i32(inv.reshape(-1)))
    return hn, hpar, hact, _i32(tn), tup, p2t


class Agents:
    """Host-side state of n agents for cpu_ref_agents: the history tables
    (built once, as the reference memoises them) and the prepared H_A / B
    transposes (computed once, as the reference caches them by content).
    Each agent is a dict with A, B, C, D (NumPy float32, no batch axis), S,
    U, O, policies [P, T, F], A_dependencies, B_dependencies, gamma and,
    for do_fpi, obs (indices) -- or qs (list of per-factor posteriors)."""

    def __init__(self, agents, num_iter=16, do_fpi=True):
        self.keep = []
        self.out = []
        memo = {}
        arr = (_Agent * len(agents))()
        for i, ag in enumerate(agents):
            S, U, O = list(ag["S"]), list(ag["U"]), list(ag["O"])
            F, M = len(S), len(O)
            pm = np.asarray(ag["policies"])
            P, T, _ = pm.shape
            adeps, bdeps = ag["A_dependencies"], ag["B_dependencies"]
            key = (id(ag["policies"]), repr(adeps), repr(bdeps))
            if key not in memo:  # agents sharing a policy matrix share the tables
                memo[key] = (ag["policies"],) + _history_tables(pm, adeps, bdeps)
            _, hn, hpar, hact, tn, tup, p2t = memo[key]
            A = [np.ascontiguousarray(x, dtype=np.float32) for x in ag["A"]]
            B = [np.ascontiguousarray(x, dtype=np.float32) for x in ag["B"]]
            C = [np.ascontiguousarray(x, dtype=np.float32) for x in ag["C"]]
            D = [np.ascontiguousarray(x, dtype=np.float32) for x in ag.get("D") or [np.ones(s, np.float32) / s
                                                                                  for s in S]]
            C_t = _i32([1 if c.ndim == 2 else 0 for c in C])
            E = ag.get("E")
            E = None if E is None else np.ascontiguousarray(E, dtype=np.float32)
            qs = np.zeros(sum(S), np.float32)
            if not do_fpi:
                qs[:] = np.concatenate([np.asarray(q, dtype=np.float32).reshape(-1) for q in ag["qs"]])
            G = np.zeros(P, np.float32)
            qpi = np.zeros(P, np.float32)
            best = np.zeros(1, np.int32)
            ints = dict(S=_i32(S), U=_i32(U), O=_i32(O), adep=_i32([d for dl in adeps for d in dl]),
                        adep_off=_i32(np.concatenate([[0], np.cumsum([len(d) for d in adeps])])),
                        bdep=_i32([d for dl in bdeps for d in dl]),
                        bdep_off=_i32(np.concatenate([[0], np.cumsum([len(d) for d in bdeps])])),
                        C_t=C_t, obs=_i32(ag.get("obs", [0] * M)), hn=hn, tn=tn)
            ptab = dict(A=_ptrs(A), B=_ptrs(B), C=_ptrs(C), D=_ptrs(D), hpar=_ptrs(hpar), hact=_ptrs(hact),
                        tup=_ptrs(tup), p2t=_ptrs(p2t))
            a = arr[i]
            a.F, a.M, a.T, a.P, a.iters, a.do_fpi = F, M, T, P, int(num_iter), int(bool(do_fpi))
            a.gamma = float(ag["gamma"])
            for k, v in ints.items():
                setattr(a, k, v.ctypes.data)
            for k, v in ptab.items():
                setattr(a, k, ctypes.cast(v, ctypes.c_void_p).value)
            a.E = E.ctypes.data if E is not None else None
            a.qs, a.G, a.qpi, a.best = qs.ctypes.data, G.ctypes.data, qpi.ctypes.data, best.ctypes.data
            n = int(lib().cpu_ref_prepared_floats(ctypes.byref(a)))
            prep = np.zeros(max(n, 1), np.float32)
            nha = sum(int(np.prod([S[d] for d in dl])) for dl in adeps)
            a.ha = prep.ctypes.data
            a.btr = prep.ctypes.data + 4 * nha
            if lib().cpu_ref_prepare(ctypes.byref(a)) != 0:
                raise ValueError("cpu_ref_prepare failed")
            self.keep.append((A, B, C, D, E, hpar, hact, tup, p2t, ints, ptab, prep))
            self.out.append(dict(qs=qs, G=G, q_pi=qpi, best=best, S=S))
        self.arr = arr
        self.n = len(agents)

    def run(self, nthreads=0):
        if lib().cpu_ref_agents(self.arr, ctypes.c_int64(self.n), ctypes.c_int32(nthreads)) != 0:
            raise ValueError("cpu_ref_agents: unsupported agent")
        return self.out
