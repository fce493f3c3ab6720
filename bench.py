"""Benchmark of the FPI + EFE hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Workloads (seeded synthetic data, generated on the device):
  headline (default)         the ragged batched step below (value, roofline,
                             plain per-agent comparison, CPU baseline), with
                             the plain_fpi_large, plain_efe and grad results
                             attached as sub-objects of the same JSON line.
  plain_fpi_large (default)  1024 agents, S=[16,16,8,8], 6 modalities x 32
                             outcomes, A Dirichlet(0.5), 16 sweeps: one step =
                             one fused infer_states launch for all agents.
  plain_efe                  64 agents, S=[32,16,8], U=[8,4,1], horizon 3
                             (32768 policies), 4 x 64 outcomes: one step =
                             one infer_policies (neg-EFE + q_pi) for all agents.
  grad                       512 agents, S=[16,8], U=[16,4], horizon 1, 3 x 16
                             outcomes, 32-step trajectories: one step = the
                             gradient of sum_t ln q_pi(a_t) w.r.t. the logits of
                             A, B, C, D through the HIP adjoints, compared with
                             torch autograd through the pure-torch forward.
  learning                   1024 agents, Dirichlet learning of A on the
                             plain_fpi_large model (6 x [32,16,16,8,8] counts
                             per agent, one time step): one step = the fused
                             pA -> (qA, E[qA]) update of every modality.
  ragged                     4096 agents of different shapes (1-4 factors of
                             2-64 states, 1-8 controls, 1-5 modalities of 2-64
                             outcomes, horizon 1-2): one step = infer_states ->
                             infer_policies -> action for every agent, one HIP
                             graph replay of the batched kernels; compared with
                             the plain per-agent entry points on a sample.

Prints ONE JSON line with value = agent-steps/s over the K timed steps,
the roofline of the dominant kernel (HIP events on the launch stream) and a
CPU baseline timed on a bounded sample on this host.

    python bench.py --calibrate-copy   # plain device copy of 1 GiB, for the
                                       # FETCH_SIZE / WRITE_SIZE scale factor
    python bench.py --write-traffic KEY --kernel SUBSTR \
        --pmc-dirs FETCH_DIR WRITE_DIR COPY_FETCH_DIR COPY_WRITE_DIR
                                       # HBM bytes per launch of the kernel
                                       # whose name contains SUBSTR (comma-
                                       # separated: summed; SUBSTR:N sums
                                       # its last N dispatches), from four
                                       # rocprofv3 --pmc csv runs, scaled by
                                       # the copy calibration; stored under KEY
                                       # in profiles/traffic.json
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6
CPU_RUNS, CPU_WARM = 20, 3


def _cpu_threads():
    return len(os.sched_getaffinity(0))


_CPU_FLAGS = None


def _cpu_native():
    """Builds the oracle's C/OpenMP restatement with -march=native for this
    host (once per process) and returns the flags."""
    global _CPU_FLAGS
    if _CPU_FLAGS is None:
        from oracle import cpu_ref
        _CPU_FLAGS = cpu_ref.use_native()
    return _CPU_FLAGS


def _log(msg):
    """Progress line on stderr (the JSON result is the only stdout line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _median_runs(fn, runs=CPU_RUNS, warm=CPU_WARM):
    """Median wall time (s) of `runs` calls of fn after `warm` untimed ones."""
    for _ in range(warm):
        fn()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    ts.sort()
    n = len(ts)
    return ts[n // 2] if n % 2 else 0.5 * (ts[n // 2 - 1] + ts[n // 2])


def _args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="headline",
                    choices=["headline", "plain_fpi_small", "plain_fpi_large", "plain_efe", "ragged", "grad",
                             "learning"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--write-traffic", metavar="KEY")
    ap.add_argument("--kernel", default="fpi_fast_kernel")
    ap.add_argument("--pmc-dirs", nargs=4)
    return ap.parse_args()


COPY_BYTES = 4 << 28  # bytes read (and written) by calibrate_copy


def _pmc_rows(d: str, counter: str):
    """(kernel name, value) per dispatch for `counter` from a rocprofv3 csv dir."""
    import csv
    import glob
    rows = []
    for path in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        with open(path, newline="") as fh:
            for r in csv.DictReader(fh):
                if r.get("Counter_Name") == counter:
                    rows.append((r.get("Kernel_Name", ""), float(r["Counter_Value"])))
    return rows


def write_traffic(a):
    """HBM bytes per launch = FETCH_SIZE + WRITE_SIZE (KiB per dispatch, last
    dispatch of the named kernel), each scaled by bytes / counter-reading of a
    plain 1 GiB device copy measured the same way."""
    fdir, wdir, cfdir, cwdir = a.pmc_dirs

    def last(d, counter, spec):
        # "name" = the last dispatch of the kernel, "name:N" = its last N
        # dispatches summed (a kernel launched N times per step)
        sub, _, n = spec.partition(":")
        n = int(n) if n else 1
        vals = [v for k, v in _pmc_rows(d, counter) if sub in k]
        if len(vals) < n:
            raise SystemExit(f"fewer than {n} {counter} rows for kernels matching {sub!r} under {d}")
        return sum(vals[-n:])

    def biggest(d, counter):
        vals = [v for _, v in _pmc_rows(d, counter)]
        if not vals:
            raise SystemExit(f"no {counter} rows under {d}")
        return max(vals)

    subs = [k for k in a.kernel.split(",") if k]
    fetch_kib = sum(last(fdir, "FETCH_SIZE", k) for k in subs)
    write_kib = sum(last(wdir, "WRITE_SIZE", k) for k in subs)
    cf_kib, cw_kib = biggest(cfdir, "FETCH_SIZE"), biggest(cwdir, "WRITE_SIZE")
    sf, sw = COPY_BYTES / (cf_kib * 1024.0), COPY_BYTES / (cw_kib * 1024.0)
    traffic = fetch_kib * 1024.0 * sf + write_kib * 1024.0 * sw
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as fh:
            db = json.load(fh)
    except (OSError, ValueError):
        db = {}
    db[a.write_traffic] = traffic
    db[a.write_traffic + " (detail)"] = {
        "kernel": a.kernel, "FETCH_SIZE_KiB": fetch_kib, "WRITE_SIZE_KiB": write_kib,
        "copy_FETCH_SIZE_KiB": cf_kib, "copy_WRITE_SIZE_KiB": cw_kib, "fetch_scale": sf, "write_scale": sw,
        "copy_bytes_each_way": COPY_BYTES}
    with open(path, "w") as fh:
        json.dump(db, fh, indent=1, sort_keys=True)
    print(json.dumps({a.write_traffic: traffic}))


def _traffic(key: str):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as fh:
            return json.load(fh).get(key)
    except (OSError, ValueError):
        return None


def calibrate_copy():
    import torch
    n = 1 << 28  # 1 GiB of float32
    x = torch.ones(n, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(10):
        y.copy_(x)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / 10
    print(json.dumps({"calibration": "device_copy", "bytes_read": 4 * n, "bytes_written": 4 * n,
                      "ms": dt * 1e3, "GBps": 8 * n / dt / 1e9}))


def bench_plain_fpi_small(a, torch):
    """64 agents, S=[4,3], O=[5,6], 16 sweeps (the reference's CPU-runnable
    case): one step = one fused infer_states launch for all agents."""
    import numpy as np

    from pymdp_amd import workloads
    from pymdp_amd.ffi import infer_states_ffi

    dev = torch.device("cuda:0")
    w = workloads.plain_fpi_small(seed=0, device=dev)
    B, S, O, deps = w["batch"], w["S"], w["O"], w["A_dependencies"]
    obs = torch.stack([o.to(torch.int32) for o in w["obs"]], -1).contiguous()
    out = torch.empty((B, sum(S)), device=dev)

    def step():
        return infer_states_ffi(w["A"], obs, w["D"], deps, w["num_iter"], distr_obs=False, out=out)

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(a.steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    ms_step = (time.perf_counter() - t0) * 1e3 / a.steps
    ms_events = e0.elapsed_time(e1) / a.steps
    from pymdp_amd import _native
    kernel = _native.lib().pymdp_hip_last_launch().decode()
    K = int(np.prod(S))
    alg_bytes = B * (len(O) * K * 4 + len(O) * 4 + 2 * sum(S) * 4)
    achieved = alg_bytes / (ms_events * 1e-3) / 1e9
    from oracle import numpy_ref as R
    q = step()
    torch.cuda.synchronize()
    e64 = 0.0
    for b in range(B):
        A_b = [x[b].cpu().numpy() for x in w["A"]]
        ob = [np.eye(o, dtype=np.float32)[int(x[b])] for o, x in zip(O, w["obs"])]
        r64 = R.run_factorized_fpi(A_b, ob, [d[b].cpu().numpy() for d in w["D"]], deps, w["num_iter"], np.float64)
        e64 = max(e64, max(float(np.abs(q[f][b].cpu().numpy() - r64[f]).max()) for f in range(len(S))))
    cpu = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline_fpi(w, B)
    return {
        "metric": "agent-steps/s (plain FPI, 64 agents, S=[4,3], O=[5,6], 16 sweeps)",
        "value": B / (ms_step * 1e-3), "unit": "agent-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 data, f64 sweep accumulation",
        "data": "synthetic (seeded Dirichlet(1) A, uniform D, uniform observation indices; generated on device)",
        "config": {"workload": "plain_fpi_small", "agents": B, "num_states": S, "num_obs": O,
                   "num_iter": w["num_iter"], "obs": "index (one-hot)"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": kernel, "kernel_ms": ms_events,
                     "alg_bytes_per_launch": alg_bytes,
                     "note": "one launch of 64 one-wave agents: the launch latency, not HBM, sets the time"},
        "parity": {"agents_checked": B, "max_abs_err_vs_fp64": e64},
        "cpu_baseline": cpu,
    }


def bench_plain_fpi_large(a, torch):
    import numpy as np

    from pymdp_amd import workloads
    from pymdp_amd.ffi import infer_states_ffi

    dev = torch.device("cuda:0")
    w = workloads.plain_fpi_large(seed=0, device=dev)
    B, S, O, deps = w["batch"], w["S"], w["O"], w["A_dependencies"]
    obs = torch.stack([o.to(torch.int32) for o in w["obs"]], -1).contiguous()
    out = torch.empty((B, sum(S)), device=dev)
    torch.cuda.synchronize()

    def step():
        return infer_states_ffi(w["A"], obs, w["D"], deps, w["num_iter"], distr_obs=False, out=out)

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(a.steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_events = e0.elapsed_time(e1) / a.steps
    ms_step = wall * 1e3 / a.steps
    from pymdp_amd import _native
    kernel = _native.lib().pymdp_hip_last_launch().decode()
    # algorithmic bytes of one launch: observed A rows + obs indices + prior in + q out
    K = int(np.prod(S))
    per_agent = len(O) * K * 4 + len(O) * 4 + 2 * sum(S) * 4
    alg_bytes = B * per_agent
    achieved = alg_bytes / (ms_events * 1e-3) / 1e9

    # parity on sampled agents against the oracle (FP32 reference path and FP64)
    from oracle import numpy_ref as R
    q = step()
    torch.cuda.synchronize()
    e32 = e64 = 0.0
    for b in (0, B // 2, B - 1):
        A_b = [x[b].cpu().numpy() for x in w["A"]]
        ob = [np.eye(o, dtype=np.float32)[int(x[b])] for o, x in zip(O, w["obs"])]
        D_b = [d[b].cpu().numpy() for d in w["D"]]
        r32 = R.run_factorized_fpi(A_b, ob, D_b, deps, w["num_iter"], np.float32)
        r64 = R.run_factorized_fpi(A_b, ob, D_b, deps, w["num_iter"], np.float64)
        for f in range(len(S)):
            g = q[f][b].cpu().numpy().astype(np.float64)
            e32 = max(e32, float(np.abs(g - r32[f]).max()))
            e64 = max(e64, float(np.abs(g - r64[f]).max()))

    # the latency floor: one more launch with the kernel's timing trace
    # (per agent: start, log-likelihood rows formed, end of each sweep, end;
    # wall-clock ticks of 100 MHz), outside the timed region
    import ctypes
    it = w["num_iter"]
    trbuf = torch.zeros(B * (it + 3), dtype=torch.int64, device=dev)
    _native.check(_native.lib().pymdp_hip_fpi_set_trace(ctypes.c_void_p(trbuf.data_ptr())))
    try:
        step()
        torch.cuda.synchronize()
    finally:
        _native.check(_native.lib().pymdp_hip_fpi_set_trace(None))
    t = trbuf.cpu().numpy().reshape(B, it + 3).astype(np.float64) / 100.0  # us
    t -= t[:, 0].min()
    sweeps = np.diff(t[:, 1:2 + it], axis=1)
    starts = np.sort(t[:, 0])
    sweep_trace = {
        "unit": "us", "kernel_span": float(t[:, 2 + it].max()),
        "loglik_rows_mean": float((t[:, 1] - t[:, 0]).mean()),
        "sweep_mean": float(sweeps.mean()), "sweep_median": float(np.median(sweeps)),
        "sweep_min": float(sweeps.min()), "sweep_max": float(sweeps.max()),
        "agent_mean": float((t[:, 2 + it] - t[:, 0]).mean()),
        "agents_started_within_1us": int((starts < starts[0] + 1.0).sum()),
        "how": "pymdp_hip_fpi_set_trace on one launch after the timed region",
    }

    cpu = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline_fpi(w)
    res = {
        "metric": "agent-steps/s (plain FPI, 1024 agents, S=[16,16,8,8], 6x32 outcomes, 16 sweeps)",
        "value": B / (ms_step * 1e-3),
        "unit": "agent-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 data, f64 sweep accumulation",
        "data": "synthetic (seeded Dirichlet A, sampled observations; generated on device)",
        "config": {"workload": "plain_fpi_large", "agents": B, "num_states": S, "num_obs": O,
                   "num_iter": w["num_iter"], "obs": "index (one-hot)"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": _traffic("fpi plain_fpi_large"),
                     "kernel": kernel, "kernel_ms": ms_events, "alg_bytes_per_launch": alg_bytes},
        "parity": {"agents_checked": 3, "max_abs_err_vs_fp32_ref": e32, "max_abs_err_vs_fp64": e64},
        "sweep_trace": sweep_trace,
        "cpu_baseline": cpu,
    }
    return res


def bench_learning(a, torch):
    """Dirichlet learning of A (pymdp/learning.py:57-117): HBM-bound, 12 bytes
    per count (pA read, qA and E[qA] written)."""
    import numpy as np
    from pymdp_amd import learning, workloads

    dev = torch.device("cuda:0")
    w = workloads.learning_large(seed=0, device=dev)
    B, S, O, deps = w["batch"], w["S"], w["O"], w["A_dependencies"]
    torch.cuda.synchronize()
    out = {}

    def step():
        qA, EA = learning.update_obs_likelihood_dirichlet(
            w["pA"], [None] * len(O), w["obs"], w["qs"], A_dependencies=deps, categorical_obs=False, num_obs=O,
            lr=1.0)
        out["r"] = (qA, EA)

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(a.steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_events = e0.elapsed_time(e1) / a.steps
    ms_step = wall * 1e3 / a.steps
    counts = sum(int(p.numel()) for p in w["pA"])
    alg_bytes = 12 * counts  # per step: six launches, one per modality
    achieved = alg_bytes / (ms_events * 1e-3) / 1e9
    # parity: two agents of every modality against the FP64 oracle
    from oracle import numpy_ref as R
    qA, EA = out["r"]
    eq = ee = 0.0
    for b in (0, B - 1):
        qs_b = [q[b].cpu().numpy() for q in w["qs"]]
        for m in range(len(O)):
            ob = np.eye(O[m])[w["obs"][m][b].cpu().numpy()]
            q64, e64 = R.update_obs_likelihood_dirichlet_m(w["pA"][m][b].cpu().numpy(), ob, qs_b, deps[m], 1.0,
                                                           np.float64)
            eq = max(eq, float(np.max(np.abs(qA[m][b].cpu().numpy() - q64) / np.maximum(1.0, np.abs(q64)))))
            ee = max(ee, float(np.abs(EA[m][b].cpu().numpy() - e64).max()))
    cpu = None
    if not a.no_cpu_baseline:
        # NumPy FP32 restatement on a bounded sample of agents (single thread)
        n, t0 = 0, time.perf_counter()
        pA_h = [p[:8].cpu().numpy() for p in w["pA"]]
        qs_h = [q[:8].cpu().numpy() for q in w["qs"]]
        obs_h = [o[:8].cpu().numpy() for o in w["obs"]]
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < min(a.cpu_seconds, 10.0):
            b = n % 8
            R.update_obs_likelihood_dirichlet([p[b] for p in pA_h], [None] * len(O), [o[b] for o in obs_h],
                                              [q[b] for q in qs_h], deps, False, O, 1.0, np.float32)
            n += 1
        el = time.perf_counter() - t0
        cpu = {"value": n / el, "unit": "agent-updates/s", "cores": 1, "kind": "port",
               "sample": f"{n} agent updates over the first 8 agents, {el:.1f} s, NumPy FP32 restatement "
                         f"of pymdp/learning.py (oracle), single thread"}
    return {
        "metric": "agent-updates/s (Dirichlet learning of A, 1024 agents, 6 x [32,16,16,8,8] counts)",
        "value": B / (ms_step * 1e-3),
        "unit": "agent-updates/s",
        "ms_per_step": ms_step,
        "steps": a.steps,
        "config": {"workload": "learning_large", "agents": B, "num_states": S, "num_obs": O, "time_steps": 1,
                   "counts_per_step": counts},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": _traffic("learning pA"),
                     "kernel": "dirichlet_kernel (6 launches)", "kernel_ms": ms_events,
                     "alg_bytes_per_launch": alg_bytes // len(O)},
        "parity": {"agents_checked": 2, "max_rel_err_counts_vs_fp64": eq, "max_abs_err_expected_vs_fp64": ee},
        "cpu_baseline": cpu,
    }


def cpu_baseline_fpi(w, n=256):
    """The oracle's C/OpenMP restatement of the reference CPU FFI FPI
    (-march=native, OpenMP over agents on every core of this host) on the
    first n agents: median of 20 runs after 3 warm-ups."""
    import numpy as np

    from oracle import cpu_ref
    flags = _cpu_native()
    n = min(n, int(w["batch"]))
    A = [x[:n].cpu().numpy() for x in w["A"]]
    D = [d[:n].cpu().numpy() for d in w["D"]]
    obs = np.stack([o[:n].cpu().numpy() for o in w["obs"]], 1).astype(np.int32)
    threads = _cpu_threads()
    t = _median_runs(lambda: cpu_ref.infer_states(A, obs, D, w["A_dependencies"], w["num_iter"], threads))
    return {"value": n / t, "unit": "agent-steps/s", "cores": threads, "kind": "port",
            "sample": f"first {n} agents of the workload per run, median of {CPU_RUNS} runs after {CPU_WARM} "
                      f"warm-ups ({t * 1e3:.2f} ms per run), oracle/cpu_baseline.c (OpenMP over agents, {flags})"}


def _efe_macs(pm, S, O, A_dependencies, B_dependencies, batch):
    """Multiply-adds of the two GEMM stages of the neg-EFE pipeline for one
    call: gemm1 ([A ; H_A] rows x inner tuple, per distinct inner history
    tuple of each level) and the score GEMM (every inner tuple x outer
    history of each level), from the action-history dedup of the policies."""
    import numpy as np

    from oracle import dedup_ref as DR
    lv = DR.factor_histories(pm, B_dependencies)
    T = pm.shape[1]
    macs = 0
    for m, deps in enumerate(A_dependencies):
        S0, inner = S[deps[0]], deps[1:]
        K_in = int(np.prod([S[d] for d in inner])) if inner else 1
        for t in range(T):
            H0 = len(lv[t][deps[0]][1])
            n_in = len(np.unique(np.stack([lv[t][d][2] for d in inner], 1), axis=0)) if inner else 1
            macs += (O[m] + 1) * S0 * K_in * n_in + (O[m] + 1) * S0 * n_in * H0
    return batch * macs


def bench_plain_efe(a, torch):
    import numpy as np

    from oracle import dedup_ref as DR
    from oracle import numpy_ref as R
    from pymdp_amd import workloads
    from pymdp_amd.ffi._efe import update_posterior_policies_ffi

    dev = torch.device("cuda:0")
    w = workloads.plain_efe(seed=0, device=dev)
    pm = R.construct_policies(w["S"], w["U"], w["T"])
    B = w["batch"]

    def step():
        return update_posterior_policies_ffi(pm, w["qs"], w["A"], w["B"], w["C"], None, w["A_dependencies"],
                                             w["B_dependencies"], gamma=w["gamma"])

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(a.steps):
        q_pi, G = step()
    e1.record(stream)
    torch.cuda.synchronize()
    ms_step = (time.perf_counter() - t0) * 1e3 / a.steps
    ms_events = e0.elapsed_time(e1) / a.steps
    macs = _efe_macs(pm, w["S"], w["O"], w["A_dependencies"], w["B_dependencies"], B)
    achieved = 2.0 * macs / (ms_events * 1e-3) / 1e12
    # parity: two agents against the FP64 dedup oracle
    eG = eq = 0.0
    for b in (0, B - 1):
        A_b = [x[b].cpu().numpy() for x in w["A"]]
        B_b = [x[b].cpu().numpy() for x in w["B"]]
        C_b = [x[b].cpu().numpy() for x in w["C"]]
        qs_b = [x[b].cpu().numpy().astype(np.float64) for x in w["qs"]]
        G64 = DR.neg_efe_dedup(pm, qs_b, A_b, B_b, C_b, w["A_dependencies"], w["B_dependencies"], dt=np.float64)
        p64 = DR.q_pi_from_G(G64, w["gamma"])
        eG = max(eG, float(np.abs(G[b].cpu().numpy() - G64).max()))
        eq = max(eq, float(np.abs(q_pi[b].cpu().numpy() - p64).max()))
    cpu = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline_efe(w, pm)
    return {
        "metric": "agent-steps/s (plain EFE, 64 agents, 32768 policies)",
        "value": B / (ms_step * 1e-3), "unit": "agent-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 data, f64 accumulation", "data": "synthetic",
        "config": {"workload": "plain_efe", "agents": B, "num_states": w["S"], "num_controls": w["U"],
                   "horizon": w["T"], "num_obs": w["O"], "num_policies": int(pm.shape[0]),
                   "efe_macs_per_step": macs},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "kernel": "efe pipeline (rollout, H_A, gemm1, score, scatter; gemm1 and score on FP64 MFMA)",
                     "kernel_ms": ms_events,
                     "alg_flops_per_launch": 2 * macs,
                     "note": "flops are the multiply-adds of the two GEMM stages (A x Kronecker weights, outcome "
                             "weights x outcome-state products), counted once per deduplicated history"},
        "parity": {"agents_checked": 2, "max_abs_err_G_vs_fp64": eG, "max_abs_err_q_pi_vs_fp64": eq},
        "cpu_baseline": cpu,
    }


def cpu_baseline_efe(w, pm):
    """The oracle's C/OpenMP restatement of the reference CPU factor-history
    neg-EFE + q_pi + action (-march=native), one agent per core: the first
    (cores) agents per run, median of 20 runs after 3 warm-ups."""
    from oracle import cpu_ref
    flags = _cpu_native()
    threads = _cpu_threads()
    n = min(threads, int(w["batch"]))
    ags = [dict(S=w["S"], U=w["U"], O=w["O"], A=[x[b].cpu().numpy() for x in w["A"]],
                B=[x[b].cpu().numpy() for x in w["B"]], C=[x[b].cpu().numpy() for x in w["C"]],
                qs=[x[b].cpu().numpy() for x in w["qs"]], policies=pm, A_dependencies=w["A_dependencies"],
                B_dependencies=w["B_dependencies"], gamma=w["gamma"]) for b in range(n)]
    host = cpu_ref.Agents(ags, 16, do_fpi=False)
    t = _median_runs(lambda: host.run(threads))
    return {"value": n / t, "unit": "agent-steps/s", "cores": threads, "kind": "port",
            "sample": f"first {n} agents per run (one per core), median of {CPU_RUNS} runs after {CPU_WARM} "
                      f"warm-ups ({t:.3f} s per run), oracle/cpu_baseline.c factor-history neg-EFE + q_pi + "
                      f"argmax (shared inner-factor contraction at every level, OpenMP over agents, {flags})"}


def _plain_agent_step(ag, tensors, torch):
    """One agent through the plain per-agent entry points (infer_states_ffi,
    update_posterior_policies_ffi, argmax action) at batch size 1."""
    from pymdp_amd.ffi import infer_states_ffi
    from pymdp_amd.ffi._efe import update_posterior_policies_ffi
    A, B, C, D, obs, pol0 = tensors
    qs = infer_states_ffi(A, obs, D, ag["A_dependencies"], 16, distr_obs=False)
    q_pi, G = update_posterior_policies_ffi(ag["policies"], qs, A, B, C, None, ag["A_dependencies"],
                                            ag["B_dependencies"], gamma=ag["gamma"])
    # index_select keeps the choice on the device (no host sync; capturable)
    return torch.index_select(pol0, 0, torch.argmax(q_pi[0]).view(1))[0]


def bench_ragged(a, torch):
    import numpy as np

    from pymdp_amd import ragged, workloads

    dev = torch.device("cuda:0")
    w = workloads.ragged(seed=0, device=dev, n_agents=4096)
    ags = w["agents"]
    rb = ragged.RaggedBatch(ags, num_iter=w["num_iter"], device=dev)
    # the agents' parameters do not change between steps (no learning), the
    # setting under which the reference caches A-derived tables
    # (ffi_cache_params_static, pymdp/agent.py:880): the sum_o A ln A table is
    # formed once and every step reuses it
    rb.set_params_static(True)
    rb.step()
    torch.cuda.synchronize()
    if not a.no_autotune:
        rb.autotune()
    g = rb.capture()
    for _ in range(a.warmup):
        g.replay()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(a.steps):
        g.replay()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_events = e0.elapsed_time(e1) / a.steps
    ms_step = wall * 1e3 / a.steps
    n = len(ags)
    alg = ragged.algorithmic_bytes(ags)
    achieved = alg / (ms_events * 1e-3) / 1e9
    # the same step when parameters may change between steps: the table is
    # formed inside every step (one more pass over A)
    rb.set_params_static(False)
    gd = rb.capture()
    for _ in range(2):
        gd.replay()
    e0.record(stream)
    for _ in range(max(3, a.steps // 4)):
        gd.replay()
    e1.record(stream)
    torch.cuda.synchronize()
    ms_dynamic = e0.elapsed_time(e1) / max(3, a.steps // 4)
    rb.set_params_static(True)
    rb.step()

    # plain per-agent entry points on a bounded sample (every 16th agent)
    sample = list(range(0, n, 16))
    tens = []
    for i in sample:
        ag = ags[i]
        tens.append(([x[None] for x in ag["A"]], [x[None] for x in ag["B"]], [x[None] for x in ag["C"]],
                     [x[None] for x in ag["D"]], [torch.tensor([o], device=dev) for o in ag["obs"]],
                     torch.as_tensor(ag["policies"][:, 0, :], device=dev)))
    for i, tt in zip(sample[:8], tens[:8]):
        _plain_agent_step(ags[i], tt, torch)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain_act = [_plain_agent_step(ags[i], tt, torch) for i, tt in zip(sample, tens)]
    torch.cuda.synchronize()
    plain_s = time.perf_counter() - t0
    plain_rate = len(sample) / plain_s
    same = all(torch.equal(pa.to(torch.int32), rb.agent_outputs(i)[3]) for pa, i in zip(plain_act, sample))
    _log("ragged: plain per-agent calls captured per agent")
    # the same per-agent calls timed as GPU time: each agent's call sequence
    # captured once in a graph (after one eager call of that shape on the
    # capture stream), all graphs replayed back to back between two events
    cap = torch.cuda.Stream()
    graphs = []
    with torch.cuda.stream(cap):
        for i, tt in zip(sample, tens):
            _plain_agent_step(ags[i], tt, torch)
            gi = torch.cuda.CUDAGraph()
            with torch.cuda.graph(gi, stream=cap):
                _plain_agent_step(ags[i], tt, torch)
            graphs.append(gi)
    torch.cuda.synchronize()
    for gi in graphs:
        gi.replay()
    torch.cuda.synchronize()
    e0.record(stream)
    for gi in graphs:
        gi.replay()
    e1.record(stream)
    torch.cuda.synchronize()
    plain_gpu_ms = e0.elapsed_time(e1)
    plain_gpu_rate = len(sample) / (plain_gpu_ms * 1e-3)
    del graphs

    # parity of a few agents against the oracle (FPI restatement + history-dedup EFE, FP64)
    from oracle import dedup_ref as DR
    from oracle import numpy_ref as R
    e64 = 0.0
    for i in (0, n // 3, 2 * n // 3, n - 1):
        ag = ags[i]
        qs, G, q_pi, act = rb.agent_outputs(i)
        A = [x.cpu().numpy() for x in ag["A"]]
        ob = [np.eye(o, dtype=np.float32)[x] for o, x in zip(ag["O"], ag["obs"])]
        q64 = R.run_factorized_fpi(A, ob, [d.cpu().numpy() for d in ag["D"]], ag["A_dependencies"], 16, np.float64)
        e64 = max(e64, max(float(np.abs(q.cpu().numpy() - r).max()) for q, r in zip(qs, q64)))
        qg = [q.cpu().numpy().astype(np.float64) for q in qs]
        G64 = DR.neg_efe_dedup(ag["policies"], qg, A, [x.cpu().numpy() for x in ag["B"]],
                               [x.cpu().numpy() for x in ag["C"]], ag["A_dependencies"], ag["B_dependencies"])
        p64 = DR.q_pi_from_G(G64, ag["gamma"])
        e64 = max(e64, float(np.abs(G.cpu().numpy() - G64).max()), float(np.abs(q_pi.cpu().numpy() - p64).max()))

    cpu = None
    if not a.no_cpu_baseline:
        _log("ragged: CPU baseline")
        cpu = cpu_baseline_ragged(ags, rb)
    info = rb.plan.info()
    return {
        "metric": "agent-steps/s (batched ragged FPI+EFE step, 4096 agents, HIP graph)",
        "value": n / (ms_step * 1e-3),
        "unit": "agent-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 data, f64 accumulation",
        "data": "synthetic (seeded ragged shapes, Dirichlet A/B, N(0,1) C; generated on device)",
        "config": {"workload": "ragged", "agents": n, "policies_total": ragged.num_policies(ags),
                   "num_iter": w["num_iter"], "gamma": ags[0]["gamma"], "fpi_units": info["units"],
                   "fpi_team_max": info["team_max"], "fpi_grid": info["grid_fpi"],
                   "efe_tiles": [info["tiles1"], info["tiles2"], info["tiles3"]],
                   "tile_sizes": rb.plan.tile_config(),
                   "efe_macs_per_step": info["efe_macs"], "fpi_macs_per_step": info["fpi_macs"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": _traffic("ragged step"),
                     "kernel": "rg_fpi + rg_roll (one per level) + rg_rows stages 1, 2, 3 + rg_assemble kernels "
                               "(one graph, static parameters)",
                     "kernel_ms": ms_events, "alg_bytes_per_launch": alg},
        "plain_per_agent": {"value": plain_rate, "unit": "agent-steps/s",
                            "sample": f"every 16th agent ({len(sample)}), batch-1 infer_states_ffi + "
                                      f"update_posterior_policies_ffi + argmax, {plain_s:.2f} s wall",
                            "same_actions_as_batched": bool(same),
                            "gpu_time": {"value": plain_gpu_rate, "unit": "agent-steps/s",
                                         "ms_total": plain_gpu_ms,
                                         "how": "each agent's calls captured in one graph, the graphs replayed "
                                                "back to back between two HIP events (no host work)"}},
        "speedup_vs_plain": (n / (ms_step * 1e-3)) / plain_rate,
        "speedup_vs_plain_gpu_time": (n / (ms_events * 1e-3)) / plain_gpu_rate,
        "params_changing_each_step": {"ms_per_step": ms_dynamic, "value": n / (ms_dynamic * 1e-3),
                                      "note": "sum_o A ln A formed inside every step (rg_ha_kernel)"},
        "parity": {"agents_checked": 4, "max_abs_err_vs_fp64": e64, "team_barrier_timeouts": rb.status()},
        "cpu_baseline": cpu,
    }


def bench_grad(a, torch):
    from oracle import autograd_ref as AB
    from pymdp_amd import differentiable as DF
    from pymdp_amd import workloads

    dev = torch.device("cuda:0")
    w = workloads.grad_fpi_efe(seed=0, device=dev)
    keys = ("A_logits", "B_logits", "C_logits", "D_logits")
    ph = [[x.clone().requires_grad_() for x in w[k]] for k in keys]
    pr = [[x.clone().requires_grad_() for x in w[k]] for k in keys]

    def run(params, fn, **kw):
        for grp in params:
            for x in grp:
                x.grad = None
        loss = fn(*params, w["obs"], w["actions"], w["policies"], w["A_dependencies"], num_iter=w["num_iter"],
                  gamma=w["gamma"], **kw)
        loss.backward()
        return loss

    def hip():
        return run(ph, DF.trajectory_loss, B_dependencies=w["B_dependencies"])

    def ref():
        return run(pr, AB.trajectory_loss)

    def timed(fn, k):
        for _ in range(max(1, a.warmup)):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(k):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / k

    ms_hip = timed(hip, a.steps)
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(a.steps):
        hip()
    e1.record(stream)
    torch.cuda.synchronize()
    ms_events = e0.elapsed_time(e1) / a.steps
    ms_ref = timed(ref, max(2, a.steps // 5))
    num = max(float((x.grad - y.grad).abs().max()) for gx, gy in zip(ph, pr) for x, y in zip(gx, gy))
    den = max(float(y.grad.abs().max()) for gy in pr for y in gy)
    B = w["batch"]
    # algorithmic bytes of one gradient: every logit read once and its
    # gradient written once, the observation and action trajectories read
    n_logits = sum(int(x.numel()) for k in keys for x in w[k])
    alg_bytes = 2 * 4 * n_logits + 4 * int(w["obs"].numel()) + 8 * int(w["actions"].numel())
    achieved = alg_bytes / (ms_events * 1e-3) / 1e9
    cpu = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline_grad(w, keys, AB, torch)
    return {
        "metric": "agent-trajectory gradients/s (512 agents x 32 steps, dL/d logits of A, B, C, D)",
        "value": B / (ms_hip * 1e-3), "unit": "agent-gradients/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms_hip, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 data, f64 adjoint accumulation",
        "data": "synthetic (N(0,1) logits, uniform observation / policy trajectories)",
        "config": {"workload": "grad", "agents": B, "steps_per_trajectory": w["T"], "num_states": w["S"],
                   "num_controls": w["U"], "num_obs": w["O"], "num_iter": w["num_iter"],
                   "adjoint": "implicit where converged, unrolled otherwise"},
        "grads_per_s": 1e3 / ms_hip,
        "autograd_baseline": {"ms_per_gradient": ms_ref, "grads_per_s": 1e3 / ms_ref,
                              "what": "torch autograd through the pure-torch FP32 forward on the same GPU"},
        "speedup_vs_autograd": ms_ref / ms_hip,
        "parity": {"max_abs_grad_diff_vs_autograd": num, "max_abs_grad": den},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "trajectory loss forward + HIP adjoints (32 time steps, all kernels of one gradient)",
                     "kernel_ms": ms_events, "alg_bytes_per_launch": alg_bytes,
                     "note": "32 dependent time steps of 512 small agents: latency-bound; the bytes are every "
                             "logit read once and its gradient written once"},
        "cpu_baseline": cpu,
    }


def cpu_baseline_grad(w, keys, AB, torch, n=32):
    """torch autograd through the pure-torch FP32 forward on the host cores
    (torch.set_num_threads = every core of this host), on the first n
    agents: median of 20 runs after 3 warm-ups."""
    threads = _cpu_threads()
    old = torch.get_num_threads()
    torch.set_num_threads(threads)
    try:
        params = [[x[:n].detach().cpu().clone().requires_grad_() for x in w[k]] for k in keys]
        obs = w["obs"][:, :n].cpu()
        actions = w["actions"][:, :n].cpu()

        def run():
            for grp in params:
                for x in grp:
                    x.grad = None
            loss = AB.trajectory_loss(*params, obs, actions, w["policies"], w["A_dependencies"],
                                      num_iter=w["num_iter"], gamma=w["gamma"])
            loss.backward()

        t = _median_runs(run)
    finally:
        torch.set_num_threads(old)
    return {"value": n / t, "unit": "agent-gradients/s", "cores": threads, "kind": "port",
            "sample": f"first {n} agents per run, median of {CPU_RUNS} runs after {CPU_WARM} warm-ups "
                      f"({t:.3f} s per run), torch autograd through the FP32 forward of oracle/autograd_ref.py "
                      f"on the host (torch.set_num_threads({threads}))"}


def cpu_baseline_ragged(ags, rb, n=1024):
    """The oracle's C/OpenMP restatement of the reference CPU path per agent
    (FPI, factor-history neg-EFE, q_pi, argmax action; -march=native,
    OpenMP over agents on every core of this host) on the first n agents:
    median of 20 runs after 3 warm-ups.  Its actions are compared with the
    batched GPU step's on the first 256 of them."""
    import numpy as np

    from oracle import cpu_ref
    flags = _cpu_native()
    threads = _cpu_threads()
    host = []
    for ag in ags[:n]:
        d = dict(ag)
        for k in "ABCD":
            d[k] = [x.cpu().numpy() for x in ag[k]]
        host.append(d)
    H = cpu_ref.Agents(host, 16, do_fpi=True)
    t = _median_runs(lambda: H.run(threads))
    same = True
    for i, o in enumerate(H.out[:256]):
        act = np.asarray(ags[i]["policies"])[int(o["best"][0]), 0, :]
        same &= bool(np.array_equal(act, rb.agent_outputs(i)[3].cpu().numpy()))
    return {"value": n / t, "unit": "agent-steps/s", "cores": threads, "kind": "port",
            "sample": f"first {n} agents per run, median of {CPU_RUNS} runs after {CPU_WARM} warm-ups "
                      f"({t * 1e3:.1f} ms per run), oracle/cpu_baseline.c FPI + factor-history neg-EFE + q_pi + "
                      f"argmax (OpenMP over agents, {flags})",
            "same_actions_as_gpu_first_256": same}


HEADLINE_METRIC = "agent-steps/s & % HBM peak: batched vs plain FPI+EFE at <=1e-5 err; grads/s"


def bench_headline(a, torch):
    import copy
    _log("ragged step")
    res = bench_ragged(a, torch)
    res["metric"] = HEADLINE_METRIC
    keep = ("value", "unit", "ms_per_step", "steps", "config", "roofline", "parity", "cpu_baseline",
            "autograd_baseline", "speedup_vs_autograd", "grads_per_s", "sweep_trace")
    b = copy.copy(a)
    b.steps, b.warmup = 50, 5
    _log("plain_fpi_small")
    small = bench_plain_fpi_small(b, torch)
    res["plain_fpi_small"] = {k: small[k] for k in keep if k in small}
    b = copy.copy(a)
    b.steps, b.warmup = 20, 3
    _log("plain_fpi_large")
    fpi = bench_plain_fpi_large(b, torch)
    res["plain_fpi_large"] = {k: fpi[k] for k in keep if k in fpi}
    b = copy.copy(a)
    b.steps, b.warmup = 5, 1
    _log("plain_efe")
    efe = bench_plain_efe(b, torch)
    res["plain_efe"] = {k: efe[k] for k in keep if k in efe}
    b = copy.copy(a)
    b.steps, b.warmup = 5, 1
    _log("grad")
    g = bench_grad(b, torch)
    res["grad"] = {k: g[k] for k in keep if k in g}
    res["grads_per_s"] = g["grads_per_s"]
    b = copy.copy(a)
    b.steps, b.warmup = 10, 2
    _log("learning")
    lrn = bench_learning(b, torch)
    res["learning"] = {k: lrn[k] for k in keep if k in lrn}
    return res


def main():
    a = _args()
    if a.write_traffic:
        write_traffic(a)
        return
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {a.gpus}); "
              "multi-GPU runs are not part of it", file=sys.stderr)
        sys.exit(2)
    if not a.no_autotune:
        os.environ.setdefault("PYMDP_HIP_AUTOTUNE", "1")
    import torch
    if a.calibrate_copy:
        calibrate_copy()
        return
    if a.workload == "headline":
        res = bench_headline(a, torch)
    elif a.workload == "plain_fpi_small":
        res = bench_plain_fpi_small(a, torch)
    elif a.workload == "plain_fpi_large":
        res = bench_plain_fpi_large(a, torch)
    elif a.workload == "ragged":
        res = bench_ragged(a, torch)
    elif a.workload == "grad":
        res = bench_grad(a, torch)
    elif a.workload == "learning":
        res = bench_learning(a, torch)
    else:
        res = bench_plain_efe(a, torch)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
