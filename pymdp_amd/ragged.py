"""Batched ragged agents: many independent agents of different shapes stepped
together (infer_states -> infer_policies -> deterministic sample_action ->
update_empirical_prior) by the persistent HIP kernels of libpymdp_hip
(pymdp_hip_ragged_*).

Per agent the step computes exactly what the reference's one-agent path
computes: the FPI posterior of pymdp/algos.py:88-155 from the agent's
observation indices and prior D, the neg-EFE of every policy
(pymdp/control.py:604-793 with utility and state information gain), q_pi =
softmax(gamma G + ln E) (pymdp/control.py:963) and the first action of the
first-index argmax policy (pymdp/control.py:115-150), and the predicted prior
B . q of that action for the next step (pymdp/agent.py:825-851); in carry mode
the next step's FPI starts from that prior instead of D, as the reference's
rollout does (pymdp/envs/rollout.py:65-82).  The batched entry has
no counterpart in the reference; the plain per-agent entry points stay the
drop-in (pymdp_amd.ffi).

Layout: one float32 parameter arena (every agent's A, B, C, D, E at offsets
the plan chooses; A_m is stored outcome-major with its dependency axes in
the plan's contraction order: [O_m][positions of the leading axes][last
axis], so that the rows of one outcome are contiguous (the FPI reads the
observed outcome's rows, the EFE stream reads every row in storage order);
B_f as [U_f][dependency states][S_f]), one int32 observation arena, one float32 output arena
(qs, G, q_pi, predicted prior per agent) and one int32 action arena.  A step allocates nothing
and does not synchronise, so it can be captured in a HIP graph
(RaggedBatch.capture).
"""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np
import torch

from . import _native


def _spec(agents: Sequence[dict]):
    hdr, S, U, O, Ar, Af, Br, Bf, Ct, Ep, gam, pol = ([] for _ in range(12))
    for ag in agents:
        F, M = len(ag["S"]), len(ag["O"])
        pm = np.ascontiguousarray(np.asarray(ag["policies"], dtype=np.int32))
        if pm.ndim != 3 or pm.shape[2] != F:
            raise ValueError("ragged: policies must be [P, T, F] int32")
        P, T = pm.shape[0], pm.shape[1]
        hdr += [F, M, T, P]
        S += list(ag["S"])
        U += list(ag["U"])
        O += list(ag["O"])
        for d in ag["A_dependencies"]:
            Ar.append(len(d))
            Af += list(d)
        for d in ag["B_dependencies"]:
            Br.append(len(d))
            Bf += list(d)
        if ag.get("C") is None:
            Ct += [0] * M
        else:
            Ct += [1 if (c.dim() == 2) else 0 for c in ag["C"]]
        Ep.append(0 if ag.get("E") is None else 1)
        gam.append(float(ag.get("gamma", 1.0)))
        pol.append(pm.reshape(-1))
    return (hdr, S, U, O, Ar, Af, Br, Bf, Ct, np.asarray(Ep, np.int32), np.asarray(gam, np.float32),
            np.concatenate(pol).astype(np.int32))


class RaggedPlan:
    """Host plan of a ragged batch (no device needed until `upload`)."""

    def __init__(self, agents: Sequence[dict], num_iter: int = 16):
        L = _native.lib()
        self.n = len(agents)
        self.sumF = sum(len(a["S"]) for a in agents)
        self.sumM = sum(len(a["O"]) for a in agents)
        hdr, S, U, O, Ar, Af, Br, Bf, Ct, Ep, gam, pol = _spec(agents)
        self._keep = [_native.I64(x) for x in (hdr, S, U, O, Ar, Af if Af else [0], Br, Bf if Bf else [0], Ct)]
        self._Ep, self._gam, self._pol = Ep, gam, pol
        h = ctypes.c_void_p()
        k = self._keep
        _native.check(L.pymdp_hip_ragged_plan_create(
            self.n, k[0].ptr, k[1].ptr, k[2].ptr, k[3].ptr, k[4].ptr, k[5].ptr, k[6].ptr, k[7].ptr, k[8].ptr,
            Ep.ctypes.data, gam.ctypes.data, pol.ctypes.data, int(num_iter), ctypes.byref(h)))
        self.handle = h
        nl = L.pymdp_hip_ragged_layout(h, None, 0)
        lay = np.zeros(nl, np.int64)
        _native.check(L.pymdp_hip_ragged_layout(h, lay.ctypes.data_as(_native._i64p), nl))
        n, sF, sM = self.n, self.sumF, self.sumM
        self.sizes = lay[:4].copy()
        pos = 4
        fields = [("A_off", sM), ("A_perm", 3 * sM), ("B_off", sF), ("C_off", sM), ("D_off", sF), ("E_off", n),
                  ("obs_off", n), ("qs_off", n), ("G_off", n), ("qpi_off", n), ("act_off", n), ("pred_off", n)]
        self.layout = {}
        for name, cnt in fields:
            self.layout[name] = lay[pos:pos + cnt].copy()
            pos += cnt
        if pos != nl:
            raise RuntimeError("ragged: unexpected layout length")
        self.layout["A_perm"] = self.layout["A_perm"].reshape(sM, 3)
        self.uploaded = False

    def info(self) -> dict:
        out = np.zeros(16, np.int64)
        _native.check(_native.lib().pymdp_hip_ragged_info(self.handle, out.ctypes.data_as(_native._i64p)))
        keys = ["agents", "units", "tiles1", "tiles2", "tiles3", "grid_fpi", "hist_doubles", "score_doubles",
                "z1_doubles", "z2_doubles", "part_doubles", "team_max", "efe_macs", "fpi_macs", "grid_cap", "ha_floats"]
        return dict(zip(keys, (int(x) for x in out)))

    def tile_config(self) -> dict:
        """EFE work-item sizes in use (pymdp_hip_ragged_tile_config)."""
        out = (ctypes.c_int32 * 8)()
        _native.check(_native.lib().pymdp_hip_ragged_tile_config(self.handle, out))
        keys = ["e1_rows", "e2_rows", "e3_rows", "autotuned", "grid_fpi"]
        return dict(zip(keys, (int(x) for x in out)))

    def set_tiles(self, e1_rows: int, e2_rows: int):
        """Sets the stage-1 and stage-2 rows per work item (128 or 256);
        results do not depend on them."""
        cfg = (ctypes.c_int32 * 2)(e1_rows, e2_rows)
        _native.check(_native.lib().pymdp_hip_ragged_set_tiles(self.handle, cfg))

    def upload(self):
        if not self.uploaded:
            _native.check(_native.lib().pymdp_hip_ragged_plan_upload(self.handle))
            self.uploaded = True

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _native.lib().pymdp_hip_ragged_plan_destroy(h)
            except Exception:  # pragma: no cover - interpreter shutdown
                pass
            self.handle = None


def pack_params(plan: RaggedPlan, agents: Sequence[dict], device) -> torch.Tensor:
    """One float32 arena with every agent's A (axes in storage order), B, C,
    D and E at the plan's offsets."""
    arena = torch.zeros(int(plan.sizes[0]), dtype=torch.float32, device=device)
    lay = plan.layout
    mi = fi = 0
    for ai, ag in enumerate(agents):
        for m, A in enumerate(ag["A"]):
            perm = [int(p) for p in lay["A_perm"][mi] if p >= 0]
            # outcome-major, then the storage axes in the plan's order:
            # [O][positions of all but the last storage axis][last axis]
            At = A.permute(0, *[1 + p for p in perm]).contiguous().reshape(-1)
            o = int(lay["A_off"][mi])
            arena[o:o + At.numel()] = At.to(device)
            C = ag["C"][m].reshape(-1)
            o = int(lay["C_off"][mi])
            arena[o:o + C.numel()] = C.to(device)
            mi += 1
        for f in range(len(ag["S"])):
            Bf = ag["B"][f]
            # [U_f][dependency states][S_f]: the rollout's lanes run over S_f
            B = Bf.permute(Bf.dim() - 1, *range(1, Bf.dim() - 1), 0).contiguous().reshape(-1)
            o = int(lay["B_off"][fi])
            arena[o:o + B.numel()] = B.to(device)
            D = ag["D"][f].reshape(-1)
            o = int(lay["D_off"][fi])
            arena[o:o + D.numel()] = D.to(device)
            fi += 1
        if ag.get("E") is not None:
            E = ag["E"].reshape(-1)
            o = int(lay["E_off"][ai])
            arena[o:o + E.numel()] = E.to(device)
    return arena


def pack_obs(plan: RaggedPlan, obs: Sequence[Sequence[int]], device) -> torch.Tensor:
    flat = np.zeros(max(1, int(plan.sizes[1])), np.int32)
    for ai, o in enumerate(obs):
        base = int(plan.layout["obs_off"][ai])
        flat[base:base + len(o)] = np.asarray(o, np.int32)
    return torch.from_numpy(flat).to(device)


class RaggedBatch:
    """Ragged batch on one device: plan + packed arenas + the step launch.

        rb = RaggedBatch(agents, num_iter=16, device="cuda:0")
        rb.step()                 # eager step on the current stream
        rb.autotune()             # optional: pick launch grids by timing
        g = rb.capture()          # HIP graph of one step; g.replay()
        qs, G, q_pi, action = rb.agent_outputs(i)
        rb.set_prior_source(carry=True)   # later steps start from B . q
        rb.reset_prior()                  # ... which for the next step is D
    """

    def __init__(self, agents: Sequence[dict], num_iter: int = 16, device="cuda:0"):
        self.device = torch.device(device)
        self.agents = agents
        self.plan = RaggedPlan(agents, num_iter)
        with torch.cuda.device(self.device):
            self.plan.upload()
        self.params = pack_params(self.plan, agents, self.device)
        self.obs = pack_obs(self.plan, [a["obs"] for a in agents], self.device)
        self.out = torch.zeros(max(1, int(self.plan.sizes[2])), dtype=torch.float32, device=self.device)
        self.act = torch.zeros(max(1, int(self.plan.sizes[3])), dtype=torch.int32, device=self.device)
        self.graph = None
        self.carry = False
        self.params_static = False
        # gather / scatter indices D -> predicted-prior slots (reset_prior)
        d_idx, p_idx = [], []
        fi = 0
        for i, ag in enumerate(agents):
            po = int(self.plan.layout["pred_off"][i])
            for s in ag["S"]:
                d_idx.append(np.arange(s) + int(self.plan.layout["D_off"][fi]))
                p_idx.append(np.arange(s) + po)
                po += s
                fi += 1
        self._d_idx = torch.as_tensor(np.concatenate(d_idx), device=self.device)
        self._p_idx = torch.as_tensor(np.concatenate(p_idx), device=self.device)

    def set_prior_source(self, carry: bool):
        """carry=False: every step's FPI prior is D.  carry=True: it is the
        predicted prior the previous step wrote (graphs captured before this
        call keep the old source)."""
        _native.check(_native.lib().pymdp_hip_ragged_set_prior_source(self.plan.handle, 1 if carry else 0))
        self.carry = bool(carry)

    def reset_prior(self):
        """Predicted-prior slots := D (stream-ordered device copy), so the next
        carry-mode step starts from D as the reference's first rollout step."""
        with torch.cuda.device(self.device):
            self.out[self._p_idx] = self.params[self._d_idx]

    def set_params_static(self, static: bool):
        """static=True: the params arena is not modified in place between
        steps, so the sum_o A ln A table of the stage-1 product is formed once
        (by the next step) and reused; False (the default): every step forms
        it.  Call again after changing parameters in place."""
        _native.check(_native.lib().pymdp_hip_ragged_set_params_static(self.plan.handle, 1 if static else 0))
        self.params_static = bool(static)

    def set_obs(self, obs: Sequence[Sequence[int]]):
        self.obs.copy_(pack_obs(self.plan, obs, self.device))

    def _args(self):
        return (self.plan.handle, self.params.data_ptr(), self.obs.data_ptr(), self.out.data_ptr(),
                self.act.data_ptr())

    def step(self):
        with torch.cuda.device(self.device):
            _native.check(_native.lib().pymdp_hip_ragged_step(*self._args(), _native.stream_handle(self.device)))

    def autotune(self) -> float:
        ms = ctypes.c_float(0.0)
        with torch.cuda.device(self.device):
            _native.check(_native.lib().pymdp_hip_ragged_autotune(*self._args(), _native.stream_handle(self.device),
                                                                  ctypes.byref(ms)))
        return float(ms.value)

    def status(self) -> int:
        """1 if an FPI team exchange of an earlier step timed out (that
        agent's posterior is NaN); synchronises the device."""
        flag = ctypes.c_int32(0)
        with torch.cuda.device(self.device):
            _native.check(_native.lib().pymdp_hip_ragged_status(self.plan.handle, ctypes.byref(flag)))
        return int(flag.value)

    def capture(self) -> torch.cuda.CUDAGraph:
        """HIP graph of one step (five kernel nodes)."""
        with torch.cuda.device(self.device):
            s = torch.cuda.Stream(self.device)
            s.wait_stream(torch.cuda.current_stream(self.device))
            with torch.cuda.stream(s):
                self.step()
            torch.cuda.current_stream(self.device).wait_stream(s)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self.step()
        self.graph = g
        return g

    def agent_outputs(self, i: int):
        ag = self.agents[i]
        lay = self.plan.layout
        P = int(np.asarray(ag["policies"]).shape[0])
        qs_o, G_o, q_o, a_o = (int(lay[k][i]) for k in ("qs_off", "G_off", "qpi_off", "act_off"))
        S = ag["S"]
        qs_flat = self.out[qs_o:qs_o + sum(S)]
        qs = list(torch.split(qs_flat, S))
        return qs, self.out[G_o:G_o + P], self.out[q_o:q_o + P], self.act[a_o:a_o + len(S)]

    def agent_prior(self, i: int):
        """Predicted prior per factor that agent i's last step wrote."""
        o = int(self.plan.layout["pred_off"][i])
        S = self.agents[i]["S"]
        return list(torch.split(self.out[o:o + sum(S)], S))


def algorithmic_bytes(agents: Sequence[dict]) -> int:
    """HBM bytes one step must move: A read in full once (the EFE needs every
    row; the FPI's observed rows come from the same lines), B, C, D, E and
    the observation indices read, qs, G, q_pi, actions and the predicted
    prior written."""
    tot = 0
    for ag in agents:
        P = int(np.asarray(ag["policies"]).shape[0])
        tot += sum(4 * a.numel() for a in ag["A"]) + sum(4 * b.numel() for b in ag["B"])
        tot += sum(4 * c.numel() for c in ag["C"]) + sum(4 * d.numel() for d in ag["D"])
        tot += 4 * len(ag["O"]) + 8 * sum(ag["S"]) + 8 * P + 4 * len(ag["S"])
        if ag.get("E") is not None:
            tot += 4 * P
    return tot


def num_policies(agents: Sequence[dict]) -> int:
    return int(sum(np.asarray(a["policies"]).shape[0] for a in agents))


__all__ = ["RaggedPlan", "RaggedBatch", "pack_params", "pack_obs", "algorithmic_bytes", "num_policies"]
