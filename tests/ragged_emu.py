"""Host emulation of the batched ragged kernels (pymdp_amd/csrc/ragged.hip).

Reads the plan's own device tables (agent records, FPI units, EFE tiles, int
arena) and replays the kernels' index arithmetic in NumPy, asserting that
every access stays inside its buffer (LDS arrays, stage buffers, arenas),
and computing the same values, so that the plan and the kernels' indexing
are checked on the CPU and their results compared with the oracle.

Test infrastructure only.
"""
from __future__ import annotations

import numpy as np

from pymdp_amd import _native

MINVAL = np.float32(2.220446049250313e-16)
kF, kM, kR, kT, kDim, kNT = 8, 8, 3, 4, 64, 256
kCap, kRowsCap, kTeamMax = 6144, 1024, 128
kE1Cols, kE1Ld = 16, 64  # stage-1 / 2 history columns staged at a time, their LDS row stride
kE3Cols, kE3Rows = 8, 256  # stage-3 columns staged at a time, row slots of a stage-3 item

AGENT_DT = np.dtype([
    ("F", "i4"), ("M", "i4"), ("T", "i4"), ("P", "i4"), ("G", "i4"), ("iters", "i4"), ("totS", "i4"),
    ("has_E", "i4"), ("team", "i4"), ("gamma", "f4"),
    ("S", "i4", (kF,)), ("S_off", "i4", (kF + 1,)), ("O", "i4", (kM,)), ("mod_group", "i4", (kM,)),
    ("gR", "i4", (kM,)), ("gfac", "i4", (kM, kR)), ("gdim", "i4", (kM, kR)),
    ("BR", "i4", (kF,)), ("Bdep", "i4", (kF, 3)), ("BK", "i4", (kF,)), ("U", "i4", (kF,)),
    ("C_time", "i4", (kM,)),
    ("H", "i4", (kT, kF)), ("hist_lvl", "i4", (kT, kF)),
    ("score_base", "i4", (kT, kM)), ("score_stride", "i4", (kT, kM, kR)),
    ("ed", "i4", (kM, 3)), ("ef", "i4", (kM, 3)), ("tstart", "i4", (kM, kT + 1)),
    ("K", "i8", (kM,)), ("A_off", "i8", (kM,)), ("B_off", "i8", (kF,)), ("C_off", "i8", (kM,)),
    ("D_off", "i8", (kF,)), ("E_off", "i8"),
    ("obs_off", "i8"), ("qs_off", "i8"), ("G_off", "i8"), ("qpi_off", "i8"), ("act_off", "i8"),
    ("hist_off", "i8"), ("p2h_off", "i8"), ("pol0_off", "i8"), ("score_off", "i8"),
    ("logit_off", "i8"), ("part_off", "i8"), ("z1_off", "i8", (kM,)), ("z2_off", "i8", (kM, kT)),
    ("ha_off", "i8", (kM,)),
    ("pred_off", "i8"),
], align=True)

UNIT_DT = np.dtype([("agent", "i4"), ("slice", "i4"), ("nseg", "i4"), ("x32", "i4"), ("g", "i4", (kM,)),
                    ("r0", "i4", (kM,)), ("r1", "i4", (kM,)), ("lds", "i4", (kM,))])
E1REC_DT = np.dtype([("a_off", "i8"), ("h_off", "i8"), ("z_off", "i8"), ("hb", "i8", (kT,)), ("ts", "i4", (kT + 1,)),
                     ("O", "i4"), ("d2", "i4"), ("d01", "i4"), ("e0", "i4"), ("e1", "i4"), ("N1", "i4"),
                     ("blk", "i4")])


def e1_tpr(d2):
    """Threads per stage-1 row (rg_e1_tpr)."""
    return 2 if d2 > 32 else 1
ROLL_DT = np.dtype([("b_off", "i8"), ("out", "i8"), ("par", "i8", (3,)), ("ent", "i8"), ("Sf", "i4"), ("KB", "i4"),
                    ("rb", "i4"), ("nh", "i4"), ("t", "i4"), ("dd", "i4", (3,))], align=True)
TILE_DT = np.dtype([("agent", "i4"), ("m", "i4"), ("t", "i4"), ("r0", "i4"), ("r1", "i4"), ("c0", "i4"),
                    ("c1", "i4"), ("pad", "i4")])


def tables(plan):
    """(agents, units, (tiles1, tiles2, tiles3), iarena) copied from the plan."""
    L = _native.lib()
    sizes = np.zeros(16, np.int64)
    _native.check(L.pymdp_hip_ragged_tables(plan.handle, sizes.ctypes.data_as(_native._i64p), None, None, None,
                                            None))
    assert sizes[0] == AGENT_DT.itemsize, (int(sizes[0]), AGENT_DT.itemsize)
    assert sizes[1] == UNIT_DT.itemsize and sizes[3] == TILE_DT.itemsize
    ag = np.zeros(plan.n, AGENT_DT)
    units = np.zeros(int(sizes[2]), UNIT_DT)
    n1, n2, n3 = int(sizes[4]), int(sizes[5]), int(sizes[6])
    tiles = np.zeros(n1 + n2 + n3, TILE_DT)
    iarena = np.zeros(max(1, int(sizes[7])), np.int32)
    _native.check(L.pymdp_hip_ragged_tables(plan.handle, None, ag.ctypes.data, units.ctypes.data,
                                            tiles.ctypes.data, iarena.ctypes.data))
    return ag, units, (tiles[:n1], tiles[n1:n1 + n2], tiles[n1 + n2:]), iarena[:int(sizes[7])]


def e1_records(plan, stage=1):
    """Stage-1, -2 or -3 work-item records copied from the plan."""
    L = _native.lib()
    fn = {1: L.pymdp_hip_ragged_e1_records, 2: L.pymdp_hip_ragged_e2_records, 3: L.pymdp_hip_ragged_e3_records}[stage]
    cnt, size = np.zeros(1, np.int64), np.zeros(1, np.int64)
    _native.check(fn(plan.handle, None, cnt.ctypes.data_as(_native._i64p), size.ctypes.data_as(_native._i64p)))
    assert int(size[0]) == E1REC_DT.itemsize, (int(size[0]), E1REC_DT.itemsize)
    rec = np.zeros(int(cnt[0]), E1REC_DT)
    _native.check(fn(plan.handle, rec.ctypes.data, None, None))
    return rec


def roll_tasks(plan):
    """(rollout task records, first task of each level) copied from the plan."""
    L = _native.lib()
    cnt, size = np.zeros(1, np.int64), np.zeros(1, np.int64)
    start = np.zeros(kT + 1, np.int64)
    _native.check(L.pymdp_hip_ragged_roll_tasks(plan.handle, None, cnt.ctypes.data_as(_native._i64p),
                                                size.ctypes.data_as(_native._i64p),
                                                start.ctypes.data_as(_native._i64p)))
    assert int(size[0]) == ROLL_DT.itemsize, (int(size[0]), ROLL_DT.itemsize)
    rec = np.zeros(int(cnt[0]), ROLL_DT)
    _native.check(L.pymdp_hip_ragged_roll_tasks(plan.handle, rec.ctypes.data, None, None, None))
    return rec, start


def _chk(idx, n, what):
    idx = np.asarray(idx)
    if idx.size:
        lo, hi = int(idx.min()), int(idx.max())
        assert lo >= 0 and hi < n, f"{what}: index range [{lo}, {hi}] outside [0, {n})"
    return idx


def _ceil_log2(n):
    l = 0
    while (1 << l) < n:
        l += 1
    return l


def _rows_lg(n):
    return 0 if n <= 8 else _ceil_log2(n) - 3


class Emu:
    """Replays one step.  `params`/`obs` are the packed arenas as NumPy."""

    def __init__(self, plan, params: np.ndarray, obs: np.ndarray, carry: bool = False, out=None):
        self.plan = plan
        self.carry = carry
        self.info = plan.info()
        self.ag, self.units, (self.t1, self.t2, self.t3), self.iarena = tables(plan)
        self.e1rec = e1_records(plan)
        self.e2rec = e1_records(plan, 2)
        assert len(self.e2rec) == len(self.t2)
        self.e3rec = e1_records(plan, 3)
        assert len(self.e3rec) == len(self.t3)
        self.rolls, self.roll_start = roll_tasks(plan)
        assert len(self.e1rec) == len(self.t1)
        self.params = params.astype(np.float32)
        self.obs = obs.astype(np.int32)
        self.hist = np.full(max(1, self.info["hist_doubles"]), np.nan)
        self.score = np.full(max(1, self.info["score_doubles"]), np.nan)
        # stage-1 / stage-2 outputs are FP32
        self.z1 = np.full(max(1, self.info["z1_doubles"]), np.nan, np.float32)
        self.z2 = np.full(max(1, self.info["z2_doubles"]), np.nan, np.float32)
        self.out = np.full(max(1, int(plan.sizes[2])), np.nan, np.float32) if out is None else out.copy()
        self.act = np.full(max(1, int(plan.sizes[3])), -1, np.int32)
        check_units(self.ag, self.units, self.info)

    # ---------------------------------------------------------------- FPI
    def _segment(self, a, g, r0, nr, X, qv, part):
        """rg_fpi_segment: X is the unit's LDS slice [kCap] at the segment."""
        S_off = a["S_off"]
        R = int(a["gR"][g])
        n = int(a["gdim"][g][R - 1])
        ld = n | 1
        fl = int(a["gfac"][g][R - 1])
        d1 = int(a["gdim"][g][1]) if R == 3 else 1
        _chk(np.arange(nr), kRowsCap, "s_y")
        rows = X(np.arange(nr)[:, None] * ld + np.arange(n)[None, :])  # [nr, n]
        qc = qv[S_off[fl]:S_off[fl] + n]
        r = r0 + np.arange(nr)
        if R == 1:
            w = np.ones(nr)
        elif R == 2:
            q0 = qv[S_off[a["gfac"][g][0]]:]
            w = q0[_chk(r, int(a["gdim"][g][0]), "w q0")]
        else:
            q0 = qv[S_off[a["gfac"][g][0]]:]
            q1 = qv[S_off[a["gfac"][g][1]]:]
            i = r // d1
            w = q0[_chk(i, int(a["gdim"][g][0]), "w q0")] * q1[_chk(r - i * d1, d1, "w q1")]
        nrg = min(kNT // n, 1 + int(np.sqrt(np.float32(nr))))
        assert nrg * n <= kNT
        part[S_off[fl]:S_off[fl] + n] += w @ rows
        if R >= 2:
            lt = _rows_lg(n)
            assert (1 << lt) <= 64
            y = rows @ qc
            if R == 2:
                f0 = int(a["gfac"][g][0])
                _chk(r, int(a["S"][f0]), "m0 (R=2)")
                part[S_off[f0] + r] += y
            else:
                i = r // d1
                j = r - i * d1
                i_lo, i_hi = r0 // d1, (r0 + nr - 1) // d1
                ni = i_hi - i_lo + 1
                # the kernel's thread-per-i / thread-per-j pass: y indices
                # rb + j over j in [max(0, -rb), min(d1, nr - rb)) and the
                # clamped m_1 rows stay inside s_y; q_0 indices stay in range
                for ii in range(ni):
                    rb = (i_lo + ii) * d1 - r0
                    j0, j1 = max(0, -rb), min(d1, nr - rb)
                    if j1 > j0:
                        _chk(np.array([rb + j0, rb + j1 - 1]), nr, "m0 y")
                _chk(np.array([i_lo, i_hi]), int(a["gdim"][g][0]), "m1 q0")
                f0, f1 = int(a["gfac"][g][0]), int(a["gfac"][g][1])
                np.add.at(part, S_off[f1] + j, q0[i] * y)
                np.add.at(part, S_off[f0] + i, q1[j] * y)

    def fpi_agent(self, ai, units):
        a = self.ag[ai]
        P = self.params
        F, M = int(a["F"]), int(a["M"])
        totS = int(a["totS"])
        assert totS <= kF * kDim and M <= kM and F <= kF
        S, S_off = a["S"], a["S_off"]
        team = int(a["team"])
        assert len(units) == team and [int(u["slice"]) for u in units] == list(range(team))
        lp = np.zeros(totS)
        for f in range(F):
            if self.carry:
                d = self.out[_chk(int(a["pred_off"]) + S_off[f] + np.arange(S[f]), self.out.size, "prior")]
            else:
                d = P[_chk(int(a["D_off"][f]) + np.arange(S[f]), P.size, "D")]
            lp[S_off[f]:S_off[f] + S[f]] = np.log(np.maximum(d, MINVAL).astype(np.float64))
        o_idx = self.obs[_chk(int(a["obs_off"]) + np.arange(M), self.obs.size, "obs")]
        lds = []
        # x32 units hold the log-likelihood as floats, 2 kCap of them
        x32 = int(units[0]["x32"])
        assert all(int(u["x32"]) == x32 for u in units) and x32 in (0, 1)
        cap = 2 * kCap if x32 else kCap
        for u in units:
            x = np.full(cap, np.nan)
            for sg in range(int(u["nseg"])):
                g = int(u["g"][sg])
                R = int(a["gR"][g])
                n = int(a["gdim"][g][R - 1])
                ld = n | 1
                r0, r1 = int(u["r0"][sg]), int(u["r1"][sg])
                nr = r1 - r0
                prod = np.ones((nr, n))
                ej = (r0 + np.arange(nr))[:, None] * n + np.arange(n)[None, :]
                for m in range(M):
                    if a["mod_group"][m] != g:
                        continue
                    o = int(o_idx[m])
                    if o < 0 or o >= a["O"][m]:
                        continue
                    Km = int(a["K"][m])
                    assert ej.max() < Km
                    # outcome-major A_m: (o, position, k) at (o P + position) n + k
                    pos = (r0 + np.arange(nr))[:, None]
                    addr = (o * (Km // n) + pos) * n + np.arange(n)[None, :]
                    v = P[_chk(int(a["A_off"][m]) + addr, P.size, "A (ll)")]
                    prod *= np.maximum(v, MINVAL).astype(np.float64)
                dst = int(u["lds"][sg]) + np.arange(nr)[:, None] * ld + np.arange(n)[None, :]
                ll = np.log(prod)
                x[_chk(dst, cap, "s_x (ll)")] = ll.astype(np.float32).astype(np.float64) if x32 else ll
            lds.append(x)

        def softmax(src):
            out = np.zeros(totS)
            for f in range(F):
                assert S[f] <= kDim
                xx = src[S_off[f]:S_off[f] + S[f]]
                ex = np.exp(xx - xx.max())
                out[S_off[f]:S_off[f] + S[f]] = ex / ex.sum()
            return out

        log_q = np.zeros(totS)
        for it in range(int(a["iters"])):
            qv = softmax(log_q)
            parts = []
            for u, x in zip(units, lds):
                part = np.zeros(totS)
                for sg in range(int(u["nseg"])):
                    base = int(u["lds"][sg])

                    def X(idx, base=base, x=x):
                        v = x[_chk(base + idx, cap, "s_x")]
                        assert np.isfinite(v).all(), "read of an unset LDS log-likelihood entry"
                        return v

                    self._segment(a, int(u["g"][sg]), int(u["r0"][sg]), int(u["r1"][sg]) - int(u["r0"][sg]),
                                  X, qv, part)
                parts.append(part)
            if team > 1:
                po = int(a["part_off"]) + (it & 1) * team * totS
                _chk(po + np.arange(team * totS), self.info["part_doubles"], "team partials")
            tot = np.zeros(totS)
            for part in parts:
                tot += part
            log_q = lp + tot
        qv = softmax(log_q)
        self.out[_chk(int(a["qs_off"]) + np.arange(totS), self.out.size, "qs out")] = qv.astype(np.float32)
        h0 = int(a["hist_off"])
        self.hist[_chk(h0 + np.arange(totS), self.hist.size, "hist post")] = qv

    # ---------------------------------------------------------------- rollout
    def roll(self, rt):
        """rg_roll_kernel for one task: output o -> state s of entry o / S_f,
        entries (h, u, parents) of the level's action-sorted history list;
        B_f stored [U_f][K_B][S_f]."""
        Sf, KB, rb, nh, t = (int(rt[k]) for k in ("Sf", "KB", "rb", "nh", "t"))
        ne = 2 + rb
        assert 1 <= nh and nh * Sf <= kNT and 1 <= rb <= 3 and Sf <= kDim
        P = self.params
        ent = self.iarena[_chk(int(rt["ent"]) + np.arange(nh * ne), self.iarena.size, "roll entries")].reshape(nh, ne)
        for i in range(nh):
            h, u = int(ent[i][0]), int(ent[i][1])
            assert u >= 0
            w = np.ones(1)
            for j in range(rb):
                d = int(rt["dd"][j])
                base = int(rt["par"][j]) + (0 if t == 0 else int(ent[i][2 + j]) * d)
                w = np.kron(w, self.hist[_chk(base + np.arange(d), self.hist.size, "hist parent")])
            assert w.size == KB and np.isfinite(w).all(), "rollout read of an unset parent vector"
            addr = int(rt["b_off"]) + (u * KB + np.arange(KB))[:, None] * Sf + np.arange(Sf)[None, :]
            Bu = P[_chk(addr, P.size, "B")].astype(np.float64)  # [K_B, S_f]
            self.hist[_chk(int(rt["out"]) + h * Sf + np.arange(Sf), self.hist.size, "hist out")] = w @ Bu

    # ---------------------------------------------------------------- EFE stages
    def _hcount(self, a, m, t, p):
        f = int(a["ef"][m][p])
        return 1 if f < 0 else int(a["H"][t][f])

    def _hist_table(self, a, m, t, p, c0, c1):
        """History vectors [c0, c1) of padded dim p of modality m at level t
        from the agent record (padding dims: one history, q = 1)."""
        d = int(a["ed"][m][p])
        f = int(a["ef"][m][p])
        if f < 0:
            assert d == 1
            return np.ones((c1 - c0, 1))
        assert d == int(a["S"][f])
        idx = int(a["hist_off"]) + int(a["hist_lvl"][t][f]) + (c0 + np.arange(c1 - c0))[:, None] * d + \
            np.arange(d)[None, :]
        q = self.hist[_chk(idx, self.hist.size, "hist table")]
        assert np.isfinite(q).all()
        return q

    def e1(self, tl, rec):
        """rg_e1_kernel: rows [e0, e1) of the modality's row space [O1][d01]
        (row e = o d01 + pos: A_m[pos, o, :] for o < O, the FP32 sum_o A ln A
        table's row pos for o = O) times every last-axis history column."""
        ai, m, e0, e1 = (int(tl[k]) for k in ("agent", "m", "r0", "r1"))
        a = self.ag[ai]
        O = int(a["O"][m])
        O1 = O + 1
        d0, d1, d2 = (int(x) for x in a["ed"][m])
        d01 = d0 * d1
        T = int(a["T"])
        f = int(a["ef"][m][2])
        N1 = int(a["tstart"][m][T])
        tpr = e1_tpr(d2)
        nv = (d2 + 3) // 4
        nvt = (nv + tpr - 1) // tpr
        kp = 4 * nvt * tpr
        assert 1 <= nv <= 16 and kp <= kE1Ld
        rows_per_item = self.plan.tile_config()["e1_rows"] // tpr
        assert 0 <= e0 < e1 <= O1 * d01 and e1 - e0 <= rows_per_item
        assert (int(tl["c0"]), int(tl["c1"])) == (0, N1)
        assert int(a["K"][m]) == d01 * d2
        # the item's record: sizes and addresses as derived from the agent
        assert (int(rec["O"]), int(rec["d2"]), int(rec["d01"]), int(rec["e0"]), int(rec["e1"]), int(rec["N1"])) == \
            (O, d2, d01, e0, e1, N1)
        assert int(rec["a_off"]) == int(a["A_off"][m])
        assert int(rec["h_off"]) == int(a["ha_off"][m])
        assert int(rec["z_off"]) == int(a["z1_off"][m])
        for t in range(kT + 1):
            assert int(rec["ts"][t]) == int(a["tstart"][m][t])
        for t in range(T):
            assert int(rec["hb"][t]) == int(a["hist_off"]) + int(a["hist_lvl"][t][f])
        e = np.arange(e0, e1)
        o, pos = e // d01, e % d01
        # every 16-byte row load (up to 3 floats past the row: the params
        # arena keeps at least 3 floats of the agent after its last A block,
        # the table 16 bytes of tail padding)
        isA = o < O
        # outcome-major A_m: row e = o d01 + pos at e d2
        start = np.where(isA, int(rec["a_off"]) + e * d2, int(rec["h_off"]) + pos * d2)
        last = start + 4 * nv - 1
        _chk(np.concatenate([start[isA], last[isA]]), self.params.size, "A row (E1)")
        _chk(np.concatenate([start[~isA], last[~isA]]), self.info["ha_floats"] + 4, "sum A ln A row (E1)")
        # the rows: A in FP32, the table formed as rg_ha_kernel (FP32 xlogx
        # terms summed in FP64, stored FP32)
        pidx = start[:, None] + np.arange(d2)[None, :]
        Xa = self.params[_chk(np.where(isA[:, None], pidx, 0), self.params.size, "A (E1)")]
        Am = self.params[_chk(int(a["A_off"][m]) + np.arange(d01 * O * d2), self.params.size, "A block (E1)")]
        Am = Am.reshape(O, d01, d2)
        xl = np.where(Am == 0, np.float32(0), Am * np.log(np.maximum(Am, MINVAL)).astype(np.float32))
        ha = xl.astype(np.float64).sum(0).astype(np.float32)  # [d01, d2]
        X = np.where(isA[:, None], Xa, ha[pos]).astype(np.float64)  # [rows, d2]
        # history columns, staged 16 at a time as s_q[n][k] (k < kp)
        cols = []
        for n in range(N1):
            t = sum(1 for lvl in range(1, kT) if n >= int(rec["ts"][lvl]))
            h = n - int(rec["ts"][t])
            assert t < T and 0 <= h < a["H"][t][f]
            base = int(rec["hb"][t]) + h * d2
            col = self.hist[_chk(base + np.arange(d2), self.hist.size, "hist (E1)")]
            np.testing.assert_array_equal(col, self._hist_table(a, m, t, 2, h, h + 1)[0])
            assert np.isfinite(col).all()
            cols.append(col)
            _chk(np.array([(n % kE1Cols) * kE1Ld + kp - 1]), kE1Cols * kE1Ld, "s_q (E1)")
        Q = np.stack(cols)  # [N1, d2]
        Y = X @ Q.T  # [rows, N1]
        zi = int(rec["z_off"]) + np.arange(N1)[None, :] * (O1 * d01) + e[:, None]
        _chk(zi - int(a["z1_off"][m]), N1 * O1 * d01, "Z1 (block)")
        self.z1[_chk(zi, self.z1.size, "Z1")] = Y

    def e2(self, tl, rec):
        """rg_rows_kernel<2>: rows [r0, r1) of level t's stage-1 output rows
        (h2, o, i) (each d1 floats, contiguous) times the level's middle-dim
        histories; output block h2 of Z2 holds H1 columns of O1 d0 rows."""
        ai, m, t, r0, r1 = (int(tl[k]) for k in ("agent", "m", "t", "r0", "r1"))
        a = self.ag[ai]
        O1 = int(a["O"][m]) + 1
        d0, d1, _ = (int(x) for x in a["ed"][m])
        assert int(a["gR"][a["mod_group"][m]]) >= 2
        H1, H2 = self._hcount(a, m, t, 1), self._hcount(a, m, t, 2)
        f = int(a["ef"][m][1])
        od = O1 * d0
        tpr = e1_tpr(d1)
        nv = (d1 + 3) // 4
        assert 0 <= r0 < r1 <= H2 * od and r1 - r0 <= self.plan.tile_config()["e2_rows"] // tpr
        assert (int(tl["c0"]), int(tl["c1"])) == (0, H1)
        z1base = int(a["z1_off"][m]) + int(a["tstart"][m][t]) * od * d1
        assert (int(rec["a_off"]), int(rec["z_off"]), int(rec["d2"]), int(rec["d01"]), int(rec["e0"]), int(rec["e1"]),
                int(rec["N1"]), int(rec["blk"])) == (z1base, int(a["z2_off"][m][t]), d1, d0, r0, r1, H1, od)
        assert int(rec["hb"][0]) == int(a["hist_off"]) + int(a["hist_lvl"][t][f])
        assert int(rec["ts"][0]) == 0 and (rec["ts"][1:] == H1).all()
        Rr = r0 + np.arange(r1 - r0)
        # row loads (up to 3 floats past a row: Z1 has 16 bytes of tail padding)
        _chk(np.array([z1base + r0 * d1, z1base + (r1 - 1) * d1 + 4 * nv - 1]), self.z1.size + 4, "Z1 row (E2)")
        src = z1base + Rr[:, None] * d1 + np.arange(d1)[None, :]
        Z = self.z1[_chk(src, self.z1.size, "Z1 (E2 read)")]
        assert np.isfinite(Z).all(), "E2 reads an unset stage-1 value"
        # the level's middle-dim histories, staged 16 at a time as s_q[n][k]
        kp = 4 * ((nv + tpr - 1) // tpr) * tpr
        assert kp <= kE1Ld
        _chk(np.array([(kE1Cols - 1) * kE1Ld + kp - 1]), kE1Cols * kE1Ld, "s_q (E2)")
        hidx = int(rec["hb"][0]) + np.arange(H1)[:, None] * d1 + np.arange(d1)[None, :]
        assert int(rec["hb"][0]) == int(a["hist_off"]) + int(a["hist_lvl"][t][f])
        Q = self.hist[_chk(hidx, self.hist.size, "hist (E2)")]  # [H1, d1]
        assert np.isfinite(Q).all()
        Y = Z.astype(np.float64) @ Q.T  # [nr, H1]
        h2 = Rr // od
        oi = Rr - h2 * od
        zi = int(a["z2_off"][m][t]) + ((h2[:, None] * H1 + np.arange(H1)[None, :]) * od + oi[:, None])
        self.z2[_chk(zi, self.z2.size, "Z2")] = Y

    def e3(self, tl, rec):
        """rg_rows_kernel<3>: whole (h2, h1) pairs [r0 / O1, r1 / O1) of level
        t (rows (pair, o) of d0 floats) times the level's first-dim histories,
        each row's term of the combination score in LDS, summed per
        (pair, history) over the pair's O1 rows."""
        ai, m, t, r0, r1 = (int(tl[k]) for k in ("agent", "m", "t", "r0", "r1"))
        a = self.ag[ai]
        O = int(a["O"][m])
        O1 = O + 1
        d0 = int(a["ed"][m][0])
        H0, H1, H2 = (self._hcount(a, m, t, p) for p in range(3))
        f = int(a["ef"][m][0])
        tpr = e1_tpr(d0)
        nv = (d0 + 3) // 4
        rs = kE3Rows // tpr  # row slots of an item
        # a padding first axis: a thread per pair, up to 256 pairs
        assert r0 % O1 == 0 and r1 % O1 == 0 and 0 <= r0 < r1 <= H2 * H1 * O1
        assert r1 - r0 <= (kE3Rows * O1 if f < 0 else rs)
        assert (int(tl["c0"]), int(tl["c1"])) == (0, H0)
        R = int(a["gR"][a["mod_group"][m]])
        coff = int(a["C_off"][m]) + (t * O if a["C_time"][m] else 0)
        base = int(a["score_off"]) + int(a["score_base"][t][m])
        assert (int(rec["a_off"]), int(rec["h_off"]), int(rec["z_off"]), int(rec["d2"]), int(rec["d01"]),
                int(rec["e0"]), int(rec["e1"]), int(rec["N1"]), int(rec["blk"]), int(rec["O"])) == \
            (int(a["z2_off"][m][t]), coff, base, d0, 1 if R == 1 else 0, r0, r1, H0, O1, O)
        assert int(rec["hb"][0]) == (-1 if f < 0 else int(a["hist_off"]) + int(a["hist_lvl"][t][f]))
        assert int(rec["ts"][0]) == 0 and (rec["ts"][1:] == H0).all()
        # LDS: staged columns and the score terms [column][row slot]
        kp = 4 * ((nv + tpr - 1) // tpr) * tpr
        assert kp <= kE1Ld
        _chk(np.array([(kE3Cols - 1) * kE1Ld + kp - 1]), kE1Cols * kE1Ld, "s_q (E3)")
        if f < 0:
            _chk(np.array([O - 1]), kE1Cols * kE1Ld, "s_c (E3 pairs)")
        else:
            _chk(np.array([(kE3Cols - 1) * rs + (r1 - r0) - 1]), kE3Cols * kE3Rows, "s_t (E3)")
        buf = self.z1 if R == 1 else self.z2
        a_off = int(rec["a_off"])
        _chk(np.array([a_off + r0 * d0, a_off + (r1 - 1) * d0 + 4 * nv - 1]), buf.size + 4, "row (E3)")
        np_ = (r1 - r0) // O1
        src = a_off + r0 * d0 + np.arange(np_ * O1 * d0)
        Z = buf[_chk(src, buf.size, "Z2 (E3 read)")].reshape(np_, O1, d0).astype(np.float64)
        assert np.isfinite(Z).all(), "E3 reads an unset stage-2 value"
        if f < 0:
            assert d0 == 1 and H0 == 1
            Q = np.ones((1, 1))
        else:
            hidx = int(rec["hb"][0]) + np.arange(H0)[:, None] * d0 + np.arange(d0)[None, :]
            Q = self.hist[_chk(hidx, self.hist.size, "hist (E3)")]
            assert np.isfinite(Q).all()
        qo = np.einsum("poi,hi->pho", Z, Q)  # [np, H0, O1]
        C = self.params[_chk(coff + np.arange(O), self.params.size, "C")].astype(np.float64)
        q = qo[..., :O]
        qf = q.astype(np.float32)
        lq = np.log(np.maximum(qf, MINVAL)).astype(np.float32).astype(np.float64)
        v = (q * C).sum(-1) - np.where(qf != 0, q * lq, 0.0).sum(-1) + qo[..., O]
        si = base + (r0 // O1 + np.arange(np_))[:, None] * H0 + np.arange(H0)[None, :]
        _chk(si - base, H0 * H1 * H2, "score (block)")
        self.score[_chk(si, self.score.size, "score")] = v

    # ---------------------------------------------------------------- assemble
    def assemble(self, ai):
        a = self.ag[ai]
        Pn, T, F, M = int(a["P"]), int(a["T"]), int(a["F"]), int(a["M"])
        p2h = self.iarena[_chk(int(a["p2h_off"]) + np.arange(Pn * T * F), self.iarena.size, "p2h")]
        p2h = p2h.reshape(Pn, T, F)
        G = np.zeros(Pn)
        for t in range(T):
            for m in range(M):
                g = int(a["mod_group"][m])
                ix = np.full(Pn, int(a["score_base"][t][m]), np.int64)
                for j in range(int(a["gR"][g])):
                    f = int(a["gfac"][g][j])
                    assert (p2h[:, t, f] < a["H"][t][f]).all()
                    ix += p2h[:, t, f].astype(np.int64) * int(a["score_stride"][t][m][j])
                G += self.score[_chk(int(a["score_off"]) + ix, self.score.size, "score read")]
        assert np.isfinite(G).all()
        self.out[_chk(int(a["G_off"]) + np.arange(Pn), self.out.size, "G out")] = G.astype(np.float32)
        logit = float(a["gamma"]) * G
        if a["has_E"]:
            E = self.params[_chk(int(a["E_off"]) + np.arange(Pn), self.params.size, "E")]
            logit = logit + np.log(np.maximum(E, MINVAL).astype(np.float64))
        _chk(int(a["logit_off"]) + np.arange(Pn), self.score.size, "logit")
        ex = np.exp(logit - logit.max())
        q = (ex / ex.sum()).astype(np.float32)
        self.out[_chk(int(a["qpi_off"]) + np.arange(Pn), self.out.size, "q_pi out")] = q
        bx = int(np.argmax(q))
        pol0 = self.iarena[_chk(int(a["pol0_off"]) + bx * F + np.arange(F), self.iarena.size, "pol0")]
        self.act[_chk(int(a["act_off"]) + np.arange(F), self.act.size, "act")] = pol0
        S, S_off = a["S"], a["S_off"]
        for f in range(F):
            h = int(p2h[bx, 0, f])
            assert h < a["H"][0][f]
            src = int(a["hist_off"]) + int(a["hist_lvl"][0][f]) + h * int(S[f]) + np.arange(S[f])
            dst = int(a["pred_off"]) + int(S_off[f]) + np.arange(S[f])
            self.out[_chk(dst, self.out.size, "pred out")] = self.hist[_chk(src, self.hist.size, "pred hist")]

    def run(self):
        by_agent = {}
        for u in self.units:
            by_agent.setdefault(int(u["agent"]), []).append(u)
        for ai, us in by_agent.items():
            self.fpi_agent(ai, us)
        check_rolls(self.ag, self.rolls, self.roll_start, self.iarena)
        for t in range(kT):
            for rt in self.rolls[int(self.roll_start[t]):int(self.roll_start[t + 1])]:
                assert int(rt["t"]) == t
                self.roll(rt)
        for tl, rec in zip(self.t1, self.e1rec):
            self.e1(tl, rec)
        for tl, rec in zip(self.t2, self.e2rec):
            self.e2(tl, rec)
        for tl, rec in zip(self.t3, self.e3rec):
            self.e3(tl, rec)
        for ai in range(self.plan.n):
            self.assemble(ai)
        return self


def check_rolls(ag, rolls, start, iarena):
    """Rollout tasks: grouped by level, and every history vector of every
    (agent, level, factor) written by exactly one task entry, with that
    factor's B block and sizes."""
    assert list(start) == sorted(start) and int(start[-1]) == len(rolls)
    for t in range(kT):
        assert (rolls["t"][int(start[t]):int(start[t + 1])] == t).all()
    want = {}
    for a in ag:
        h0 = int(a["hist_off"])
        for t in range(int(a["T"])):
            for f in range(int(a["F"])):
                Sf = int(a["S"][f])
                for h in range(int(a["H"][t][f])):
                    want[h0 + int(a["hist_lvl"][t][f]) + h * Sf] = (int(a["B_off"][f]), Sf, int(a["BK"][f]),
                                                                     int(a["BR"][f]))
    got = {}
    for rt in rolls:
        ne = 2 + int(rt["rb"])
        b_base = int(rt["b_off"])
        for i in range(int(rt["nh"])):
            h = int(iarena[int(rt["ent"]) + i * ne])
            o = int(rt["out"]) + h * int(rt["Sf"])
            assert o not in got, "history vector written twice"
            got[o] = (b_base, int(rt["Sf"]), int(rt["KB"]), int(rt["rb"]))
    assert got == want


def check_units(ag, units, info):
    """FPI units: every agent's stage-0 rows covered exactly once by units of
    its team, segments inside the LDS slice, teams consecutive in the queue,
    the largest team below the resident grid."""
    seen = {}
    last_agent, order = None, []
    for u in units:
        ai = int(u["agent"])
        if ai != last_agent:
            assert ai not in order, f"units of agent {ai} are not consecutive"
            order.append(ai)
            last_agent = ai
        a = ag[ai]
        assert 1 <= int(u["nseg"]) <= kM
        used = []
        for sg in range(int(u["nseg"])):
            g = int(u["g"][sg])
            assert 0 <= g < int(a["G"])
            R = int(a["gR"][g])
            n = int(a["gdim"][g][R - 1])
            ld = n | 1
            rows = int(np.prod(a["gdim"][g][:R - 1])) if R > 1 else 1
            r0, r1 = int(u["r0"][sg]), int(u["r1"][sg])
            assert 0 <= r0 < r1 <= rows and r1 - r0 <= kRowsCap
            lo, hi = int(u["lds"][sg]), int(u["lds"][sg]) + (r1 - r0) * ld
            assert 0 <= lo and hi <= (2 * kCap if int(u["x32"]) else kCap)
            used.append((lo, hi))
            seen.setdefault((ai, g), []).append((r0, r1))
        used.sort()
        for (a0, b0), (a1, _) in zip(used, used[1:]):
            assert b0 <= a1, "overlapping LDS segments"
    assert len(order) == len(ag)
    for ai, a in enumerate(ag):
        assert 1 <= int(a["team"]) <= kTeamMax
        for g in range(int(a["G"])):
            R = int(a["gR"][g])
            rows = int(np.prod(a["gdim"][g][:R - 1])) if R > 1 else 1
            segs = sorted(seen[(ai, g)])
            assert segs[0][0] == 0 and segs[-1][1] == rows
            for (x0, x1), (y0, _) in zip(segs, segs[1:]):
                assert x1 == y0
        if int(a["team"]) > 1:
            assert int(a["part_off"]) + 2 * int(a["team"]) * int(a["totS"]) <= info["part_doubles"]
        # float slices only for an agent whose stage-0 rows do not fit one
        # slice of kCap doubles
        x32 = {int(u["x32"]) for u in units if int(u["agent"]) == ai}
        assert len(x32) == 1
        if x32 == {1}:
            need, longest = 0, 0
            for g in range(int(a["G"])):
                R = int(a["gR"][g])
                rows = int(np.prod(a["gdim"][g][:R - 1])) if R > 1 else 1
                need += rows * (int(a["gdim"][g][R - 1]) | 1)
                longest = max(longest, rows)
            assert need > kCap or longest > kRowsCap or int(a["G"]) > kM, ai


def check_bounds_only(plan):
    """Index-range checks of every agent, unit and tile of a (large) plan
    without data: the extreme index of each access, from the same formulas."""
    info = plan.info()
    ag, units, (t1, t2, t3), iarena = tables(plan)
    check_units(ag, units, info)
    params_n = int(plan.sizes[0])
    for a in ag:
        assert int(a["totS"]) <= kF * kDim
        for m in range(int(a["M"])):
            assert int(a["A_off"][m]) + int(a["O"][m]) * int(a["K"][m]) <= params_n
        for f in range(int(a["F"])):
            assert int(a["B_off"][f]) + int(a["S"][f]) * int(a["BK"][f]) * int(a["U"][f]) <= params_n
        hl = int(a["totS"]) + sum(int(a["H"][t][f]) * int(a["S"][f]) for t in range(int(a["T"]))
                                  for f in range(int(a["F"])))
        assert int(a["hist_off"]) + hl <= info["hist_doubles"]
        assert int(a["logit_off"]) + int(a["P"]) <= info["score_doubles"]
        assert int(a["p2h_off"]) + int(a["P"]) * int(a["T"]) * int(a["F"]) <= iarena.size
        assert int(a["pol0_off"]) + int(a["P"]) * int(a["F"]) <= iarena.size
        T = int(a["T"])
        for m in range(int(a["M"])):
            d0, d1, d2 = (int(x) for x in a["ed"][m])
            O1 = int(a["O"][m]) + 1
            assert d0 * d1 * d2 == int(a["K"][m])
            N1 = int(a["tstart"][m][T])
            assert int(a["z1_off"][m]) + N1 * O1 * d0 * d1 <= info["z1_doubles"]
            R = int(a["gR"][a["mod_group"][m]])
            for t in range(T):
                H = [1 if a["ef"][m][p] < 0 else int(a["H"][t][a["ef"][m][p]]) for p in range(3)]
                if R == 1:
                    assert int(a["z2_off"][m][t]) + H[2] * O1 <= info["z1_doubles"]
                else:
                    assert int(a["z2_off"][m][t]) + H[2] * H[1] * O1 * d0 <= info["z2_doubles"]
                blk = int(a["score_base"][t][m]) + H[0] * H[1] * H[2]
                assert int(a["score_off"]) + blk <= info["score_doubles"]
    rolls, start = roll_tasks(plan)
    check_rolls(ag, rolls, start, iarena)
    for rt in rolls:
        Sf, KB, rb, nh = (int(rt[k]) for k in ("Sf", "KB", "rb", "nh"))
        ne = 2 + rb
        assert 1 <= nh and 1 <= rb <= 3
        assert int(rt["ent"]) + nh * ne <= iarena.size
        ent = iarena[int(rt["ent"]):int(rt["ent"]) + nh * ne].reshape(nh, ne)
        assert ent[:, 1].min() >= 0
        assert nh * Sf <= kNT
        assert int(rt["b_off"]) + (int(ent[:, 1].max()) + 1) * KB * Sf <= params_n
        for j in range(rb):
            hi = 0 if int(rt["t"]) == 0 else int(ent[:, 2 + j].max())
            assert int(rt["par"][j]) + (hi + 1) * int(rt["dd"][j]) <= info["hist_doubles"]
        assert int(rt["out"]) + (int(ent[:, 0].max()) + 1) * Sf <= info["hist_doubles"]
    e1_rows = plan.tile_config()["e1_rows"]
    for tl in t1:
        a = ag[int(tl["agent"])]
        m = int(tl["m"])
        O1 = int(a["O"][m]) + 1
        d0, d1, d2 = (int(x) for x in a["ed"][m])
        e0, e1 = int(tl["r0"]), int(tl["r1"])
        assert 0 <= e0 < e1 <= O1 * d0 * d1 and e1 - e0 <= e1_rows // e1_tpr(d2) and 1 <= d2 <= kDim
        assert (int(tl["c0"]), int(tl["c1"])) == (0, int(a["tstart"][m][int(a["T"])]))
    # stage-1 records: every row load (up to 3 floats past a row), history
    # column and output element stays inside its buffer
    rec = e1_records(plan)
    assert len(rec) == len(t1)
    if len(rec):
        O_, d2_, d01_ = rec["O"].astype(np.int64), rec["d2"].astype(np.int64), rec["d01"].astype(np.int64)
        e0_, e1_ = rec["e0"].astype(np.int64), rec["e1"].astype(np.int64)
        nv4 = 4 * ((d2_ + 3) // 4)
        assert (rec["e0"] == t1["r0"]).all() and (rec["e1"] == t1["r1"]).all()
        # A rows: the last A row an item touches (A_m is outcome-major, row
        # e = o d01 + pos at e d2)
        hasA = e0_ < O_ * d01_
        last_a = rec["a_off"] + (np.minimum(e1_, O_ * d01_) - 1) * d2_ + nv4 - 1
        assert (rec["a_off"] >= 0).all() and (last_a[hasA] < params_n).all()
        hasH = e1_ > O_ * d01_
        last_h = rec["h_off"] + (e1_ - 1 - O_ * d01_) * d2_ + nv4 - 1
        assert (rec["h_off"] >= 0).all() and (last_h[hasH] < info["ha_floats"] + 4).all()
        N1 = rec["N1"].astype(np.int64)
        for t in range(kT):
            lvl = rec["ts"][:, t + 1] > rec["ts"][:, t]
            last_hist = rec["hb"][:, t] + (rec["ts"][:, t + 1] - rec["ts"][:, t]) * d2_
            assert (rec["hb"][lvl, t] >= 0).all() and (last_hist[lvl] <= info["hist_doubles"]).all()
        assert (rec["ts"][:, 0] == 0).all() and (rec["ts"][:, kT] == N1).all()
        last = rec["z_off"] + (N1 - 1) * (O_ + 1) * d01_ + e1_ - 1
        assert (rec["z_off"] >= 0).all() and (last < info["z1_doubles"]).all()
    e2_rows = plan.tile_config()["e2_rows"]
    rec2 = e1_records(plan, 2)
    assert len(rec2) == len(t2)
    for tl, rc in zip(t2, rec2):
        a = ag[int(tl["agent"])]
        m, t = int(tl["m"]), int(tl["t"])
        O1 = int(a["O"][m]) + 1
        d0, d1, _ = (int(x) for x in a["ed"][m])
        H = [1 if a["ef"][m][p] < 0 else int(a["H"][t][a["ef"][m][p]]) for p in range(3)]
        r0, r1 = int(tl["r0"]), int(tl["r1"])
        od = O1 * d0
        assert 0 <= r0 < r1 <= H[2] * od and r1 - r0 <= e2_rows // e1_tpr(d1)
        assert (int(tl["c0"]), int(tl["c1"])) == (0, H[1])
        assert (int(rc["e0"]), int(rc["e1"]), int(rc["d2"]), int(rc["N1"]), int(rc["blk"])) == (r0, r1, d1, H[1], od)
        last_row = int(rc["a_off"]) + (r1 - 1) * d1 + 4 * ((d1 + 3) // 4) - 1
        assert int(rc["a_off"]) >= 0 and last_row < info["z1_doubles"] + 4
        assert int(rc["hb"][0]) >= 0 and int(rc["hb"][0]) + H[1] * d1 <= info["hist_doubles"]
        h2 = (r1 - 1) // od
        last = int(rc["z_off"]) + h2 * H[1] * od + (H[1] - 1) * od + (r1 - 1 - h2 * od)
        assert int(rc["z_off"]) >= 0 and last < info["z2_doubles"]
    rec3 = e1_records(plan, 3)
    assert len(rec3) == len(t3)
    for tl, rc in zip(t3, rec3):
        a = ag[int(tl["agent"])]
        m, t = int(tl["m"]), int(tl["t"])
        O1 = int(a["O"][m]) + 1
        d0 = int(a["ed"][m][0])
        H = [1 if a["ef"][m][p] < 0 else int(a["H"][t][a["ef"][m][p]]) for p in range(3)]
        r0, r1 = int(tl["r0"]), int(tl["r1"])
        assert r0 % O1 == 0 and r1 % O1 == 0 and 0 <= r0 < r1 <= H[2] * H[1] * O1
        pad = a["ef"][m][0] < 0
        assert r1 - r0 <= (kE3Rows * O1 if pad else kE3Rows // e1_tpr(d0)) and (int(tl["c0"]), int(tl["c1"])) == (0, H[0])
        assert (int(rc["e0"]), int(rc["e1"]), int(rc["d2"]), int(rc["N1"]), int(rc["blk"])) == (r0, r1, d0, H[0], O1)
        lim = info["z1_doubles"] if int(rc["d01"]) else info["z2_doubles"]
        assert int(rc["a_off"]) >= 0 and int(rc["a_off"]) + (r1 - 1) * d0 + 4 * ((d0 + 3) // 4) - 1 < lim + 4
        assert int(rc["h_off"]) >= 0 and int(rc["h_off"]) + O1 - 1 <= params_n
        if int(rc["hb"][0]) >= 0:
            assert int(rc["hb"][0]) + H[0] * d0 <= info["hist_doubles"]
        else:
            assert d0 == 1 and H[0] == 1
        assert int(rc["z_off"]) >= 0 and int(rc["z_off"]) + (r1 // O1 - 1) * H[0] + H[0] - 1 < info["score_doubles"]
    return info
