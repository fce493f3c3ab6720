"""Host emulation of the batched ragged kernels (pymdp_amd/csrc/ragged.hip).

Reads the plan's own device tables (agent records, work items, int arena)
and replays the three kernels' index arithmetic in NumPy, asserting that
every access stays inside its buffer (LDS arrays, workspace slots, arenas),
and computing the same values, so that the plan and the kernels' indexing
are checked on the CPU and their results compared with the oracle.

Test infrastructure only.
"""
from __future__ import annotations

import ctypes

import numpy as np

from pymdp_amd import _native

MINVAL = np.float32(2.220446049250313e-16)
kF, kM, kR, kT, kDim, kNT = 8, 8, 3, 4, 64, 256
kHC, kFpiScr, kTileFloats, kHaPart = 16, 2048, 4096, 512
TILE_LDS = kTileFloats + 2 * kDim

AGENT_DT = np.dtype([
    ("F", "i4"), ("M", "i4"), ("T", "i4"), ("P", "i4"), ("G", "i4"), ("iters", "i4"), ("totS", "i4"),
    ("has_E", "i4"), ("gamma", "f4"), ("scr_lds", "i4"), ("ll_total", "i4"), ("xa_off", "i4"),
    ("xb_off", "i4"), ("scr_total", "i4"),
    ("S", "i4", (kF,)), ("S_off", "i4", (kF + 1,)), ("O", "i4", (kM,)), ("mod_group", "i4", (kM,)),
    ("gR", "i4", (kM,)), ("gfac", "i4", (kM, kR)), ("gdim", "i4", (kM, kR)), ("gll", "i4", (kM,)),
    ("BR", "i4", (kF,)), ("Bdep", "i4", (kF, 3)), ("BK", "i4", (kF,)), ("U", "i4", (kF,)),
    ("C_time", "i4", (kM,)),
    ("H", "i4", (kT, kF)), ("hist_lvl", "i4", (kT, kF)), ("roll_lvl", "i4", (kT, kF)),
    ("score_base", "i4", (kT, kM)), ("score_stride", "i4", (kT, kM, kR)),
    ("K", "i8", (kM,)), ("A_off", "i8", (kM,)), ("B_off", "i8", (kF,)), ("C_off", "i8", (kM,)),
    ("D_off", "i8", (kF,)), ("E_off", "i8"),
    ("obs_off", "i8"), ("qs_off", "i8"), ("G_off", "i8"), ("qpi_off", "i8"), ("act_off", "i8"),
    ("hist_off", "i8"), ("roll_off", "i8"), ("p2h_off", "i8"), ("pol0_off", "i8"), ("score_off", "i8"),
    ("logit_off", "i8"),
], align=True)


def tables(plan):
    L = _native.lib()
    sizes = np.zeros(8, np.int64)
    _native.check(L.pymdp_hip_ragged_tables(plan.handle, sizes.ctypes.data_as(_native._i64p), None, None, None,
                                            None))
    assert sizes[0] == AGENT_DT.itemsize, (int(sizes[0]), AGENT_DT.itemsize)
    assert sizes[1] == 20
    ag = np.zeros(plan.n, AGENT_DT)
    order = np.zeros(plan.n, np.int32)
    items = np.zeros((int(sizes[2]), 5), np.int32)
    iarena = np.zeros(max(1, int(sizes[3])), np.int32)
    _native.check(L.pymdp_hip_ragged_tables(plan.handle, None, ag.ctypes.data, order.ctypes.data,
                                            items.ctypes.data, iarena.ctypes.data))
    return ag, order, items, iarena[:int(sizes[3])], int(sizes[4])


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


class Emu:
    """Replays one step.  `params`/`obs` are the packed arenas as NumPy."""

    def __init__(self, plan, params: np.ndarray, obs: np.ndarray):
        self.plan = plan
        self.info = plan.info()
        self.ag, self.order, self.items, self.iarena, self.grid_max = tables(plan)
        self.params = params.astype(np.float32)
        self.obs = obs.astype(np.int32)
        self.hist = np.full(max(1, self.info["hist_doubles"]), np.nan)
        self.score = np.full(max(1, self.info["score_doubles"]), np.nan)
        self.out = np.full(max(1, int(plan.sizes[2])), np.nan, np.float32)
        self.act = np.full(max(1, int(plan.sizes[3])), -1, np.int32)
        assert sorted(self.order.tolist()) == list(range(plan.n))
        assert self.info["grid_fpi"] <= self.grid_max and self.info["grid_efe"] <= self.grid_max

    # ---------------------------------------------------------------- FPI
    def fpi(self, ai: int):
        a = self.ag[ai]
        P = self.params
        F, M, G = int(a["F"]), int(a["M"]), int(a["G"])
        totS = int(a["totS"])
        assert totS <= kF * kDim and M <= kM and F <= kF
        S, S_off = a["S"], a["S_off"]
        scr_n = kFpiScr if a["scr_lds"] else self.info["ws_fpi"]
        assert int(a["scr_total"]) <= scr_n
        scr = np.zeros(scr_n)
        lp = np.zeros(totS)
        for f in range(F):
            d = P[_chk(int(a["D_off"][f]) + np.arange(S[f]), P.size, "D")]
            lp[S_off[f]:S_off[f] + S[f]] = np.log(np.maximum(d, MINVAL).astype(np.float64))
        o_idx = self.obs[_chk(int(a["obs_off"]) + np.arange(M), self.obs.size, "obs")]
        for g in range(G):
            R = int(a["gR"][g])
            Kg = int(np.prod(a["gdim"][g][:R]))
            e = np.arange(Kg)
            prod = np.ones(Kg)
            for m in range(M):
                if a["mod_group"][m] != g:
                    continue
                o = int(o_idx[m])
                if o < 0 or o >= a["O"][m]:
                    continue
                assert int(a["K"][m]) == Kg
                x = P[_chk(int(a["A_off"][m]) + o * int(a["K"][m]) + e, P.size, "A (ll)")]
                prod *= np.maximum(x, MINVAL).astype(np.float64)
            scr[_chk(int(a["gll"][g]) + e, int(a["scr_total"]), "ll")] = np.log(prod)
        log_q = np.zeros(totS)
        qv = np.zeros(totS)

        def softmax(src):
            out = np.zeros(totS)
            for f in range(F):
                assert S[f] <= kDim
                x = src[S_off[f]:S_off[f] + S[f]]
                ex = np.exp(x - x.max())
                out[S_off[f]:S_off[f] + S[f]] = ex / ex.sum()
            return out

        for _ in range(int(a["iters"])):
            qv = softmax(log_q)
            log_q = lp.copy()
            for g in range(G):
                R = int(a["gR"][g])
                src_off, src_len = int(a["gll"][g]), int(np.prod(a["gdim"][g][:R]))
                for k in range(R):
                    c = R - 1 - k
                    n = int(a["gdim"][g][c])
                    f = int(a["gfac"][g][c])
                    lg = _ceil_log2(n)
                    assert lg <= 6
                    rows = int(np.prod(a["gdim"][g][:c])) if c > 0 else 1
                    assert rows * n == src_len
                    X = scr[_chk(src_off + np.arange(rows * n), int(a["scr_total"]), "stage src")].reshape(rows, n)
                    if c == 0:
                        w = np.ones(1)
                    elif c == 1:
                        w = qv[_chk(S_off[a["gfac"][g][0]] + np.arange(rows), totS, "w1")]
                    else:
                        d1 = int(a["gdim"][g][1])
                        r = np.arange(rows)
                        w = qv[_chk(S_off[a["gfac"][g][0]] + r // d1, totS, "w2a")] * \
                            qv[_chk(S_off[a["gfac"][g][1]] + r % d1, totS, "w2b")]
                    log_q[S_off[f]:S_off[f] + n] += w @ X
                    if c > 0:
                        dst_off = int(a["xb_off"]) if (k & 1) else int(a["xa_off"])
                        lim = int(a["scr_total"]) if (k & 1) else int(a["xb_off"])
                        _chk(dst_off + np.arange(rows), lim, "stage dst")
                        scr[dst_off:dst_off + rows] = X @ qv[S_off[f]:S_off[f] + n]
                        src_off, src_len = dst_off, rows
        qv = softmax(log_q)
        self.out[_chk(int(a["qs_off"]) + np.arange(totS), self.out.size, "qs out")] = qv.astype(np.float32)
        h0 = int(a["hist_off"])
        self.hist[_chk(h0 + np.arange(totS), self.hist.size, "hist post")] = qv
        for t in range(int(a["T"])):
            for f in range(F):
                Sf, rb, KB, Uf = int(S[f]), int(a["BR"][f]), int(a["BK"][f]), int(a["U"][f])
                nh = int(a["H"][t][f])
                Bf = P[_chk(int(a["B_off"][f]) + np.arange(Sf * KB * Uf), P.size, "B")].reshape(Sf, KB, Uf)
                for h in range(nh):
                    rec = self.iarena[_chk(int(a["roll_off"]) + int(a["roll_lvl"][t][f]) + h * (1 + rb)
                                           + np.arange(1 + rb), self.iarena.size, "roll rec")]
                    u = int(rec[0])
                    assert 0 <= u < Uf
                    w = np.ones(1)
                    for j in range(rb):
                        d = int(a["Bdep"][f][j])
                        if t == 0:
                            base = int(S_off[d])
                        else:
                            assert 0 <= rec[1 + j] < a["H"][t - 1][d]
                            base = int(a["hist_lvl"][t - 1][d]) + int(rec[1 + j]) * int(S[d])
                        qp = self.hist[_chk(h0 + base + np.arange(S[d]), self.hist.size, "hist parent")]
                        w = np.kron(w, qp)
                    assert w.size == KB
                    dst = h0 + int(a["hist_lvl"][t][f]) + h * Sf + np.arange(Sf)
                    self.hist[_chk(dst, self.hist.size, "hist out")] = Bf[:, :, u].astype(np.float64) @ w

    # ---------------------------------------------------------------- EFE item
    def efe_item(self, it):
        ai, m, t, c0, c1 = (int(x) for x in it)
        a = self.ag[ai]
        P = self.params
        g = int(a["mod_group"][m])
        R = int(a["gR"][g])
        O = int(a["O"][m])
        O1 = O + 1
        K = int(a["K"][m])
        n = int(a["gdim"][g][R - 1])
        pre = K // n
        assert pre * n == K
        hc = c1 - c0
        assert 1 <= hc <= kHC and c1 <= a["H"][t][a["gfac"][g][R - 1]]
        h0 = int(a["hist_off"])
        fR = int(a["gfac"][g][R - 1])
        qh_idx = h0 + int(a["hist_lvl"][t][fR]) + c0 * n + np.arange(hc * n)
        _chk(np.arange(hc * n), kHC * kDim, "s_qh")
        qh = self.hist[_chk(qh_idx, self.hist.size, "qh")].reshape(hc, n)
        assert np.isfinite(qh).all()
        ws_n = self.info["ws_efe"]
        Z = np.full(ws_n, np.nan)
        np1 = n + 1
        pt = kTileFloats // (O * np1)
        pt = min(pt, kHaPart // np1)
        pt = max(pt, 1)
        pt = min(pt, pre)
        for p0 in range(0, pre, pt):
            np_ = min(pt, pre - p0)
            ncol = np_ * n
            oc = min(max(kNT // ncol, 1), O)
            _chk(np.array([0, (O * np_ - 1) * np1 + n - 1]), TILE_LDS, "s_tile")
            _chk(np.array([0, (np_ - 1) * np1 + n - 1]), kHaPart, "s_ha")
            _chk(np.array([0, oc * ncol - 1]), kHaPart, "s_hap")
            gidx = int(a["A_off"][m]) + np.arange(O)[:, None] * K + p0 * n + np.arange(np_ * n)[None, :]
            tile = P[_chk(gidx, P.size, "A tile")].reshape(O, np_, n).astype(np.float64)
            t32 = tile.astype(np.float32)
            xl = np.where(t32 == 0, 0.0, t32.astype(np.float64) * np.log(np.maximum(t32, MINVAL).astype(np.float64)))
            ha = xl.sum(0)  # [np, n]
            X = np.concatenate([tile, ha[None]], 0)  # [O1, np, n]
            Z1 = np.einsum("opi,hi->hop", X, qh)  # [hc, O1, np]
            zi = (np.arange(hc)[:, None, None] * O1 + np.arange(O1)[None, :, None]) * pre + p0 + \
                np.arange(np_)[None, None, :]
            Z[_chk(zi, ws_n, "Z1")] = Z1
        zin_off, Bin, rows_in = 0, hc, pre
        for j in range(R - 2, -1, -1):
            nj = int(a["gdim"][g][j])
            fj = int(a["gfac"][g][j])
            Hj = int(a["H"][t][fj])
            qj = self.hist[_chk(h0 + int(a["hist_lvl"][t][fj]) + np.arange(Hj * nj), self.hist.size, "qj")]
            qj = qj.reshape(Hj, nj)
            rows_out = rows_in // nj
            assert rows_out * nj == rows_in
            rt = min((TILE_LDS * 4 // 8) // (nj + 1), O1 * rows_out)
            _chk(np.array([0, (rt - 1) * (nj + 1) + nj - 1]), TILE_LDS * 4 // 8, "stage tile")
            _chk(np.array([0, min(kHC, Hj) * nj - 1]), kHC * kDim, "stage qh")
            zout_off = zin_off + Bin * O1 * rows_in
            zin = Z[_chk(zin_off + np.arange(Bin * O1 * rows_in), ws_n, "zin")].reshape(Bin, O1, rows_out, nj)
            assert np.isfinite(zin).all()
            zo = np.einsum("borj,hj->bhor", zin, qj)
            _chk(zout_off + np.arange(Bin * Hj * O1 * rows_out), ws_n, "zout")
            Z[zout_off:zout_off + zo.size] = zo.reshape(-1)
            zin_off, Bin, rows_in = zout_off, Bin * Hj, rows_out
        assert rows_in == 1
        zf = Z[zin_off:zin_off + Bin * O1].reshape(Bin, O1)
        assert np.isfinite(zf).all()
        coff = int(a["C_off"][m]) + (t * O if a["C_time"][m] else 0)
        C = P[_chk(coff + np.arange(O), P.size, "C")].astype(np.float64)
        q = zf[:, :O]
        qf = q.astype(np.float32)
        lq = np.log(np.maximum(qf, MINVAL).astype(np.float64))
        ent = -np.where(qf != 0, q * lq, 0.0).sum(1)
        v = ent + q @ C + zf[:, O]
        stride = int(a["score_stride"][t][m][R - 1])
        base = int(a["score_off"]) + int(a["score_base"][t][m]) + c0 * stride
        blk = int(a["score_base"][t][m]) + int(np.prod([a["H"][t][a["gfac"][g][j]] for j in range(R)]))
        _chk(base + np.arange(Bin), int(a["score_off"]) + blk, "score (block)")
        self.score[_chk(base + np.arange(Bin), self.score.size, "score")] = v

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

    def run(self):
        for ai in self.order:
            self.fpi(int(ai))
        for it in self.items:
            self.efe_item(it)
        for ai in range(self.plan.n):
            self.assemble(ai)
        return self


def check_bounds_only(plan, total_hist=None):
    """Index-range checks of every agent and item of a (large) plan without
    data: the extreme index of each access, from the same formulas."""
    info = plan.info()
    ag, order, items, iarena, grid_max = tables(plan)
    params_n = int(plan.sizes[0])
    for a in ag:
        assert int(a["scr_total"]) <= (kFpiScr if a["scr_lds"] else info["ws_fpi"])
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
    for it in items:
        ai, m, t, c0, c1 = (int(x) for x in it)
        a = ag[ai]
        g = int(a["mod_group"][m])
        R = int(a["gR"][g])
        O1 = int(a["O"][m]) + 1
        K = int(a["K"][m])
        n = int(a["gdim"][g][R - 1])
        hc = c1 - c0
        assert 1 <= hc <= kHC
        need = hc * O1 * (K // n)
        Bin, rows = hc, K // n
        for j in range(R - 2, -1, -1):
            Hj = int(a["H"][t][a["gfac"][g][j]])
            rows //= int(a["gdim"][g][j])
            Bin *= Hj
            need += Bin * O1 * rows
        assert need <= info["ws_efe"], (need, info["ws_efe"])
        np1 = n + 1
        pt = max(1, min(kTileFloats // ((O1 - 1) * np1), kHaPart // np1))
        pt = min(pt, K // n)
        assert ((O1 - 1) * pt - 1) * np1 + n - 1 < TILE_LDS
        assert (pt - 1) * np1 + n - 1 < kHaPart
        ncol = pt * n
        assert min(max(kNT // ncol, 1), O1 - 1) * ncol <= kHaPart
    return info
