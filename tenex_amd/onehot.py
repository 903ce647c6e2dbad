"""One-hot FP4 MFMA counts for the pairs of few levels (csrc/onehot.hip).

A pair (x, y) whose (b_x^2 - 1) * max(b_y - 1, 1) indicator cells are few
costs far less as an exact contraction of 0/1 indicator rows (FP4 operands,
FP32 sums, exact below 2^24 samples) than as one LDS atomic per sample.
The host plan here

  * orders the sources by level count as the histogram planner does
    (tnx_te_plan: ascending, ties by list position) and packs them into
    128-column groups (src_cols(b_y) columns each, no source split);
  * packs the targets, ascending by level count, into 128-row groups
    (b_x^2 - 1 rows each);
  * gives row group r the leading column groups g whose pairs all satisfy
    (BX_r^2 - 1) * src_cols(BY_g) <= thr (BX, BY the groups' largest
    level counts),
    so each target's one-hot sources are a prefix of the sorted sources:
    src_skip[x] tells the histogram planner to leave exactly those out;
  * deals tiles of two row groups x two column groups to the workgroups.

Every pair keeps the same float TE on either path (exact integer counts and
the same fixed-point entropy sums), so the split changes only the time.
No reference counterpart: the reference counts every pair with per-sample
atomics (tenex/csrc/full_smem.cu:31-142, tenex/csrc/adaptive_smem.cu:30-149).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from ._lib import call, ptr

GROUP = 128            # rows (columns) of a row (column) group
ONEHOT_THR = 400       # largest (b_x^2 - 1) * b_y a one-hot pair may have (long series)
ONEHOT_THR_SHORT = 1000   # the same for series of at most SHORT_L samples
SHORT_L = 6000
ONEHOT_MAX_LEVELS = 64  # sources (a target's b_x^2 - 1 rows fit one 128-row group: b_x <= 11)
ONEHOT_MAX_L = 65535   # counts are staged as u16
BAND = 4               # row panels per band of the tile order


@dataclass
class OnehotPlan:
    args: _lib.OnehotArgs
    src_skip: np.ndarray          # int32 (n_vars,): sources left to this plan per target
    n_pairs: int                  # pairs counted here (target != source)
    n_tiles: int
    cells: int                    # indicator cells of all tiles' active sub-tiles
    keep: list = field(default_factory=list)   # device arrays the args point into
    host: dict = field(default_factory=dict)   # host copies of the plan arrays
    sources: np.ndarray | None = None          # the source list the plan was made for


def default_threshold(L: int) -> int:
    """One-hot threshold for series of L samples: the histogram kernels'
    per-pair work (clearing and reducing the histogram) weighs more against
    the counting on short series, which favours the contraction there."""
    return ONEHOT_THR_SHORT if L <= SHORT_L else ONEHOT_THR


def _groups(sizes: np.ndarray, cap: int = GROUP):
    """Consecutive items into groups of at most cap rows: (group of each
    item, first row of each item inside its group, number of groups)."""
    g = np.empty(len(sizes), np.int64)
    off = np.empty(len(sizes), np.int64)
    cur, used = 0, 0
    for i, sz in enumerate(sizes):
        if used + sz > cap:
            cur += 1
            used = 0
        g[i] = cur
        off[i] = used
        used += sz
    return g, off, (cur + 1 if len(sizes) else 0)


def src_cols(b, src_ie: bool):
    """Operand rows (stored columns) of a source of b levels."""
    return np.maximum(b - 1, 1) if src_ie else b


def plan(npv_host: np.ndarray, L: int, device, targets=None, sources=None,
         thr: int = ONEHOT_THR, src_ie: bool = True) -> OnehotPlan | None:
    """One-hot plan for the pairs of targets x sources that qualify, or None
    when none does (thr <= 0, L too long, or no small pairs).  A pair
    qualifies when its (b_x^2 - 1) * cols(b_y) indicator cells are at most
    thr, cols(b) = b, or with src_ie max(b - 1, 1) (the source's last level
    then follows from the target's marginals)."""
    if thr <= 0 or L < 1 or L > ONEHOT_MAX_L:
        return None
    npv = np.asarray(npv_host, dtype=np.int64)
    n = len(npv)
    tg = np.arange(n, dtype=np.int64) if targets is None else np.asarray(targets, np.int64)
    sr = np.arange(n, dtype=np.int64) if sources is None else np.asarray(sources, np.int64)
    s_sorted = sr[np.argsort(npv[sr], kind="stable")]
    # sources that can pair with the cheapest targets (b_x = 2: 3 rows)
    bs = npv[s_sorted]
    cs = src_cols(bs, src_ie)
    n_src = int(np.searchsorted(cs, thr // 3, side="right"))   # cs ascends with bs
    n_src = int(np.searchsorted(bs[:n_src], ONEHOT_MAX_LEVELS, side="right"))
    if n_src == 0:
        return None
    s_use = s_sorted[:n_src]
    s_g, s_col, n_cg = _groups(src_cols(npv[s_use], src_ie))
    cg_max = np.zeros(n_cg, np.int64)
    np.maximum.at(cg_max, s_g, src_cols(npv[s_use], src_ie))
    cg_end = np.searchsorted(s_g, np.arange(n_cg), side="right")   # sources before group end
    # targets: 2 <= b_x, b_x^2 - 1 <= GROUP, (b_x^2 - 1) * cols(b_min) <= thr
    cmin = int(cs[0])
    t_ok = tg[(npv[tg] >= 2) & (npv[tg] ** 2 - 1 <= GROUP) &
              ((npv[tg] ** 2 - 1) * max(cmin, 1) <= thr)]
    t_use = t_ok[np.argsort(npv[t_ok], kind="stable")]
    if len(t_use) == 0:
        return None
    t_g, t_row, n_rg = _groups(npv[t_use] ** 2 - 1)
    rg_max = np.zeros(n_rg, np.int64)
    np.maximum.at(rg_max, t_g, npv[t_use])
    # column groups of each row group: the leading g with (BX^2 - 1) CY <= thr
    # (CY: the group's largest column count)
    ncg = np.array([int(np.searchsorted((rg_max[r] ** 2 - 1) * cg_max, thr, side="right"))
                    for r in range(n_rg)], np.int64)
    # drop row groups with no column group (only trailing ones: ncg falls as r grows)
    keep_rg = int(np.count_nonzero(ncg > 0))
    if keep_rg == 0:
        return None
    t_sel = t_g < keep_rg
    t_use, t_g, t_row = t_use[t_sel], t_g[t_sel], t_row[t_sel]
    n_rg, ncg, rg_max = keep_rg, ncg[:keep_rg], rg_max[:keep_rg]
    src_skip = np.zeros(n, np.int32)
    src_skip[t_use] = cg_end[ncg[t_g] - 1]
    # pairs covered (target != source)
    pos = np.full(n, -1, np.int64)
    pos[s_use] = np.arange(n_src)
    n_pairs = int(src_skip[t_use].astype(np.int64).sum() -
                  np.count_nonzero((pos[t_use] >= 0) & (pos[t_use] < src_skip[t_use])))
    # tiles: row panels (2i, 2i + 1) x column panels (2j, 2j + 1), in bands
    # of BAND row panels crossed with the column panels, so that the tiles
    # an XCD runs at the same time (a contiguous run of this order) share
    # both operands in its L2
    tiles = []
    cells = 0
    panels = list(range(0, n_rg, 2))
    for b0 in range(0, len(panels), BAND):
        band = panels[b0: b0 + BAND]
        for j in range(0, int(ncg[band[0]]), 2):
            c = [j, j + 1 if j + 1 < n_cg else -1]
            for i in band:
                if j >= ncg[i]:
                    continue
                r = [i, i + 1 if i + 1 < n_rg else -1]
                mask = 0
                for wr in range(2):
                    for wc in range(2):
                        if r[wr] >= 0 and c[wc] >= 0 and c[wc] < ncg[r[wr]]:
                            mask |= 1 << (2 * wr + wc)
                            cells += GROUP * GROUP
                if mask:
                    tiles.append([r[0], r[1], c[0], c[1], mask, 0, 0, 0])
    tiles = np.asarray(tiles, np.int32).reshape(-1, 8)
    # operand row descriptions
    rows_a, rows_b = n_rg * GROUP, n_cg * GROUP
    arow = np.zeros((rows_a, 2), np.int32)
    arow[:, 0] = -1
    for x, g, r0 in zip(t_use, t_g, t_row):
        bx = int(npv[x])
        k = np.arange(bx * bx - 1)
        code = np.where(k < (bx - 1) * bx, ((k // bx) << 8) | (k % bx),
                        (255 << 8) | (k - (bx - 1) * bx))
        base = g * GROUP + r0
        arow[base: base + len(k), 0] = x
        arow[base: base + len(k), 1] = code
    brow = np.zeros((rows_b, 2), np.int32)
    brow[:, 0] = -1
    for y, g, c0 in zip(s_use, s_g, s_col):
        cy = int(src_cols(int(npv[y]), src_ie))
        base = g * GROUP + c0
        brow[base: base + cy, 0] = y
        brow[base: base + cy, 1] = np.arange(cy)
    rg_off = np.searchsorted(t_g, np.arange(n_rg + 1)).astype(np.int32)
    cg_off = np.searchsorted(s_g, np.arange(n_cg + 1)).astype(np.int32)
    mvars = np.unique(np.concatenate([t_use, s_use])).astype(np.int32)
    moff = np.full(n, -1, np.int64)
    sizes = npv[mvars] ** 2 + npv[mvars]
    moff[mvars] = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    Kp = -(-int(L) // 128) * 128   # two 64-sample FP4 chunks per K stage
    host = dict(arow=arow, brow=brow, moff=moff, msize=int(sizes.sum()), tiles=tiles,
                rg_off=rg_off, t_var=t_use.astype(np.int32), t_row=t_row.astype(np.int32),
                cg_off=cg_off, s_var=s_use.astype(np.int32), s_col=s_col.astype(np.int32),
                mvars=mvars, Kp=Kp, n_rg=n_rg, n_cg=n_cg, src_ie=int(bool(src_ie)))
    args, keep = _device_args(host, device)
    return OnehotPlan(args=args, src_skip=src_skip, n_pairs=n_pairs, n_tiles=len(tiles),
                      cells=cells, keep=keep, host=host,
                      sources=None if sources is None else np.asarray(sources, np.int64))


def _device_args(host: dict, device):
    """Device arrays (operand buffers, marginals and the plan) and the
    OnehotArgs pointing into them."""
    Kp, rows_a, rows_b = host["Kp"], len(host["arow"]), len(host["brow"])
    h = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)  # noqa: E731
    keep = [torch.empty(rows_a * Kp // 2, dtype=torch.int8, device=device),
            torch.empty(rows_b * Kp // 2, dtype=torch.int8, device=device),
            h(host["arow"]), h(host["brow"]),
            torch.zeros(host["msize"], dtype=torch.int32, device=device),
            h(host["moff"]), h(host["tiles"]), h(host["rg_off"]), h(host["t_var"]),
            h(host["t_row"]), h(host["cg_off"]), h(host["s_var"]), h(host["s_col"]),
            h(host["mvars"])]
    A, B, ar, br, marg, mo, tl, ro, tv, tr, co, sv, sc, mv = keep
    args = _lib.OnehotArgs(
        A=ptr(A), B=ptr(B), Kp=Kp, rows_a=rows_a, rows_b=rows_b, arow=ptr(ar), brow=ptr(br),
        marg=ptr(marg), moff=ptr(mo), tiles=ptr(tl), n_tiles=len(host["tiles"]),
        rg_off=ptr(ro), t_var=ptr(tv), t_row=ptr(tr), cg_off=ptr(co), s_var=ptr(sv),
        s_col=ptr(sc), mvars=ptr(mv), n_mvars=len(host["mvars"]), src_ie=host["src_ie"])
    return args, keep


def replicate(op: OnehotPlan, K: int, m: int, device) -> OnehotPlan:
    """The plan of op (over m variables) for K stacked copies, variable
    k * m + j being copy k of variable j: every copy's pairs stay inside
    the copy (row group g of copy k is group k * n_rg + g, likewise the
    column groups)."""
    H = op.host
    n_rg, n_cg = H["n_rg"], H["n_cg"]
    ks = np.arange(K, dtype=np.int64)

    def rows(r):   # variable ids of copy k shifted by k * m; padding (-1) kept
        out = np.tile(r, (K, 1)).reshape(K, -1, 2)
        v = out[:, :, 0]
        out[:, :, 0] = np.where(v >= 0, v + (ks * m)[:, None], -1)
        return out.reshape(-1, 2).astype(np.int32)

    tiles = np.tile(H["tiles"], (K, 1)).reshape(K, -1, 8)
    for c, step in ((0, n_rg), (1, n_rg), (2, n_cg), (3, n_cg)):
        t = tiles[:, :, c]
        tiles[:, :, c] = np.where(t >= 0, t + (ks * step)[:, None], -1)
    nt0, ns0 = len(H["t_var"]), len(H["s_var"])
    moff = np.tile(H["moff"], K).reshape(K, -1)
    moff = np.where(moff >= 0, moff + (ks * H["msize"])[:, None], -1)
    host = dict(
        arow=rows(H["arow"]), brow=rows(H["brow"]), moff=moff.reshape(-1),
        msize=K * H["msize"], tiles=tiles.reshape(-1, 8).astype(np.int32),
        rg_off=np.concatenate([H["rg_off"][:-1] + k * nt0 for k in range(K)] +
                              [[K * nt0]]).astype(np.int32),
        t_var=(np.tile(H["t_var"], (K, 1)) + (ks * m)[:, None]).reshape(-1).astype(np.int32),
        t_row=np.tile(H["t_row"], K).astype(np.int32),
        cg_off=np.concatenate([H["cg_off"][:-1] + k * ns0 for k in range(K)] +
                              [[K * ns0]]).astype(np.int32),
        s_var=(np.tile(H["s_var"], (K, 1)) + (ks * m)[:, None]).reshape(-1).astype(np.int32),
        s_col=np.tile(H["s_col"], K).astype(np.int32),
        mvars=(np.tile(H["mvars"], (K, 1)) + (ks * m)[:, None]).reshape(-1).astype(np.int32),
        Kp=H["Kp"], n_rg=K * n_rg, n_cg=K * n_cg, src_ie=H["src_ie"])
    args, keep = _device_args(host, device)
    return OnehotPlan(args=args, src_skip=np.tile(op.src_skip, K), n_pairs=K * op.n_pairs,
                      n_tiles=len(host["tiles"]), cells=K * op.cells, keep=keep, host=host)


def covered_pairs(op: OnehotPlan, npv_host: np.ndarray, sources=None) -> np.ndarray:
    """[target, source] pairs the plan counts (host; tests and planning).
    sources defaults to the list the plan was made for."""
    npv = np.asarray(npv_host, np.int64)
    n = len(npv)
    if sources is None:
        sources = op.sources
    sr = np.arange(n, dtype=np.int64) if sources is None else np.asarray(sources, np.int64)
    s_sorted = sr[np.argsort(npv[sr], kind="stable")]
    out = []
    for x in np.nonzero(op.src_skip)[0]:
        s = s_sorted[: op.src_skip[x]]
        s = s[s != x]
        out.append(np.stack([np.full(len(s), x), s], 1))
    return np.concatenate(out).astype(np.int64) if out else np.zeros((0, 2), np.int64)


def run(prep, op: OnehotPlan, te_args: _lib.TeArgs, stream: int):
    """Operands and marginals from prep's bins, then the TE of every pair of
    the plan through te_args' output mode."""
    call("tnx_onehot_pack", C.byref(op.args), ptr(prep.bins), prep.T, prep.ld, prep.tau,
         ptr(prep.npv), stream)
    call("tnx_onehot_te", C.byref(op.args), C.byref(te_args), stream)
