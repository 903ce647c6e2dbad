"""One-hot INT8 MFMA counts for pairs of few levels (csrc/onehot.hip).

Host (no GPU): the one-hot plan and the histogram plan built with its
src_skip cover every ordered pair exactly once; the tiles enumerate exactly
the covered pairs; and the kernel's scheme (indicator rows from the plan's
row codes, their contraction, the marginal subtractions of the epilogue)
reproduces the brute-force joint counts of the oracle.
GPU: counts dumped from the MFMA launch equal the oracle's, and TE matrices
with the one-hot split equal the histogram-only matrices bit for bit
(within 1e-6 of the FP64 oracle).
"""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tenex_amd import _lib, core, onehot


def _bins(npv, T, seed):
    rng = np.random.default_rng(seed)
    out = np.empty((len(npv), T), np.int8)
    for g, b in enumerate(npv):
        row = rng.integers(0, b, size=T)
        row[rng.choice(T, size=b, replace=False)] = np.arange(b)
        out[g] = row
    return out


def _mixed_levels(n, seed):
    rng = np.random.default_rng(seed)
    u = rng.random(n)
    b = np.where(u < 0.75, rng.integers(2, 9, size=n),
                 np.where(u < 0.95, rng.integers(9, 25, size=n), rng.integers(25, 50, size=n)))
    return b.astype(np.int32)


def _hist_pairs(npv, L, skip, tg=None, sr=None):
    info = _lib.PlanInfo()
    npv = np.ascontiguousarray(npv, np.int32)
    tgp = None if tg is None else np.ascontiguousarray(tg, np.int32)
    srp = None if sr is None else np.ascontiguousarray(sr, np.int32)
    sk = np.ascontiguousarray(skip, np.int32)
    args = (_lib.ptr(npv), len(npv), L, _lib.ptr(tgp), 0 if tg is None else len(tg),
            _lib.ptr(srp), 0 if sr is None else len(sr), 0, 0, _lib.ptr(sk))
    _lib.call("tnx_te_plan", *args, None, None, None, 0, C.byref(info))
    n_items = info.n_wave_items + info.n_block_items
    tgo = np.empty(info.n_tgt, np.int32)
    sro = np.empty(info.n_src, np.int32)
    work = np.empty((max(1, n_items), 4), np.int32)
    _lib.call("tnx_te_plan", *args, _lib.ptr(tgo), _lib.ptr(sro), _lib.ptr(work), n_items,
              C.byref(info))
    got = []
    for tb, tc, sb, se in work[:n_items]:
        for ti in range(tb, tb + tc):
            for sj in range(sb, se):
                if tgo[ti] != sro[sj]:
                    got.append((int(tgo[ti]), int(sro[sj])))
    return info, got


def _tile_pairs(op):
    tiles, rg_off, t_var, cg_off, s_var = (op.keep[i].cpu().numpy() for i in (6, 7, 8, 10, 11))
    out = []
    for rg0, rg1, cg0, cg1, mask, *_ in tiles:
        for w in range(4):
            if not (mask >> w) & 1:
                continue
            rg, cg = (rg0, rg1)[w >> 1], (cg0, cg1)[w & 1]
            for x in t_var[rg_off[rg]: rg_off[rg + 1]]:
                for y in s_var[cg_off[cg]: cg_off[cg + 1]]:
                    if x != y:
                        out.append((int(x), int(y)))
    return out


@pytest.mark.parametrize("thr", [24, 120, 400, 2000])
@pytest.mark.parametrize("subset", [False, True])
@pytest.mark.parametrize("src_ie", [False, True])
def test_onehot_split_covers_every_pair_once(thr, subset, src_ie):
    npv = _mixed_levels(400, 3)
    n, L = len(npv), 999
    tg = sr = None
    if subset:
        rng = np.random.default_rng(4)
        tg = np.sort(rng.choice(n, 150, replace=False)).astype(np.int32)
        sr = rng.choice(n, 220, replace=False).astype(np.int32)
    op = onehot.plan(npv, L, "cpu", tg, sr, thr=thr, src_ie=src_ie)
    assert op is not None
    cov = onehot.covered_pairs(op, npv, sr)
    cov_set = {(int(x), int(y)) for x, y in cov}
    assert len(cov_set) == len(cov) == op.n_pairs
    for x, y in cov:  # every one-hot pair satisfies the rule
        assert (npv[x] ** 2 - 1) * onehot.src_cols(npv[y], src_ie) <= thr
    assert sorted(_tile_pairs(op)) == sorted(cov_set)
    info, hist = _hist_pairs(npv, L, op.src_skip, tg, sr)
    assert info.n_pairs == len(hist) == len(set(hist))
    tgs = range(n) if tg is None else tg
    srs = range(n) if sr is None else sr
    exp = {(int(x), int(y)) for x in tgs for y in srs if x != y}
    assert not (cov_set & set(hist))
    assert cov_set | set(hist) == exp


def _emulate_counts(op, bins, npv, L, tau):
    """The MFMA kernel's scheme on the host: indicator rows from the plan's
    codes, their products, the epilogue's marginal subtractions.  Returns
    {(x, y): N (b_x, b_x, b_y)}."""
    arow, brow, moff, tiles, rg_off, t_var, t_row, cg_off, s_var, s_col = (
        op.keep[i].cpu().numpy() for i in (2, 3, 5, 6, 7, 8, 9, 10, 11, 12))
    d = bins.astype(np.int64)

    def rows(desc, target):
        R = np.zeros((len(desc), L), np.int64)
        for r, (v, code) in enumerate(desc):
            if v < 0:
                continue
            if target:
                a, b = code >> 8, code & 0xFF
                R[r] = (d[v, :L] == b) if a == 255 else ((d[v, tau:tau + L] == a) & (d[v, :L] == b))
            else:
                R[r] = d[v, :L] == code
        return R

    G = rows(arow, True) @ rows(brow, False).T
    nyc, nab = {}, {}
    for v in range(len(npv)):
        if moff[v] >= 0:
            nyc[v] = np.bincount(d[v, :L], minlength=npv[v])
            nab[v] = np.bincount(d[v, tau:tau + L] * npv[v] + d[v, :L],
                                 minlength=npv[v] ** 2).reshape(npv[v], npv[v])
    src_ie = bool(op.host["src_ie"])
    out = {}
    for rg0, rg1, cg0, cg1, mask, *_ in tiles:
        for w in range(4):
            if not (mask >> w) & 1:
                continue
            rg, cg = (rg0, rg1)[w >> 1], (cg0, cg1)[w & 1]
            for ti in range(rg_off[rg], rg_off[rg + 1]):
                for sj in range(cg_off[cg], cg_off[cg + 1]):
                    x, y = int(t_var[ti]), int(s_var[sj])
                    if x == y:
                        continue
                    bx, by = int(npv[x]), int(npv[y])
                    c1 = max(by - 1, 1) if src_ie else by
                    g = G[rg * 128 + t_row[ti]: rg * 128 + t_row[ti] + bx * bx - 1,
                          cg * 128 + s_col[sj]: cg * 128 + s_col[sj] + c1]
                    if c1 < by:   # the left-out level: row totals less the stored columns
                        tot = np.concatenate([nab[x][: bx - 1].ravel(), nyc[x][: bx - 1]])
                        g = np.concatenate([g, (tot - g.sum(1))[:, None]], 1)
                    N = np.zeros((bx, bx, by), np.int64)
                    for c in range(by):
                        sumM = 0
                        for b in range(bx):
                            M = g[(bx - 1) * bx + b, c] if b < bx - 1 else nyc[y][c] - sumM
                            sumM += M
                            sa = 0
                            for a in range(bx - 1):
                                N[a, b, c] = g[a * bx + b, c]
                                sa += N[a, b, c]
                            N[bx - 1, b, c] = M - sa
                    out[(x, y)] = N
    return out


@pytest.mark.parametrize("tau", [1, 3])
@pytest.mark.parametrize("src_ie", [False, True])
def test_onehot_scheme_reproduces_joint_counts(tau, src_ie):
    npv = _mixed_levels(60, 7)
    npv[:3] = 1          # constant variables: one level, one stored column
    T = 301
    bins = _bins(npv, T, 8)
    L = T - tau
    op = onehot.plan(npv, L, "cpu", thr=150, src_ie=src_ie)
    got = _emulate_counts(op, bins, npv, L, tau)
    assert len(got) == op.n_pairs > 100
    d = bins.astype(np.int32)
    for (x, y), N in got.items():
        assert np.array_equal(N, O.joint_counts(d, npv, x, y, tau=tau)), (x, y)


def test_onehot_plan_declines():
    assert onehot.plan(np.full(10, 70, np.int32), 500, "cpu") is None   # no small variable
    assert onehot.plan(np.full(10, 3, np.int32), 500, "cpu", thr=0) is None
    assert onehot.plan(np.full(10, 3, np.int32), 70000, "cpu") is None  # counts above u16


# ── GPU ──────────────────────────────────────────────────────────────────

def _prep(bins, tau, dev):
    b, npv, _, T = core.upload_bins(bins, dev)
    return core.prepare(b, npv, T, tau)


@pytest.mark.gpu
@pytest.mark.parametrize("tau", [1, 2])
def test_onehot_counts_bit_exact(dev, tau):
    npv = _mixed_levels(300, 11)
    bins = _bins(npv, 1501, 12)
    prep = _prep(bins, tau, dev)
    oh = core.get_onehot_plan(prep)
    assert oh is not None and oh.n_pairs > 1000
    plan = core.get_plan(prep, src_skip=oh.src_skip)
    cov = onehot.covered_pairs(oh, prep.npv_host)
    rng = np.random.default_rng(5)
    sel = cov[rng.choice(len(cov), 200, replace=False)]
    npvh = prep.npv_host.astype(np.int64)
    mc = int((npvh[sel[:, 0]] ** 2 * npvh[sel[:, 1]]).max())
    counts, m = core.counts_plan(prep, plan, sel, mc, oh=oh)
    counts = counts.cpu().numpy()
    d = bins.astype(np.int32)
    for r, (t, s) in enumerate(sel):
        exp = O.joint_counts(d, prep.npv_host, int(t), int(s), tau=tau).ravel()
        assert np.array_equal(counts[r, : exp.size], exp), (t, s)
    ref = O.te_matrix(d, prep.npv_host, tau=tau).T  # internal layout [target, source]
    assert np.abs(m.cpu().numpy() - ref).max() < 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["mixed", "binary", "subset"])
def test_onehot_matrix_equals_histogram_path(dev, case, monkeypatch):
    if case == "binary":
        rng = np.random.default_rng(2)
        bins = (rng.random((257, 777)) < 0.4).astype(np.int8)
        bins[:, :2] = [0, 1]
    else:
        npv = _mixed_levels(350, 21)
        bins = _bins(npv, 1234, 22)
    prep = _prep(bins, 1, dev)
    tg = sr = None
    if case == "subset":
        rng = np.random.default_rng(3)
        tg = np.sort(rng.choice(350, 120, replace=False))
        sr = rng.choice(350, 200, replace=False)
    assert core.get_onehot_plan(prep, tg, sr) is not None
    m1 = core.te_matrix_internal(prep, tg, sr).cpu().numpy()
    monkeypatch.setenv("TNX_ONEHOT", "0")
    m0 = core.te_matrix_internal(prep, tg, sr).cpu().numpy()
    assert np.array_equal(m1, m0)
    ref = O.te_matrix(bins.astype(np.int32), prep.npv_host).T  # internal layout
    if tg is None:
        assert np.abs(m1 - ref).max() < 1e-6
    else:
        mask = np.zeros_like(m1, dtype=bool)
        mask[np.ix_(tg, sr)] = True
        np.fill_diagonal(mask, False)
        assert np.abs(m1[mask] - ref[mask]).max() < 1e-6
        assert np.all(m1[~mask] == 0)


def test_onehot_replicate_stays_inside_copies():
    npv = _mixed_levels(70, 31)
    m, K, T, tau = len(npv), 3, 257, 1
    op = onehot.plan(npv, T - tau, "cpu", thr=300)
    opk = onehot.replicate(op, K, m, "cpu")
    cov = onehot.covered_pairs(op, npv)
    exp = {(int(x) + k * m, int(y) + k * m) for k in range(K) for x, y in cov}
    got = _tile_pairs(opk)
    assert len(got) == len(set(got)) == opk.n_pairs and set(got) == exp
    # the stacked copies' counts (different data per copy) through the scheme
    npvK = np.tile(npv, K)
    binsK = np.concatenate([_bins(npv, T, 40 + k) for k in range(K)])
    counts = _emulate_counts(opk, binsK, npvK, T - tau, tau)
    d = binsK.astype(np.int32)
    for (x, y), N in counts.items():
        assert x // m == y // m
        assert np.array_equal(N, O.joint_counts(d, npvK, x, y, tau=tau)), (x, y)
