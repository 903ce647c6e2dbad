"""Host-side logic without a GPU: planner (C-ABI, host-only), registry,
device contract, inference utilities."""
import ctypes as C

import numpy as np
import pytest
import torch

import tenex_amd as tnx
from oracle import oracle as O
from tenex_amd import _lib, core
from tenex_amd.inference import (bh_fdr_gpu, build_pairs, fdr_threshold_cpu, make_grn,
                                 threshold_pairs)
from tenex_amd.kernels import auto_select, get_kernel


def _plan_host(npv, L, tg=None, sr=None, cap_wave=0, cap_block=0, skip=None):
    npv = np.ascontiguousarray(npv, np.int32)
    info = _lib.PlanInfo()
    tgp = None if tg is None else np.ascontiguousarray(tg, np.int32)
    srp = None if sr is None else np.ascontiguousarray(sr, np.int32)
    args = (_lib.ptr(npv), len(npv), L, _lib.ptr(tgp), 0 if tg is None else len(tg),
            _lib.ptr(srp), 0 if sr is None else len(sr), cap_wave, cap_block,
            _lib.ptr(skip))
    _lib.call("tnx_te_plan", *args, None, None, None, 0, C.byref(info))
    n_items = info.n_wave_items + info.n_block_items
    tgo = np.empty(info.n_tgt, np.int32)
    sro = np.empty(info.n_src, np.int32)
    work = np.empty((max(1, n_items), 4), np.int32)
    _lib.call("tnx_te_plan", *args, _lib.ptr(tgo), _lib.ptr(sro), _lib.ptr(work), n_items,
              C.byref(info))
    return info, tgo, sro, work[:n_items]


def _expand(info, tgo, sro, work):
    wave, block = [], []
    for k, (tb, tc, sb, se) in enumerate(work):
        dst = wave if k < info.n_wave_items else block
        for ti in range(tb, tb + tc):
            for sj in range(sb, se):
                if tgo[ti] != sro[sj]:
                    dst.append((int(tgo[ti]), int(sro[sj])))
    return wave, block


@pytest.mark.parametrize("cap_wave", [0, 64, 512, 2048])
def test_planner_covers_every_pair_once(cap_wave):
    npv = synthetic_levels(300)
    info, tgo, sro, work = _plan_host(npv, 999, cap_wave=cap_wave)
    wave, block = _expand(info, tgo, sro, work)
    allp = wave + block
    assert len(allp) == len(set(allp)) == 300 * 299 == info.n_pairs
    cap = info.cap_wave
    for x, y in wave:
        assert npv[x] ** 2 * npv[y] <= cap
    for x, y in block:
        assert npv[x] ** 2 * npv[y] > cap
        assert npv[x] * npv[y] <= info.max_block_cols
    # sources ascending by levels, targets descending
    assert np.all(np.diff(npv[sro]) >= 0) and np.all(np.diff(npv[tgo]) <= 0)


def test_planner_sources_and_targets():
    npv = synthetic_levels(50)
    src = np.array([3, 9, 20, 41], np.int32)
    info, tgo, sro, work = _plan_host(npv, 500, sr=src)
    wave, block = _expand(info, tgo, sro, work)
    exp = {(t, s) for t in range(50) for s in src if s != t}
    assert set(wave + block) == exp and len(wave + block) == len(exp)
    tg = np.array([0, 1, 2], np.int32)
    info, tgo, sro, work = _plan_host(npv, 500, tg=tg)
    wave, block = _expand(info, tgo, sro, work)
    assert {p[0] for p in wave + block} == {0, 1, 2}


def test_planner_small_plans_split_items():
    """Small plans get short items (enough to fill the chip); large ones keep
    up to 256 pairs per wave item and 16 per block item."""
    npv = synthetic_levels(300)
    small, _, _, w_small = _plan_host(npv, 999, tg=np.arange(8, dtype=np.int32),
                                      sr=np.arange(100, 160, dtype=np.int32))
    pairs = lambda w: (w[:, 1] * (w[:, 3] - w[:, 2]))
    assert pairs(w_small[:small.n_wave_items]).max() <= 4
    big, _, _, w_big = _plan_host(npv, 999)
    assert pairs(w_big[:big.n_wave_items]).max() > 4


def test_plan_pairs_and_copies():
    """plan_pairs enumerates the items' pairs; replicate_plan's K copies of
    an m-variable plan cover the same pairs offset by k * m."""
    npv = synthetic_levels(40)
    src = np.array([1, 5, 9, 30], np.int32)
    plan = core.make_plan(npv, 500, "cpu", sources=src, cap_wave=256)
    assert plan.info.n_block_items > 0
    x, y = core.plan_pairs(plan)
    exp = {(t, s) for t in range(40) for s in src if s != t}
    assert set(zip(x.tolist(), y.tolist())) == exp and len(x) == len(exp)
    rep = core.replicate_plan(plan, 3, 40)
    xr, yr = core.plan_pairs(rep)
    assert rep.info.n_wave_items == 3 * plan.info.n_wave_items
    assert rep.info.n_block_items == 3 * plan.info.n_block_items
    got = set(zip(xr.tolist(), yr.tolist()))
    assert got == {(t + 40 * k, s + 40 * k) for k in range(3) for t, s in exp}
    # items of the copies stay wave-first, then block
    w = rep.work_host
    nw = rep.info.n_wave_items
    cells = lambda it: [int(npv[rep.tgt_host[t] % 40]) ** 2 * int(npv[rep.src_host[s] % 40])
                        for t in range(it[0], it[0] + it[1]) for s in range(it[2], it[3])]
    assert all(max(cells(it)) <= 256 for it in w[:nw])
    assert all(min(cells(it)) > 256 for it in w[nw:])
    pl = core.te_pairlist_plan(npv, np.array([[3, 7], [3, 1], [12, 30], [0, 5]]), "cpu")
    xp, yp = core.plan_pairs(pl)
    assert set(zip(xp.tolist(), yp.tolist())) == {(3, 7), (3, 1), (12, 30), (0, 5)}


def test_planner_rejects_bad_levels():
    with pytest.raises(ValueError):
        _plan_host(np.array([2, 0, 3], np.int32), 10)
    with pytest.raises(ValueError):
        _plan_host(np.array([2, 200], np.int32), 10)


def synthetic_levels(n):
    from tenex_amd import synthetic
    return synthetic.mixed_cardinality_levels(n, seed=n)


# ── registry and device contract (tests/test_cpu_fallback.py:12-16,
#    tests/test_device_contract.py:22-60 of the reference) ──────────────────

def test_registry_names_and_lookup():
    names = [k.name for k in tnx.registered_kernels()]
    assert names == ["GEMM-B2", "Full-SMEM", "Adaptive-SMEM"]
    assert get_kernel("full-smem").name == "Full-SMEM"
    with pytest.raises(KeyError):
        get_kernel("scatter_add")


def test_auto_select_priorities():
    lds = 160 * 1024
    from tenex_amd.kernels import compute_smem_bytes
    sel = lambda b, src=False: auto_select(b, True, lds, compute_smem_bytes(b, 100, 1, 1), None,
                                           src).name
    assert sel(2) == "GEMM-B2"
    assert sel(2, src=True) == "Full-SMEM"
    assert sel(8) == "Full-SMEM"
    assert sel(34) == "Full-SMEM"   # (34^3 + 34^2) * 4 B fits 160 KiB
    assert sel(35) == "Adaptive-SMEM"
    assert sel(127) == "Adaptive-SMEM"
    with pytest.raises(RuntimeError):
        auto_select(8, False, 0, 0, None, False)


def test_available_methods_match_reference():
    assert tnx.available_methods() == ["clr", "fdr", "nd", "point", "surrogate_test", "trace"]


def test_engine_validation():
    x = np.zeros((3, 10), np.float32)
    names = np.array(["a", "b", "c"])
    with pytest.raises(ValueError):
        tnx.TransferEntropyEngine(x[0], names)
    with pytest.raises(ValueError):
        tnx.TransferEntropyEngine(np.full((3, 10), np.nan, np.float32), names)
    with pytest.raises(ValueError):
        tnx.TransferEntropyEngine(x, np.array(["a", "a", "c"]))
    with pytest.raises(ValueError):
        tnx.TransferEntropyEngine(x, names, sources=["zz"])
    with pytest.raises(ValueError):
        tnx.TransferEntropyEngine(x, names, sources=["a", "a"])


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the no-GPU contract")
def test_device_contract_without_gpu():
    x = np.random.default_rng(0).random((4, 30)).astype(np.float32)
    eng = tnx.TransferEntropyEngine(x, np.array(list("abcd")))
    with pytest.raises(RuntimeError):
        eng.compute(accelerator="gpu")
    with pytest.raises(RuntimeError):   # GPU-only build: no silent CPU path
        eng.compute(accelerator="cpu")
    with pytest.raises(RuntimeError):
        eng.compute(accelerator="auto")
    with pytest.raises(ValueError):
        eng.compute(accelerator="tpu")
    with pytest.raises(RuntimeError):
        core.require_gpu()


def test_tau_validation_before_device():
    x = np.random.default_rng(0).random((4, 30)).astype(np.float32)
    eng = tnx.TransferEntropyEngine(x, np.array(list("abcd")))
    with pytest.raises(ValueError):
        eng._resolve_accelerator_and_devices("bogus")


# ── inference utilities ───────────────────────────────────────────────────

def test_build_pairs_orders():
    names = np.array(list("abcd"))
    p = build_pairs(names)
    assert p.shape == (12, 2) and not np.any(p[:, 0] == p[:, 1])
    assert np.array_equal(p[:3], [[0, 1], [0, 2], [0, 3]])
    ps = build_pairs(names, ["c", "a"])
    assert set(map(tuple, ps)) == {(0, 1), (0, 2), (0, 3), (2, 0), (2, 1), (2, 3)}


def test_fdr_threshold_cpu_and_topk():
    rng = np.random.default_rng(1)
    te = rng.random(1000).astype(np.float32) * 0.01
    te[:5] = 1.0
    pairs = np.stack([np.arange(1000), np.arange(1000)], 1)
    t, p = fdr_threshold_cpu(te, pairs, 0.05)
    assert set(p[:, 0]) == set(range(5))
    t, p = threshold_pairs(te, pairs, torch.device("cpu"), links=3)
    assert len(t) == 3 and np.all(t == 1.0)
    tt = torch.from_numpy(te)
    t2, p2 = threshold_pairs(tt, torch.from_numpy(pairs), torch.device("cpu"), links=3)
    assert len(t2) == 3


def test_netweaver_fdr_on_cpu_matrix():
    rng = np.random.default_rng(2)
    M = (rng.random((20, 20)) * 0.01).astype(np.float32)
    M[3, 7] = 1.0
    np.fill_diagonal(M, 0)
    names = np.array([f"g{i}" for i in range(20)])
    grn, trimmed = tnx.NetWeaver((M, names), fdr=0.05).infer("fdr", device="cpu")
    assert (3, 7) in set(map(tuple, grn.pairs.tolist()))
    assert len(trimmed) <= len(grn)
    assert grn.to_sif().shape == (len(grn), 3)


def test_surrogate_decision_restatement():
    from tenex_amd.inference.surrogate_test import surrogate_decision
    rng = np.random.default_rng(3)
    n, ns = 15, 20
    obs = (rng.random((n, n)) * 0.05).astype(np.float32)
    np.fill_diagonal(obs, 0)
    tk = rng.random((n, n, ns)) * 0.02
    s, s2 = tk.sum(-1), (tk ** 2).sum(-1)
    cg = (tk.astype(np.float32) >= obs[..., None]).sum(-1).astype(np.int32)
    mask = ~np.eye(n, dtype=bool)
    for pm in ("parametric", "mc"):
        mean, std, eff, p, sig = surrogate_decision(obs, s, s2, cg, ns, pm, 0.05, mask,
                                                    torch.device("cpu"))
        ref = O.surrogate_decision(obs, s, s2, cg, ns, p_method=pm, fdr=0.05)
        assert np.array_equal(mean, ref["mean"]) and np.array_equal(std, ref["std"])
        assert np.array_equal(p, ref["p"]) and np.array_equal(sig, ref["sig"])


def test_multi_lag_combine():
    from tenex_amd.inference.surrogate_test import SurrogateTestResult
    from tenex_amd.pipeline import combine_surrogate_multi_lag
    rng = np.random.default_rng(4)
    n = 6
    res = []
    for k in range(3):
        eff = rng.random((n, n)).astype(np.float32) - 0.3
        p = rng.random((n, n)).astype(np.float32) * 0.01
        np.fill_diagonal(p, 1.0)
        grn = make_grn(np.array(list("abcdef")), np.empty((0, 2), np.int64), np.empty(0))
        res.append(SurrogateTestResult(eff + 0.3, np.full((n, n), 0.3, np.float32),
                                       np.ones((n, n), np.float32), eff, p, grn, 10, "block",
                                       5, "parametric", 0.05))
    for comb in ("max", "adaptive_select", "weighted_sum"):
        r = combine_surrogate_multi_lag(res, [1, 2, 3], comb, np.array(list("abcdef")),
                                        torch.device("cpu"))
        assert r.p_values.shape == (n, n) and np.all(np.diag(r.p_values) == 1)
        pv = np.stack([x.p_values for x in res])
        exp = np.minimum(1.0, pv.min(0) * 3).astype(np.float32)
        np.fill_diagonal(exp, 1)
        assert np.array_equal(r.p_values, exp)


def test_bh_reject_matches_adjusted_pvalues():
    """bh_reject_gpu(p, a) is bh_fdr_gpu(p, a) < a (same float32 arithmetic),
    ties, values at the threshold and p = 1 included."""
    from tenex_amd.inference import bh_reject_gpu
    rng = np.random.default_rng(7)
    for m in (1, 2, 17, 1000, 40000):
        p = rng.random(m).astype(np.float32) ** 3
        p[rng.random(m) < 0.2] = np.float32(1.0)
        if m > 10:
            p[:5] = p[5]                      # ties
            p[6] = np.float32(0.05 * 3 / m)   # near the step-up line
        t = torch.from_numpy(p)
        for a in (0.01, 0.05, 0.2):
            assert torch.equal(bh_reject_gpu(t, a), bh_fdr_gpu(t, a) < a), (m, a)
    assert bh_reject_gpu(torch.zeros(0), 0.05).numel() == 0


def test_trim_dpi_known_answers_and_cpu_path():
    """DPI trim: a -> c survives only when A[a,c] >= min(A[a,b], A[b,c]); the
    oracle (fdr.py:61-133 restated) and the CPU device path agree on random
    sparse networks with negative weights, small row blocks (whole blocks
    kept when no edge reaches a mediator) and a nonzero threshold."""
    from tenex_amd.inference.fdr import trim_dpi
    names = np.array([f"g{i}" for i in range(4)])
    A = np.zeros((4, 4), np.float32)
    A[0, 1], A[1, 2], A[0, 2] = 0.5, 0.4, 0.3
    pairs, vals = O.trim_dpi(A)
    assert {tuple(p) for p in pairs.tolist()} == {(0, 1), (1, 2)}
    A[0, 2] = 0.45
    pairs, vals = O.trim_dpi(A)
    assert {tuple(p) for p in pairs.tolist()} == {(0, 1), (1, 2), (0, 2)}
    pairs, vals = O.trim_dpi(A, trim_threshold=0.1)   # 0.45 < 0.4 + 0.1
    assert {tuple(p) for p in pairs.tolist()} == {(0, 1), (1, 2)}
    assert O.trim_dpi(-np.abs(A)) is None                # no mediator: input unchanged
    rng = np.random.default_rng(11)
    for n, dens, bs, thr in [(40, 0.1, 512, 0.0), (90, 0.05, 16, 0.0), (70, 0.2, 8, 0.05)]:
        A = np.where(rng.random((n, n)) < dens, rng.random((n, n)), 0).astype(np.float32)
        A[rng.random((n, n)) < 0.02] *= -1
        np.fill_diagonal(A, 0)
        A[: n // 4] = np.where(A[: n // 4] > 0, 0, A[: n // 4])   # rows without mediators
        p = np.argwhere(A != 0)
        ref = O.trim_dpi(A, batch_size=bs, trim_threshold=thr)
        got = trim_dpi(A, np.array([f"g{i}" for i in range(n)]), p, A[p[:, 0], p[:, 1]],
                       torch.device("cpu"), batch_size=bs, trim_threshold=thr)
        assert np.array_equal(got.pairs, ref[0]) and np.array_equal(got.te, ref[1]), (n, bs)


def test_tested_pairs_match_mask_nonzero():
    """The learned null's host-built pair list equals nonzero over the
    rows x cols mask off the diagonal (row-major), and its sources and
    targets equal the unique ones of that list, overlaps and singletons
    included."""
    from tenex_amd.inference.learned_null import _tested_pairs
    rng = np.random.default_rng(0)
    n = 40
    cases = [(rng.choice(n, 7, replace=False), rng.choice(n, 12, replace=False)),
             (np.arange(n), np.arange(n)), (np.array([5]), np.array([5, 9])),
             (np.array([5, 9]), np.array([5])), (np.array([3]), np.array([3])),
             (np.array([1, 2, 3]), np.array([2, 3, 4, 30]))]
    for rows, cols in cases:
        m = np.zeros((n, n), bool)
        m[np.ix_(rows, cols)] = True
        np.fill_diagonal(m, False)
        exp = np.argwhere(m)
        pub, srcs, tgts = _tested_pairs(rows, cols)
        assert np.array_equal(pub, exp), (rows, cols)
        pub_t, _, _ = _tested_pairs(rows, cols, device="cpu")
        assert np.array_equal(pub_t.numpy(), exp), (rows, cols)
        assert np.array_equal(srcs, np.unique(exp[:, 0]))
        assert np.array_equal(tgts, np.unique(exp[:, 1]))


def test_upload_bins_rejects_out_of_range_levels():
    """Bins outside 0 <= bin <= 126 would reach the kernels as wrapped int8
    bytes: upload_bins refuses them (checked on the host tensor)."""
    ok = np.array([[0, 1, 2], [3, 0, 1]], np.int32)
    bins, npv, b_max, T = core.upload_bins(ok, "cpu")
    assert b_max == 4 and T == 3 and list(npv.numpy()) == [3, 4]
    for bad in (np.array([[0, -1, 2]], np.int32), np.array([[0, 127, 3]], np.int32),
                np.array([[0, 200, 3]], np.int64)):
        with pytest.raises(ValueError):
            core.upload_bins(bad, "cpu")
    with pytest.raises(ValueError):
        core.upload_bins(np.zeros((2, 3), np.float32), "cpu")


def test_learned_null_requires_parametric_p_values():
    from tenex_amd.inference.surrogate_test import SurrogateTestMethod
    names = np.array(list("abc"))
    kw = dict(bin_data=np.zeros((3, 40), np.int32), n_per_var=np.ones(3, np.int32), b_max=1,
              kernel="Adaptive-SMEM", learned="mlp", p_method="mc")
    with pytest.raises(ValueError, match="parametric"):
        SurrogateTestMethod().infer(np.zeros((3, 3), np.float32), names, None, **kw)


def test_upload_bins_rejects_out_of_range_levels():
    """Bins outside 0 <= bin <= 126 cannot be int8 histogram keys: upload_bins
    raises instead of passing wrapped bytes to the kernels (checked here with
    the host as the device)."""
    from tenex_amd import core
    ok = np.array([[0, 1, 2, 1], [3, 0, 0, 1]], np.int32)
    bins, npv, b_max, T = core.upload_bins(ok, "cpu")
    assert T == 4 and b_max == 4 and npv.tolist() == [3, 4]
    assert np.array_equal(bins[:, :4].numpy(), ok)
    for bad in (np.array([[0, -1, 2]], np.int32), np.array([[0, 127, 2]], np.int32),
                np.array([[0, 200, 1]], np.int64)):
        with pytest.raises(ValueError, match="int8 bins cannot hold"):
            core.upload_bins(bad, "cpu")
    with pytest.raises(ValueError, match="integer"):
        core.upload_bins(np.zeros((2, 3), np.float32), "cpu")
