"""GPU parity of the fused surrogate kernel and the surrogate test."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tenex_amd import core, synthetic
from tenex_amd.inference.surrogate_test import surrogate_accumulate

pytestmark = pytest.mark.gpu
TOL = 1e-6


def _small(n=30, T=400, seed=0):
    x = synthetic.zero_inflated_lognormal(n, T, seed=seed)
    d, nb, npv, bm = O.discretize(x)
    return d, npv


def _prep(d, npv, dev, tau=1):
    bins, _, _, T = core.upload_bins(d, dev)
    return core.prepare(bins, torch.from_numpy(npv).to(dev), T, tau)


@pytest.mark.parametrize("bl,method,tau", [(20, "block", 1), (33, "block", 1), (1, "random", 1),
                                           (25, "block", 2)])
def test_fused_te_k_vs_oracle(dev, bl, method, tau):
    d, npv = _small()
    n, T = d.shape
    prep = _prep(d, npv, dev, tau)
    obs_int = core.te_matrix_internal(prep).cpu().numpy()
    obs_pub = obs_int.T.copy()
    n_surr = 6
    perm, eff_bl, nb = core.block_perm_device(n, T, method, bl, 0, n_surr, 42, dev, dev)
    plan = core.surrogate_plan(prep)
    s = torch.zeros((n, n), dtype=torch.float64, device=dev)
    s2 = torch.zeros_like(s)
    cg = torch.zeros((n, n), dtype=torch.int32, device=dev)
    tk = torch.zeros((n * n, n_surr), dtype=torch.float32, device=dev)
    core.surrogate_run(prep, plan, perm, eff_bl, torch.from_numpy(obs_pub).to(dev), s, s2, cg, tk)
    pairs = O.pairfree_pairs(n)
    obs_pairs = obs_int[pairs[:, 0], pairs[:, 1]]
    os_, os2, ocg, otk = O.surrogate_stats(d, npv, perm.cpu().numpy(), eff_bl, pairs, obs_pairs,
                                           tau=tau, return_te_k=True)
    slot = pairs[:, 1] * n + pairs[:, 0]
    tk = tk.cpu().numpy()[slot]
    assert np.abs(tk - otk).max() < TOL
    sp = s.cpu().numpy().ravel()[slot]
    assert np.abs(sp - os_).max() < n_surr * TOL
    assert np.abs(s2.cpu().numpy().ravel()[slot] - os2).max() < n_surr * 2 * TOL
    # counts: identical except where a te_k sits within rounding of te_obs
    cgp = cg.cpu().numpy().ravel()[slot]
    near = np.any(np.abs(otk - obs_pairs[:, None]) < 1e-5, axis=1)
    assert np.array_equal(cgp[~near], ocg[~near])


def test_pairlist_plan_single_target_items(dev):
    """Explicit pair lists give one target per work item: the kernel's
    one-target mode (waves split the sources) against the oracle."""
    d, npv = _small(n=25, T=350, seed=9)
    n, T = d.shape
    prep = _prep(d, npv, dev)
    obs_int = core.te_matrix_internal(prep).cpu().numpy()
    rng = np.random.default_rng(1)
    allp = O.pairfree_pairs(n)
    pairs = allp[rng.choice(len(allp), 150, replace=False)]  # [target, source]
    cap = core.surrogate_cap(prep.npv_host, pairs[:, 0], pairs[:, 1])
    plan = core.pairlist_plan(prep.npv_host, pairs, dev, cap)
    assert np.all(plan.work_host[:, 1] == 1)
    n_surr = 5
    perm, eff_bl, nb = core.block_perm_device(n, T, "block", 30, 0, n_surr, 7, dev, dev)
    s = torch.zeros((n, n), dtype=torch.float64, device=dev)
    s2 = torch.zeros_like(s)
    cg = torch.zeros((n, n), dtype=torch.int32, device=dev)
    tk = torch.zeros((n * n, n_surr), dtype=torch.float32, device=dev)
    obs_pub = torch.from_numpy(obs_int.T.copy()).to(dev)
    core.surrogate_run(prep, plan, perm, eff_bl, obs_pub, s, s2, cg, tk)
    obs_pairs = obs_int[pairs[:, 0], pairs[:, 1]]
    os_, os2, ocg, otk = O.surrogate_stats(d, npv, perm.cpu().numpy(), eff_bl, pairs, obs_pairs,
                                           return_te_k=True)
    slot = pairs[:, 1] * n + pairs[:, 0]
    assert np.abs(tk.cpu().numpy()[slot] - otk).max() < TOL
    assert np.abs(s.cpu().numpy().ravel()[slot] - os_).max() < n_surr * TOL
    # slots outside the list stay untouched
    mask = np.ones(n * n, bool)
    mask[slot] = False
    assert np.all(s.cpu().numpy().ravel()[mask] == 0)


def test_block_shuffle_bit_exact(dev):
    """tnx_block_shuffle against the shuffle schedule restated in numpy:
    blocks reordered by perm, tail and padding unchanged."""
    rng = np.random.default_rng(4)
    for n, T, bl in [(7, 103, 10), (5, 64, 16), (9, 250, 1), (4, 99, 99), (6, 333, 7)]:
        raw = rng.integers(0, 9, size=(n, T)).astype(np.int32)
        bins, _, _, T_ = core.upload_bins(raw, dev)
        nb = T // bl
        perm = torch.stack([torch.randperm(nb, generator=torch.Generator().manual_seed(i))
                            for i in range(n)]).to(torch.int32).to(dev)
        out = torch.full_like(bins, -1)
        core.block_shuffle(bins, T, perm, bl, out)
        got = out.cpu().numpy()[:, :T]
        p = perm.cpu().numpy()
        exp = raw.copy()
        head = nb * bl
        for v in range(n):
            exp[v, :head] = raw[v, :head].reshape(nb, bl)[p[v]].ravel()
        assert np.array_equal(got, exp.astype(np.int8)), (n, T, bl)
        assert np.all(out.cpu().numpy()[:, T:] == 0)
        # K stacked copies: copy k of row v uses perm row k * n + v
        K = 3
        permK = torch.stack([torch.randperm(nb, generator=torch.Generator().manual_seed(100 + i))
                             for i in range(K * n)]).to(torch.int32).to(dev)
        outK = torch.full((K * n, bins.shape[1]), -1, dtype=torch.int8, device=dev)
        core.block_shuffle(bins, T, permK, bl, outK, copies=K)
        for k in range(K):
            one = torch.full_like(bins, -1)
            core.block_shuffle(bins, T, permK[k * n:(k + 1) * n], bl, one)
            assert torch.equal(outK[k * n:(k + 1) * n], one), (n, T, bl, k)


@pytest.mark.parametrize("bl,tau", [(20, 1), (33, 2)])
def test_materialized_stats_vs_oracle(dev, bl, tau):
    """Default path: shuffled copies through the TE kernels' statistics
    epilogue, against the oracle; a pair list with block-kernel pairs too."""
    d, npv = _small(n=30, T=420, seed=2)
    n, T = d.shape
    prep = _prep(d, npv, dev, tau)
    obs_int = core.te_matrix_internal(prep).cpu().numpy()
    obs_pub = torch.from_numpy(obs_int.T.copy()).to(dev)
    n_surr = 5
    s, s2, cg = surrogate_accumulate(prep, obs_pub, n_surr, "block", bl, 11, dev)
    perm, eff_bl, nb = core.block_perm_device(n, T, "block", bl, 0, n_surr, 11, dev, dev)
    pairs = O.pairfree_pairs(n)
    obs_pairs = obs_int[pairs[:, 0], pairs[:, 1]]
    os_, os2, ocg = O.surrogate_stats(d, npv, perm.cpu().numpy(), eff_bl, pairs, obs_pairs,
                                      tau=tau)
    slot = pairs[:, 1] * n + pairs[:, 0]
    assert np.abs(s.cpu().numpy().ravel()[slot] - os_).max() < n_surr * TOL
    assert np.abs(s2.cpu().numpy().ravel()[slot] - os2).max() < n_surr * 2 * TOL
    assert np.all(np.diag(s.cpu().numpy()) == 0)
    # a pair list whose large pairs take the block kernel (tiny wave cap)
    rng = np.random.default_rng(3)
    sub = pairs[rng.choice(len(pairs), 120, replace=False)]
    cells = npv[sub[:, 0]].astype(np.int64) ** 2 * npv[sub[:, 1]]
    plan = core.te_pairlist_plan(prep.npv_host, sub, dev, cap_wave=int(np.median(cells)))
    assert plan.info.n_block_items > 0 and plan.info.n_wave_items > 0
    S = torch.zeros((n, n), dtype=torch.float64, device=dev)
    S2 = torch.zeros_like(S)
    CG = torch.zeros((n, n), dtype=torch.int32, device=dev)
    core.surrogate_stats_materialized(prep, plan, perm, eff_bl, obs_pub, S, S2, CG)
    sl = sub[:, 1] * n + sub[:, 0]
    ref = dict(zip(slot.tolist(), os_.tolist()))
    assert max(abs(S.cpu().numpy().ravel()[k] - ref[k]) for k in sl.tolist()) < n_surr * TOL
    mask = np.ones(n * n, bool)
    mask[sl] = False
    assert np.all(S.cpu().numpy().ravel()[mask] == 0)


def test_staged_batch_equals_per_surrogate_runs(dev):
    """K surrogates in one launch (TNX_OUT_STAGE + tnx_stats_accumulate) and
    K one-surrogate runs (TNX_OUT_STATS) make the same additions in the same
    order: identical bits, wave and block pairs."""
    d, npv = _small(n=26, T=380, seed=9)
    n, T = d.shape
    prep = _prep(d, npv, dev)
    obs = torch.from_numpy(core.te_matrix_internal(prep).cpu().numpy().T.copy()).to(dev)
    perm, eff_bl, nb = core.block_perm_device(n, T, "block", 19, 0, 6, 3, dev, dev)
    cells = npv[:, None].astype(np.int64) ** 2 * npv[None, :]
    plan = core.get_plan(prep, cap_wave=int(np.median(cells)))
    assert plan.info.n_block_items > 0
    z = lambda dt: torch.zeros((n, n), dtype=dt, device=dev)
    a = (z(torch.float64), z(torch.float64), z(torch.int32))
    b = (z(torch.float64), z(torch.float64), z(torch.int32))
    core.surrogate_stats_materialized(prep, plan, perm, eff_bl, obs, *a)
    work = {}
    for k in range(perm.shape[0]):
        core.surrogate_stats_materialized(prep, plan, perm[k:k + 1], eff_bl, obs, *b, work)
    for u, v in zip(a, b):
        assert torch.equal(u, v)
    assert float(a[0].abs().sum()) > 0


def test_subset_stats_equal_full(dev):
    """Targets x sources touching a subset of the variables: only the subset
    is shuffled (the same block orders as rows of the full draw), and every
    pair's statistics equal the full computation's bit for bit."""
    from tenex_amd.inference.learned_null import plain_stats_pairs
    d, npv = _small(n=40, T=450, seed=6)
    n, T = d.shape
    prep = _prep(d, npv, dev)
    obs = torch.from_numpy(core.te_matrix_internal(prep).cpu().numpy().T.copy()).to(dev)
    full = surrogate_accumulate(prep, obs, 7, "block", 25, 5, dev)
    tg = np.array([3, 9, 17, 30, 31], np.int64)
    sr = np.array([0, 9, 22, 39], np.int64)
    part = surrogate_accumulate(prep, obs, 7, "block", 25, 5, dev, targets=tg, sources=sr)
    sel = np.zeros((n, n), bool)
    sel[np.ix_(sr, tg)] = True           # public layout [source, target]
    sel[np.eye(n, dtype=bool)] = False
    for a, b in zip(full, part):
        a, b = a.cpu().numpy(), b.cpu().numpy()
        assert np.array_equal(a[sel], b[sel])
        assert np.all(b[~sel] == 0)
    pairs = np.array([[30, 0], [3, 22], [17, 39], [31, 9]], np.int64)   # [target, source]
    s, s2, cg = plain_stats_pairs(prep, pairs, obs, 7, "block", 25, 5, dev)
    for got, ref in zip((s, s2, cg), full):
        assert np.array_equal(got.cpu().numpy(), ref.cpu().numpy()[pairs[:, 1], pairs[:, 0]])


def test_materialized_matches_fused_and_loop(dev):
    d, npv = _small(n=24, T=500, seed=5)
    prep = _prep(d, npv, dev)
    obs = torch.from_numpy(core.te_matrix_internal(prep).cpu().numpy().T.copy()).to(dev)
    m = surrogate_accumulate(prep, obs, 8, "block", 22, 42, dev)
    f = surrogate_accumulate(prep, obs, 8, "block", 22, 42, dev, fused=True)
    off = ~np.eye(24, dtype=bool)
    assert (np.abs(m[0].cpu().numpy() - f[0].cpu().numpy())[off] / 8).max() < 1e-8
    near = np.abs(m[2].cpu().numpy() - f[2].cpu().numpy())[off]
    assert near.max() <= 1  # float32 rounding of one te_k against observed at most


def test_identity_permutation_gives_observed(dev):
    d, npv = _small(n=20, T=300, seed=3)
    n, T = d.shape
    prep = _prep(d, npv, dev)
    obs = core.te_matrix_internal(prep).cpu().numpy().T.copy()
    perm = torch.zeros((3, n, 1), dtype=torch.int32, device=dev)  # one block of length T
    plan = core.surrogate_plan(prep)
    s = torch.zeros((n, n), dtype=torch.float64, device=dev)
    s2 = torch.zeros_like(s)
    cg = torch.zeros((n, n), dtype=torch.int32, device=dev)
    core.surrogate_run(prep, plan, perm, T, torch.from_numpy(obs).to(dev), s, s2, cg)
    off = ~np.eye(n, dtype=bool)
    assert np.abs(s.cpu().numpy()[off] / 3 - obs[off]).max() < TOL
    assert np.all(cg.cpu().numpy()[off] == 3)


def test_fused_matches_loop(dev):
    d, npv = _small(n=24, T=500, seed=5)
    prep = _prep(d, npv, dev)
    obs = torch.from_numpy(core.te_matrix_internal(prep).cpu().numpy().T.copy()).to(dev)
    a = surrogate_accumulate(prep, obs, 8, "block", 22, 42, dev, fused=True)
    b = surrogate_accumulate(prep, obs, 8, "block", 22, 42, dev, fused=False)
    off = ~np.eye(24, dtype=bool)
    # mean surrogate TE of the two paths (the reference states <= 2e-9 for its
    # loop vs fused kernels); both are within 1e-6 of the FP64 oracle
    assert (np.abs(a[0].cpu().numpy() - b[0].cpu().numpy())[off] / 8).max() < 1e-8
    assert (np.abs(a[1].cpu().numpy() - b[1].cpu().numpy())[off] / 8).max() < 1e-8
    ca, cb = a[2].cpu().numpy(), b[2].cpu().numpy()
    assert np.array_equal(ca[off], cb[off])


def test_surrogate_test_end_to_end(dev):
    import tenex_amd as tnx
    x = synthetic.zero_inflated_lognormal(40, 600, seed=8)
    # plant a lagged dependence 3 -> 7
    x[7, 1:] = x[3, :-1] + 0.1 * np.random.default_rng(1).random(599).astype(np.float32)
    names = np.array([f"g{i}" for i in range(40)])
    res = tnx.TransferEntropyEngine(x, names).compute(accelerator="gpu")
    st = tnx.NetWeaver(res, fdr=0.05).infer(method="surrogate_test", n_surrogates=20,
                                            block_length=25)
    # oracle decision from oracle statistics on the same permutations
    d, nb, npv, bm = O.discretize(x)
    perm, eff_bl, nbk = O.block_perm(40, 600, "block", 25, 0, 20, 42, device=dev)
    pairs = O.pairfree_pairs(40)
    obs = res.matrix  # public layout [src, tgt]
    obs_pairs = obs[pairs[:, 1], pairs[:, 0]]
    s, s2, cg = O.surrogate_stats(d, npv, perm, eff_bl, pairs, obs_pairs)
    S = np.zeros((40, 40)); S2 = np.zeros((40, 40)); CG = np.zeros((40, 40), np.int32)
    S[pairs[:, 1], pairs[:, 0]] = s
    S2[pairs[:, 1], pairs[:, 0]] = s2
    CG[pairs[:, 1], pairs[:, 0]] = cg
    dec = O.surrogate_decision(obs, S, S2, CG, 20, fdr=0.05)
    assert np.abs(st.mean_surrogate_te - dec["mean"]).max() < 1e-6
    assert np.abs(st.std_surrogate_te - dec["std"]).max() < 1e-5
    off = ~np.eye(40, dtype=bool)
    assert np.abs(st.p_values - dec["p"])[off].max() < 1e-3
    got = set(map(tuple, st.grn.pairs.tolist()))
    exp = set(map(tuple, np.argwhere(dec["sig"]).tolist()))
    assert len(got ^ exp) <= 1
    assert (3, 7) in got


def test_pair_stats_batches_straddle_draws(dev, monkeypatch):
    """Pair-list statistics with a few copies per launch, so the
    launches straddle the 32-surrogate block-order draws: every pair's
    statistics still equal the full computation's bit for bit."""
    from tenex_amd.inference.learned_null import plain_stats_pairs
    d, npv = _small(n=30, T=400, seed=8)
    prep = _prep(d, npv, dev)
    obs = torch.from_numpy(core.te_matrix_internal(prep).cpu().numpy().T.copy()).to(dev)
    full = surrogate_accumulate(prep, obs, 40, "block", 20, 11, dev)
    rng = np.random.default_rng(1)
    allp = np.array([[t, s] for t in range(30) for s in range(30) if t != s], np.int64)
    pairs = allp[rng.choice(len(allp), 50, replace=False)]
    m = len(np.unique(pairs))
    monkeypatch.setattr(core, "STAGE_FLOATS", 7 * m * m)   # 7 copies per launch
    s, s2, cg = plain_stats_pairs(prep, pairs, obs, 40, "block", 20, 11, dev)
    for got, ref in zip((s, s2, cg), full):
        assert np.array_equal(got.cpu().numpy(), ref.cpu().numpy()[pairs[:, 1], pairs[:, 0]])
    with pytest.raises(ValueError):
        core.surrogate_stats_pairs(prep, np.array([[3, 3]]), obs[:1, 0], 4, "block", 20, 11, dev)
    with pytest.raises(ValueError):
        core.surrogate_stats_pairs(prep, np.array([[3, 4], [3, 4]]), obs[:2, 0], 4, "block", 20,
                                   11, dev)


def test_decision_pairs_matches_dense(dev):
    """The decision on the tested pairs' sums gives the dense decision's
    matrices bit for bit, untested entries included."""
    from tenex_amd.inference.surrogate_test import surrogate_decision, surrogate_decision_pairs
    n, K = 60, 100
    rng = np.random.default_rng(3)
    obs = (rng.random((n, n)) * 0.05).astype(np.float32)
    obs[rng.random((n, n)) < 0.1] = 0.0
    obs[0, 5], obs[1, 7], obs[2, 9] = 1e-20, -1e-3, 3e-13
    np.fill_diagonal(obs, 0.0)
    mask = rng.random((n, n)) < 0.3
    np.fill_diagonal(mask, False)
    pairs = np.argwhere(mask)
    mean = rng.random(len(pairs)) * 0.03
    sd = rng.random(len(pairs)) * 0.004 + 1e-4
    S = mean * K
    S2 = (sd * sd + mean * mean) * K
    Sd = np.zeros((n, n))
    S2d = np.zeros((n, n))
    Sd[pairs[:, 0], pairs[:, 1]] = S
    S2d[pairs[:, 0], pairs[:, 1]] = S2
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    a = surrogate_decision(t(obs), t(Sd), t(S2d), None, K, "parametric", 0.05, t(mask), dev,
                           host=False)
    b = surrogate_decision_pairs(t(obs), t(pairs.astype(np.int64)), t(S), t(S2), K, 0.05)
    assert int(a[4].sum()) > 0
    for x, y in zip(a, b):
        assert x.dtype == y.dtype and torch.equal(x, y)


def test_uncounted_stats_and_pending_upload(dev, monkeypatch):
    """observed None skips count_ge and leaves the sums unchanged (staged
    batches and one-surrogate runs); the observed matrix uploaded on a side
    stream arrives intact."""
    d, npv = _small(n=28, T=360, seed=12)
    n = d.shape[0]
    prep = _prep(d, npv, dev)
    obs_h = core.te_matrix_internal(prep).cpu().numpy().T.copy()
    up = core.PendingUpload(obs_h, dev)
    obs = up.get()
    assert np.array_equal(obs.cpu().numpy(), obs_h)
    tg = np.array([1, 4, 9, 20], np.int64)
    sr = np.array([0, 4, 13, 27], np.int64)
    a = surrogate_accumulate(prep, obs, 9, "block", 21, 8, dev, targets=tg, sources=sr)
    b = surrogate_accumulate(prep, None, 9, "block", 21, 8, dev, targets=tg, sources=sr)
    c = surrogate_accumulate(prep, None, 9, "block", 21, 8, dev)
    full = surrogate_accumulate(prep, obs, 9, "block", 21, 8, dev)
    assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])
    assert int(b[2].abs().sum()) == 0 and int(a[2].sum()) > 0
    assert torch.equal(c[0], full[0]) and torch.equal(c[1], full[1])
    assert int(c[2].abs().sum()) == 0
    monkeypatch.setattr(core, "STAGE_FLOATS", n * n)     # one surrogate per run (TNX_OUT_STATS)
    d1 = surrogate_accumulate(prep, None, 9, "block", 21, 8, dev)
    assert torch.equal(d1[0], full[0]) and torch.equal(d1[1], full[1])
    assert int(d1[2].abs().sum()) == 0
