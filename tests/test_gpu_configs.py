"""The benchmark workloads at their full sizes, checked against the oracle.

  4,000 x 10,000 mixed cardinality (seed 2): TE of 20,480 stratified pairs
      within 1e-6 of the FP64 oracle; raw joint counts of 256 sampled wave-
      and block-role pairs (single and multi-pass) bit-exact, dumped from
      the production launch of the whole matrix; with the one-hot split,
      counts of 128 MFMA pairs bit-exact and the whole matrix equal to the
      histogram-only one bit for bit.
  The same shape from continuous data (4,000 x 10,000 zero-inflated
      log-normal, seed 2): FSBW-L bins, remap and level counts on the GPU
      equal numpy's bit for bit; TE of 2,048 pairs (12-21 levels, nearly all
      block-role) within 1e-6 of the FP64 oracle and the counts of 64 of
      them bit-exact.
  1,000 x 3,000 surrogate test (100 block shuffles of length 50, seed 42,
      torch CPU generator): per-pair FP64 sums, sums of squares and counts of
      2,000 sampled pairs against the oracle's shuffle-and-count.
  6,000 x 3,000 learned null (200 TF x 2,000 targets): edge-set Jaccard
      >= 0.98 against the plain test; the pairs the learned test measures
      (training grid and re-runs) carry the plain test's statistics bit for
      bit.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tenex_amd import _lib, core, synthetic

pytestmark = pytest.mark.gpu
TOL = 1e-6


def _stratified_pairs(npv, n_sample, seed):
    """Random ordered pairs, 60% / 30% / 10% with b_t^2 b_s <= 2048,
    <= 16384, larger."""
    rng = np.random.default_rng(seed)
    n = len(npv)
    tg = rng.integers(0, n, size=8 * n_sample)
    sr = rng.integers(0, n, size=8 * n_sample)
    keep = tg != sr
    tg, sr = tg[keep], sr[keep]
    cost = npv[tg].astype(np.int64) ** 2 * npv[sr]
    cls = np.digitize(cost, [2048, 16384])
    pick = np.concatenate([np.nonzero(cls == c)[0][: int(n_sample * f)]
                           for c, f in ((0, 0.6), (1, 0.3), (2, 0.1))])
    return np.stack([tg[pick], sr[pick]], axis=1).astype(np.int64)


def test_c2_mixed_cardinality_4000x10000(dev):
    n, T = 4000, 10000
    b = synthetic.mixed_cardinality_bins(n, T, seed=2)
    bins, npv, b_max, T = core.upload_bins(b, dev)
    prep = core.prepare(bins, npv, T, 1)
    npvh = prep.npv_host.astype(np.int64)
    plan = core.get_plan(prep)
    assert core.persistent_ok(prep, plan), "the headline plan runs as one persistent launch"
    d = b.astype(np.int32)
    # counts: 128 wave-role pairs, 128 block-role pairs (incl. several passes)
    pairs = O.pairfree_pairs(n)
    rng = np.random.default_rng(11)
    cand = pairs[rng.choice(len(pairs), 400000, replace=False)]
    cells = npvh[cand[:, 0]] ** 2 * npvh[cand[:, 1]]
    cap_b = core.block_cap(prep, plan)
    wave = cand[cells <= plan.info.cap_wave][:128]
    single = cand[(cells > plan.info.cap_wave) & (cells <= cap_b)][:96]
    multi = cand[cells > cap_b][:32]
    assert len(wave) == 128 and len(single) == 96 and len(multi) == 32
    sel = np.concatenate([wave, single, multi])
    mc = int((npvh[sel[:, 0]] ** 2 * npvh[sel[:, 1]]).max())
    counts, m = core.counts_plan(prep, plan, sel, mc)
    counts = counts.cpu().numpy()
    for r, (t, s) in enumerate(sel):
        exp = O.joint_counts(d, prep.npv_host, int(t), int(s)).ravel()
        assert np.array_equal(counts[r, : exp.size], exp), (t, s, exp.size)
    # TE: 20,480 stratified pairs against the FP64 oracle
    m = m.cpu().numpy()
    smp = _stratified_pairs(prep.npv_host, 20480, seed=2)
    assert len(smp) == 20480
    te64 = O.te_pairs(d, prep.npv_host, smp)
    err = np.abs(m[smp[:, 0], smp[:, 1]].astype(np.float64) - te64)
    assert err.max() < TOL, err.max()
    assert np.all(np.diag(m) == 0)
    # the production split: pairs of few levels on the one-hot MFMA kernel
    # (counts of 128 of them dumped from that launch), the rest in the
    # histogram launch; the matrix equals the histogram-only one bit for bit
    from tenex_amd import onehot
    plan2 = core.get_te_plan(prep)
    oh = plan2.cache["onehot"]
    assert oh.n_pairs > 5_000_000 and plan2.info.n_pairs + oh.n_pairs == n * (n - 1)
    cov = onehot.covered_pairs(oh, prep.npv_host)
    sel2 = cov[np.random.default_rng(12).choice(len(cov), 128, replace=False)]
    mc2 = int((npvh[sel2[:, 0]] ** 2 * npvh[sel2[:, 1]]).max())
    counts2, m2 = core.counts_plan(prep, plan2, sel2, mc2, oh=oh)
    counts2 = counts2.cpu().numpy()
    for r, (t, s) in enumerate(sel2):
        exp = O.joint_counts(d, prep.npv_host, int(t), int(s)).ravel()
        assert np.array_equal(counts2[r, : exp.size], exp), (t, s, exp.size)
    assert np.array_equal(m2.cpu().numpy(), m)


def test_c2_continuous_4000x10000(dev):
    n, T = 4000, 10000
    x = synthetic.zero_inflated_lognormal(n, T, seed=2)
    raw_g, nb_g = core.fsbw_l_gpu(torch.from_numpy(x).to(dev))
    raw_o, nb_o, _, _ = O.fsbw_l_numpy(x)
    bad = np.nonzero((raw_g.cpu().numpy() != raw_o).any(1))[0]
    assert len(bad) == 0, f"{len(bad)} rows differ from numpy, first {bad[:8]}"
    assert np.array_equal(nb_g.cpu().numpy(), nb_o)
    bins, npv, b_max = core.remap_gpu(raw_g)
    d, npv_o, bm_o = O.remap(raw_o)
    assert np.array_equal(bins[:, :T].cpu().numpy().astype(np.int32), d)
    assert np.array_equal(npv.cpu().numpy(), npv_o) and b_max == bm_o
    prep = core.prepare(bins, npv, T, 1)
    npvh = prep.npv_host.astype(np.int64)
    assert npvh.min() >= 9        # nothing for the one-hot path: histogram kernels only
    m = core.te_matrix_internal(prep).cpu().numpy()
    rng = np.random.default_rng(4)
    pairs = rng.integers(0, n, size=(3000, 2))
    pairs = pairs[pairs[:, 0] != pairs[:, 1]][:2048]
    te64 = O.te_pairs(d, npv_o, pairs)
    err = np.abs(m[pairs[:, 0], pairs[:, 1]].astype(np.float64) - te64)
    assert err.max() < TOL, err.max()
    plan = core.get_te_plan(prep)
    sel = pairs[:64]
    mc = int((npvh[sel[:, 0]] ** 2 * npvh[sel[:, 1]]).max())
    counts, m2 = core.counts_plan(prep, plan, sel, mc)
    counts = counts.cpu().numpy()
    for r, (t, s) in enumerate(sel):
        exp = O.joint_counts(d, npv_o, int(t), int(s)).ravel()
        assert np.array_equal(counts[r, : exp.size], exp), (t, s)
    assert np.array_equal(m2.cpu().numpy(), m)


def test_c3_surrogate_stats_1000x3000(dev):
    from tenex_amd.inference.surrogate_test import surrogate_accumulate
    n, T, n_surr, bl = 1000, 3000, 100, 50
    x = synthetic.zero_inflated_lognormal(n, T, seed=3)
    d, nb, npv, bm = O.discretize(x)
    d = d.astype(np.int32)
    bins, _, _, T = core.upload_bins(d, dev)
    prep = core.prepare(bins, torch.from_numpy(npv).to(dev), T, 1)
    obs_int = core.te_matrix_internal(prep)
    obs_pub = obs_int.t().contiguous()
    s, s2, cg = surrogate_accumulate(prep, obs_pub, n_surr, "block", bl, 42, "cpu")
    perm, eff, nbl = core.block_perm_device(n, T, "block", bl, 0, n_surr, 42, "cpu", "cpu")
    assert eff == bl and nbl == 60
    rng = np.random.default_rng(3)
    allp = O.pairfree_pairs(n)
    pairs = allp[rng.choice(len(allp), 2000, replace=False)]          # [target, source]
    obs_pairs = obs_int.cpu().numpy()[pairs[:, 0], pairs[:, 1]]
    os_, os2, ocg, otk = O.surrogate_stats(d, npv, perm.numpy(), bl, pairs, obs_pairs,
                                           return_te_k=True)
    slot = pairs[:, 1] * n + pairs[:, 0]                               # public [source, target]
    sp = s.cpu().numpy().ravel()[slot]
    sp2 = s2.cpu().numpy().ravel()[slot]
    cgp = cg.cpu().numpy().ravel()[slot]
    assert np.abs(sp - os_).max() < n_surr * TOL
    assert np.abs(sp2 - os2).max() < n_surr * 2 * TOL
    # counts of te_k >= observed (float32 compare): equal wherever no te_k
    # lies within the TE tolerance of the observed value
    near = np.any(np.abs(otk - obs_pairs[:, None]) < 2 * TOL, axis=1)
    assert near.mean() < 0.05
    assert np.array_equal(cgp[~near], ocg[~near])


def test_c4_learned_vs_plain_6000x3000(dev):
    import tenex_amd as tnx
    n, T = 6000, 3000
    x = synthetic.zero_inflated_lognormal(n, T, seed=4)
    rng = np.random.default_rng(4)
    names = np.array([f"g{i}" for i in range(n)])
    tf = rng.choice(n, 200, replace=False)
    rest = np.setdiff1d(np.arange(n), tf)
    tg = rng.choice(rest, 2000, replace=False)
    for s, t in zip(tf[:100], tg[:100]):       # planted regulation
        x[t, 1:] = x[s, :-1] * (0.5 + rng.random(T - 1).astype(np.float32))
    res = tnx.TransferEntropyEngine(x, names, sources=names[tf]).compute(accelerator="gpu")
    nw = tnx.NetWeaver(res, fdr=0.05)
    kw = dict(n_surrogates=100, block_length=50, targets=names[tg], pairs_scope="sources")
    plain = nw.infer("surrogate_test", **kw)
    learned = nw.infer("surrogate_test", learned="mlp", **kw)
    a = set(map(tuple, plain.grn.pairs.tolist()))
    b = set(map(tuple, learned.grn.pairs.tolist()))
    jac = len(a & b) / max(1, len(a | b))
    assert len(a) >= 100 and jac >= 0.98, (len(a), len(b), jac)
    info = learned.learned_info
    assert info["n_pairs"] == 200 * 2000
    rr = info["rerun_pairs"]                   # public [source, target]
    assert len(rr) > 0
    for f in ("mean_surrogate_te", "std_surrogate_te"):
        pm, lm = getattr(plain, f), getattr(learned, f)
        assert np.array_equal(pm[rr[:, 0], rr[:, 1]], lm[rr[:, 0], rr[:, 1]]), f
