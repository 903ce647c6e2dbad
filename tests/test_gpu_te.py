"""GPU parity of the LDS histogram TE path against the FP64 oracle.

Counts must match bit for bit; TE within 1e-6 absolute of the FP64 oracle.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tenex_amd import _lib, core, synthetic

pytestmark = pytest.mark.gpu

TOL = 1e-6


def _prep_from_dense(bins_np, dev, tau=1):
    bins, npv, b_max, T = core.upload_bins(bins_np, dev)
    return core.prepare(bins, npv, T, tau)


def test_discretize_bit_exact_c0(dev):
    x = synthetic.zero_inflated_lognormal(100, 500, seed=0)
    raw_g, nb_g = core.fsbw_l_gpu(torch.from_numpy(x).to(dev))
    raw_o, nb_o, _, _ = O.fsbw_l_numpy(x)
    assert np.array_equal(raw_g.cpu().numpy(), raw_o)
    assert np.array_equal(nb_g.cpu().numpy(), nb_o)
    bins, npv, b_max = core.remap_gpu(raw_g)
    d_o, npv_o, bmax_o = O.remap(raw_o)
    assert np.array_equal(bins[:, :500].cpu().numpy().astype(np.int32), d_o)
    assert np.array_equal(npv.cpu().numpy(), npv_o) and b_max == bmax_o


@pytest.mark.parametrize("n,T,seed", [(37, 9999, 1), (64, 3000, 2), (16, 8193, 3), (50, 130, 4),
                                      (600, 10000, 2)])
def test_discretize_bit_exact_lengths(dev, n, T, seed):
    x = synthetic.zero_inflated_lognormal(n, T, seed=seed)
    raw_g, nb_g = core.fsbw_l_gpu(torch.from_numpy(x).to(dev))
    raw_o, nb_o, _, _ = O.fsbw_l_numpy(x)
    assert np.array_equal(raw_g.cpu().numpy(), raw_o)
    assert np.array_equal(nb_g.cpu().numpy(), nb_o)


def test_gene_prep_entropies(dev):
    b = synthetic.mixed_cardinality_bins(60, 777, seed=5)
    prep = _prep_from_dense(b, dev)
    s1, s2 = O.gene_entropies(b.astype(np.int32), prep.npv_host)
    np.testing.assert_allclose(prep.S1.cpu().numpy(), s1, rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(prep.S2a.cpu().numpy(), s2, rtol=1e-12, atol=1e-9)
    st = prep.states.cpu().numpy().view(np.uint16)[:, : prep.L].astype(np.int64)
    exp = b[:, 1:].astype(np.int64) * prep.npv_host[:, None] + b[:, :-1]
    assert np.array_equal(st, exp)


def test_counts_bit_exact(dev):
    x = synthetic.zero_inflated_lognormal(100, 500, seed=0)
    d, nb, npv, bm = O.discretize(x)
    prep = _prep_from_dense(d, dev)
    rng = np.random.default_rng(0)
    pairs = rng.integers(0, 100, size=(200, 2))
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    counts, te = core.counts_debug(prep, pairs, max_cells=2048)
    counts = counts.cpu().numpy()
    for p, (t, s) in enumerate(pairs):
        exp = O.joint_counts(d, npv, int(t), int(s)).ravel()
        assert np.array_equal(counts[p, : exp.size], exp), (t, s)
    te64 = O.te_pairs(d, npv, pairs)
    assert np.abs(te.cpu().numpy() - te64).max() < TOL


def test_te_matrix_c0_vs_oracle(dev):
    """Config C0: 100 genes x 500 cells, all 9,900 pairs."""
    x = synthetic.zero_inflated_lognormal(100, 500, seed=0)
    d, nb, npv, bm = O.discretize(x)
    prep = _prep_from_dense(d, dev)
    m = core.te_matrix_internal(prep).cpu().numpy()
    ref = O.te_matrix(d, npv)  # public layout
    got = m.T
    assert np.abs(got.astype(np.float64) - ref).max() < TOL
    assert np.all(np.diag(got) == 0)


@pytest.mark.parametrize("cap_wave", [0, 64, 512])
def test_te_mixed_cardinality_vs_oracle(dev, cap_wave):
    """Mixed cardinality incl. b up to 64: wave and multi-pass block kernels."""
    b = synthetic.mixed_cardinality_bins(120, 2500, seed=7)
    prep = _prep_from_dense(b, dev)
    plan = core.get_plan(prep, cap_wave=cap_wave)
    assert plan.info.n_block_items > 0
    m = core.te_matrix_internal(prep, cap_wave=cap_wave).cpu().numpy()
    pairs = O.pairfree_pairs(120)
    te64 = O.te_pairs(b.astype(np.int32), prep.npv_host, pairs)
    got = m[pairs[:, 0], pairs[:, 1]]
    err = np.abs(got.astype(np.float64) - te64)
    assert err.max() < TOL, err.max()


@pytest.mark.parametrize("T,cap_wave", [(500, 0), (500, 64), (4100, 0), (4100, 512), (13000, 0), (13000, 512)])
def test_hot_zero_key_diversion(dev, T, cap_wave):
    """Zero-inflated data (key 0 holds about a quarter of the samples): the
    kernels with key 0 diverted to per-lane words and the plain kernels give
    the same TE bit for bit, within 1e-6 of the FP64 oracle, in the wave
    (register and memory target) and the multi-pass block kernels."""
    x = synthetic.zero_inflated_lognormal(60, T, seed=11)
    d, nb, npv, bm = O.discretize(x)
    prep = _prep_from_dense(d, dev)
    assert prep.hot_zero, core.key0_fraction(prep.states, prep.bins, prep.L)
    if cap_wave:
        assert core.get_plan(prep, cap_wave=cap_wave).info.n_block_items > 0
    m_hot = core.te_matrix_internal(prep, cap_wave=cap_wave).cpu().numpy()
    prep.hot_zero = False
    m_plain = core.te_matrix_internal(prep, cap_wave=cap_wave).cpu().numpy()
    assert np.array_equal(m_hot, m_plain)
    pairs = O.pairfree_pairs(60)
    te64 = O.te_pairs(d.astype(np.int32), prep.npv_host, pairs)
    err = np.abs(m_hot[pairs[:, 0], pairs[:, 1]].astype(np.float64) - te64)
    assert err.max() < TOL, err.max()


def test_hot_zero_flag_uniform_levels(dev):
    """Near-uniform levels keep key 0 rare: the plain kernels are chosen."""
    prep = _prep_from_dense(synthetic.mixed_cardinality_bins(50, 3000, seed=2), dev)
    assert not prep.hot_zero


def test_pairfree_ranges(dev):
    b = synthetic.mixed_cardinality_bins(40, 900, seed=9)
    prep = _prep_from_dense(b, dev)
    full = core.te_pairfree(prep).cpu().numpy()
    te64 = O.te_pairs(b.astype(np.int32), prep.npv_host, O.pairfree_pairs(40))
    assert np.abs(full - te64).max() < TOL
    for off, cnt in [(0, 1), (5, 77), (39, 39), (700, 860), (1559, 1)]:
        part = core.te_pairfree(prep, off, cnt).cpu().numpy()
        assert np.array_equal(part, full[off:off + cnt])


def test_sources_filter_matrix(dev):
    b = synthetic.mixed_cardinality_bins(50, 600, seed=11)
    prep = _prep_from_dense(b, dev)
    src = np.array([3, 7, 20, 49], np.int32)
    m = core.te_matrix_internal(prep, None, src).cpu().numpy()
    pairs = np.array([(t, s) for t in range(50) for s in src if s != t], np.int64)
    te64 = O.te_pairs(b.astype(np.int32), prep.npv_host, pairs)
    assert np.abs(m[pairs[:, 0], pairs[:, 1]] - te64).max() < TOL
    mask = np.zeros((50, 50), bool)
    mask[pairs[:, 0], pairs[:, 1]] = True
    assert np.all(m[~mask] == 0)


def test_short_and_edge_series(dev):
    # L smaller than a wave step, constant genes (b = 1), tau > 1
    rng = np.random.default_rng(3)
    b = rng.integers(0, 3, size=(12, 7)).astype(np.int8)
    b[0] = 0
    d, npv, _ = O.remap(b.astype(np.int32))
    for tau in (1, 2, 5):
        prep = _prep_from_dense(d, dev, tau=tau)
        m = core.te_matrix_internal(prep).cpu().numpy()
        pairs = O.pairfree_pairs(12)
        te64 = O.te_pairs(d, npv, pairs, tau=tau)
        assert np.abs(m[pairs[:, 0], pairs[:, 1]] - te64).max() < TOL
        assert np.abs(m[0, :]).max() < TOL  # a constant target: TE 0 from every source
        # constant source gives zero TE
        col = m[pairs[:, 0], pairs[:, 1]][pairs[:, 1] == 0]
        assert np.all(np.abs(col) < TOL)


def test_hip_graph_replay_matches_eager(dev):
    """The preparation and TE launches captured in a HIP graph: several
    replays, with eager work between them and the bins changed in place,
    each equal bit for bit to the eager result on the same bins."""
    b = synthetic.mixed_cardinality_bins(90, 1500, seed=21)
    prep = _prep_from_dense(b, dev)
    plan = core.get_plan(prep, cap_wave=512)
    assert plan.info.n_block_items > 0
    out = torch.zeros((90, 90), dtype=torch.float32, device=dev)
    g = core.capture_te_step(prep, plan, out, _lib.TNX_OUT_MATRIX, ld_out=90)
    rng = np.random.default_rng(4)
    for rep in range(3):
        # permute each variable's time axis in place (levels stay dense)
        perm = torch.from_numpy(rng.permutation(prep.T)).to(dev)
        prep.bins[:, :prep.T] = prep.bins[:, :prep.T][:, perm]
        g.replay()
        torch.cuda.synchronize()
        got = out.clone()
        eager = core.te_matrix_internal(prep, cap_wave=512)
        assert torch.equal(got, eager), rep


@pytest.mark.parametrize("kind", ["pairs", "grid"])
def test_item_per_wave_matches_split(dev, kind):
    """A wave per work item (the pair-list default once the items fill the
    chip) and four waves splitting each item give the same bits; the TE
    matches the FP64 oracle."""
    b = synthetic.mixed_cardinality_bins(90, 1500, seed=21)
    prep = _prep_from_dense(b, dev)
    n = prep.n_vars
    if kind == "pairs":
        rng = np.random.default_rng(5)
        allp = O.pairfree_pairs(n)
        pairs = allp[rng.choice(len(allp), 2500, replace=False)]     # [target, source]
        plan = core.te_pairlist_plan(prep.npv_host, pairs, dev)
        assert np.all(plan.work_host[:plan.info.n_wave_items, 1] == 1)
    else:
        plan = core.get_plan(prep)
        pairs = O.pairfree_pairs(n)
    outs = []
    for ipw in (False, True):
        out = torch.full((n, n), -1.0, dtype=torch.float32, device=dev)
        args = core.te_args(prep, plan, out, _lib.TNX_OUT_MATRIX, ld_out=n, item_per_wave=ipw)
        core.te_run(prep, plan, out, _lib.TNX_OUT_MATRIX, args=args)
        outs.append(out.cpu().numpy())
    assert np.array_equal(outs[0], outs[1])
    te64 = O.te_pairs(b.astype(np.int32), prep.npv_host, pairs)
    got = outs[1][pairs[:, 0], pairs[:, 1]]
    assert np.abs(got.astype(np.float64) - te64).max() < TOL
