"""The persistent single launch (te_persist.hip) against the FP64 oracle and
against the two-launch kernels.

Counts are compared bit for bit through the kernels' own count dump (the
production code path, wave and block roles, single and multi-pass, with and
without the key-0 diversion); TE within 1e-6 absolute of the FP64 oracle.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tenex_amd import _lib, core, synthetic

pytestmark = pytest.mark.gpu

TOL = 1e-6


def _prep(bins_np, dev, tau=1):
    bins, npv, b_max, T = core.upload_bins(bins_np, dev)
    return core.prepare(bins, npv, T, tau)


def _matrix(prep, plan, persistent):
    n = prep.n_vars
    out = torch.full((n, n), -1.0, dtype=torch.float32, device=prep.device)
    args = core.te_args(prep, plan, out, _lib.TNX_OUT_MATRIX, ld_out=n, persistent=persistent)
    core.te_run(prep, plan, out, _lib.TNX_OUT_MATRIX, args=args)
    return out.cpu().numpy()


def _wave_mask(prep, plan):
    npv = prep.npv_host.astype(np.int64)
    return npv[:, None] ** 2 * npv[None, :] <= plan.info.cap_wave


@pytest.mark.parametrize("cap_wave", [0, 1536])
def test_persistent_matches_two_launch_and_oracle(dev, cap_wave):
    """The same counts as the two-launch kernels, hence TE that differs only
    by the order of the compensated sums: wave pairs within 1e-9, block pairs
    (256 instead of 512 threads per pair) within 1e-6; all pairs within
    1e-6 of the FP64 oracle."""
    b = synthetic.mixed_cardinality_bins(150, 3000, seed=31)
    prep = _prep(b, dev)
    plan = core.get_plan(prep, cap_wave=cap_wave)
    assert plan.info.n_block_items > 0 and core.persistent_ok(prep, plan)
    m_p = _matrix(prep, plan, True)
    m_2 = _matrix(prep, plan, False)
    w = _wave_mask(prep, plan)
    np.fill_diagonal(w, False)
    assert np.abs(m_p[w].astype(np.float64) - m_2[w]).max() < 1e-9
    off = ~np.eye(150, dtype=bool)
    assert np.abs(m_p[off].astype(np.float64) - m_2[off]).max() < TOL
    pairs = O.pairfree_pairs(150)
    te64 = O.te_pairs(b.astype(np.int32), prep.npv_host, pairs)
    err = np.abs(m_p[pairs[:, 0], pairs[:, 1]].astype(np.float64) - te64)
    assert err.max() < TOL, err.max()


# one series length per register variant NI (1024-sample chunks per lane)
@pytest.mark.parametrize("T,ni", [(700, 1), (1500, 2), (3000, 3), (4001, 4), (5500, 6),
                                  (7777, 8), (10000, 10), (12289, 12)])
@pytest.mark.parametrize("hot", [False, True])
def test_register_variants_vs_oracle(dev, T, ni, hot):
    """Every register variant of the persistent kernel, on data with and
    without a frequent key 0 (the diverting and the plain kernels), wave and
    block pairs, against the FP64 oracle."""
    n = 28
    if hot:
        d, nb, npv, bm = O.discretize(synthetic.zero_inflated_lognormal(n, T, seed=ni))
    else:
        d = synthetic.mixed_cardinality_bins(n, T, seed=ni).astype(np.int32)
    prep = _prep(d, dev)
    assert -(-prep.L // 1024) <= ni and (ni == 1 or -(-prep.L // 1024) > {2: 1, 3: 2, 4: 3, 6: 4,
                                                                          8: 6, 10: 8, 12: 10}[ni])
    if hot:
        assert prep.hot_zero
    plan = core.get_plan(prep)
    assert core.persistent_ok(prep, plan)
    m = _matrix(prep, plan, True)
    pairs = O.pairfree_pairs(n)
    te64 = O.te_pairs(d.astype(np.int32), prep.npv_host, pairs)
    err = np.abs(m[pairs[:, 0], pairs[:, 1]].astype(np.float64) - te64)
    assert err.max() < TOL, err.max()


def _bounded_levels(n, T, b_hi, zero_frac, seed):
    """Dense bins with 2..b_hi levels per variable, every level present; a
    fraction zero_frac of the samples forced to level 0 (a frequent key 0)."""
    rng = np.random.default_rng(seed)
    b = rng.integers(2, b_hi + 1, size=n)
    d = (rng.random((n, T)) * b[:, None]).astype(np.int32)
    d[rng.random((n, T)) < zero_frac] = 0
    for g in range(n):
        d[g, rng.choice(T, b[g], replace=False)] = np.arange(b[g])
    return d


@pytest.mark.parametrize("hot", [False, True])
def test_counts_dump_wave_and_block_roles(dev, hot):
    """The kernels' count dump on a full plan: pairs of the wave role and of
    the block role (single pass and several passes: a small wave histogram
    leaves a small block histogram) equal the oracle's brute-force counts."""
    n, T = 70, 2500
    d = _bounded_levels(n, T, 16, 0.45 if hot else 0.0, seed=8)
    prep = _prep(d, dev)
    assert prep.hot_zero == hot
    plan = core.get_plan(prep, cap_wave=128)
    assert core.persistent_ok(prep, plan) and plan.info.n_block_items > 0
    npv = prep.npv_host.astype(np.int64)
    pairs = O.pairfree_pairs(n)
    cells = npv[pairs[:, 0]] ** 2 * npv[pairs[:, 1]]
    capB = core.block_cap(prep, plan)
    rng = np.random.default_rng(1)
    pick = np.concatenate([rng.choice(np.nonzero(cells <= 128)[0], 60, replace=False),
                           rng.choice(np.nonzero((cells > 128) & (cells <= capB))[0], 30,
                                      replace=False),
                           rng.choice(np.nonzero(cells > capB)[0],
                                      min(30, int((cells > capB).sum())), replace=False)])
    assert (cells[pick] > capB).sum() > 0, "no multi-pass block pair in the sample"
    sel = pairs[pick]
    counts, m = core.counts_plan(prep, plan, sel, int(cells[pick].max()))
    counts = counts.cpu().numpy()
    for r, (t, s) in enumerate(sel):
        exp = O.joint_counts(d, prep.npv_host, int(t), int(s)).ravel()
        assert np.array_equal(counts[r, : exp.size], exp), (t, s, exp.size)
        assert counts[r, exp.size:].max(initial=0) == 0
    te64 = O.te_pairs(d, prep.npv_host, pairs)
    got = m.cpu().numpy()[pairs[:, 0], pairs[:, 1]]
    assert np.abs(got.astype(np.float64) - te64).max() < TOL


def test_counts_debug_goes_through_production_launch(dev):
    """core.counts_debug runs a pair-list plan in the persistent kernel:
    duplicate pairs, wave and block pairs, bit-exact counts."""
    d = synthetic.mixed_cardinality_bins(50, 1200, seed=3).astype(np.int32)
    prep = _prep(d, dev)
    rng = np.random.default_rng(2)
    pairs = rng.integers(0, 50, size=(150, 2))
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    pairs = np.concatenate([pairs, pairs[:5]])  # duplicates
    npv = prep.npv_host.astype(np.int64)
    mc = int((npv[pairs[:, 0]] ** 2 * npv[pairs[:, 1]]).max())
    counts, te = core.counts_debug(prep, pairs, max_cells=mc)
    counts = counts.cpu().numpy()
    for r, (t, s) in enumerate(pairs):
        exp = O.joint_counts(d, prep.npv_host, int(t), int(s)).ravel()
        assert np.array_equal(counts[r, : exp.size], exp), (t, s)
    te64 = O.te_pairs(d, prep.npv_host, pairs)
    assert np.abs(te.cpu().numpy().astype(np.float64) - te64).max() < TOL


def test_persistent_output_modes(dev):
    """Pair-free output ranges through the persistent launch: every range
    equals the same slice of the full pair-free vector."""
    b = synthetic.mixed_cardinality_bins(40, 900, seed=9)
    prep = _prep(b, dev)
    full = core.te_pairfree(prep).cpu().numpy()
    te64 = O.te_pairs(b.astype(np.int32), prep.npv_host, O.pairfree_pairs(40))
    assert np.abs(full - te64).max() < TOL
    for off, cnt in [(0, 1), (5, 77), (700, 860), (1559, 1)]:
        part = core.te_pairfree(prep, off, cnt).cpu().numpy()
        assert np.array_equal(part, full[off:off + cnt])


def test_persistent_graph_replays(dev):
    """The persistent launch (queue counters cleared by a kernel, not a
    memset) captured in a HIP graph: replays with eager work between them
    equal the eager result bit for bit."""
    b = synthetic.mixed_cardinality_bins(90, 1500, seed=21)
    prep = _prep(b, dev)
    plan = core.get_plan(prep, cap_wave=1536)
    assert plan.info.n_block_items > 0 and core.persistent_ok(prep, plan)
    out = torch.zeros((90, 90), dtype=torch.float32, device=dev)
    g = core.capture_te_step(prep, plan, out, _lib.TNX_OUT_MATRIX, ld_out=90)
    rng = np.random.default_rng(4)
    for rep in range(3):
        perm = torch.from_numpy(rng.permutation(prep.T)).to(dev)
        prep.bins[:, :prep.T] = prep.bins[:, :prep.T][:, perm]
        g.replay()
        torch.cuda.synchronize()
        got = out.clone()
        eager = core.te_matrix_internal(prep, cap_wave=1536)
        assert torch.equal(got, eager), rep


@pytest.mark.parametrize("cap_wave", [0, 256])
def test_hot_states_edge_levels(dev, cap_wave):
    """Hot states with targets of one and two levels (every state hot, or
    the only state hot), sources of one level, a short last chunk, and with a
    small wave histogram the block role in one and several passes: the hot
    kernels equal the plain ones bit for bit and the FP64 oracle within
    1e-6."""
    n, T = 40, 2345
    d, nb, npv, bm = O.discretize(synthetic.zero_inflated_lognormal(n, T, seed=21))
    d = d.astype(np.int32)
    d[0] = 0                                   # constant (one level)
    d[1] = (np.arange(T) % 3 == 0)             # two levels
    d[2] = (np.arange(T) % 7 == 0) * 2         # levels {0, 2}: relabelled to {0, 1} below
    d[2] = np.unique(d[2], return_inverse=True)[1]
    prep = _prep(d, dev)
    assert prep.hot_zero and prep.hot[0] > 0
    plan = core.get_plan(prep, cap_wave=cap_wave)
    assert core.persistent_ok(prep, plan)
    if cap_wave:
        npvh = prep.npv_host.astype(np.int64)
        cells = npvh[:, None] ** 2 * npvh[None, :]
        assert plan.info.n_block_items > 0 and (cells > core.block_cap(prep, plan)).any()
    m_hot = _matrix(prep, plan, True)
    prep.hot_zero = False
    assert prep.hot == (0, 0)
    plan0 = core.get_plan(prep, cap_wave=cap_wave)
    m_plain = _matrix(prep, plan0, True)
    assert np.array_equal(m_hot, m_plain)
    pairs = O.pairfree_pairs(n)
    te64 = O.te_pairs(d, prep.npv_host, pairs)
    err = np.abs(m_hot[pairs[:, 0], pairs[:, 1]].astype(np.float64) - te64)
    assert err.max() < TOL, err.max()
