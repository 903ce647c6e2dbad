"""Variables of more than 127 levels (int16 bins) and the occupied-key
kernel (tnx_te_sort_pairs / tnx_gene_prep_sort) against the FP64 oracle."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tenex_amd import core

pytestmark = pytest.mark.gpu
TOL = 1e-6


def _wide_bins(n, T, levels, seed):
    """Dense rows: variable g has levels[g] levels, every level present."""
    rng = np.random.default_rng(seed)
    d = np.empty((n, T), np.int32)
    for g, b in enumerate(levels):
        d[g] = rng.integers(0, b, T)
        d[g, rng.choice(T, b, replace=False)] = np.arange(b)
    return d


@pytest.mark.parametrize("T,tau", [(1200, 1), (3000, 2), (9000, 1)])
def test_wide_matrix_vs_oracle(dev, T, tau):
    """A matrix mixing variables of 2..40 and 128..600 levels: pairs with a
    wide variable on the occupied-key kernel, the rest on the histogram
    kernels; every pair within 1e-6 of the FP64 oracle, wide S1 / S2a from
    sorted keys equal to the oracle's entropies."""
    n = 24
    rng = np.random.default_rng(T)
    levels = np.where(np.arange(n) % 4 == 0, rng.integers(128, 600, n), rng.integers(2, 40, n))
    d = _wide_bins(n, T, levels, T)
    d[5, tau:] = d[0, :-tau]                       # a wide source driving a narrow target
    b8, b16, npv, b_max, T2 = core.upload_bins_any(d, dev)
    assert b16 is not None and b_max == int(levels.max())
    prep = core.prepare(b8, npv, T, tau, bins16=b16)
    assert len(prep.wide) == int((d.max(axis=1) + 1 > 127).sum())   # row 5 copies wide row 0
    m = core.te_matrix_internal(prep).cpu().numpy()
    pairs = O.pairfree_pairs(n)
    te64 = O.te_pairs_sparse(d, pairs, tau=tau)
    err = np.abs(m[pairs[:, 0], pairs[:, 1]].astype(np.float64) - te64)
    assert err.max() < TOL, err.max()
    assert m[5, 0] > 1.0
    S1h = prep.S1.cpu().numpy()
    S2h = prep.S2a.cpu().numpy()
    for g in prep.wide:
        x, xf = d[g, :T - tau], d[g, tau:]
        c1 = np.bincount(x).astype(np.float64)
        c2 = np.unique(np.stack([xf, x]), axis=1, return_counts=True)[1].astype(np.float64)
        f = lambda c: float(np.sum(c[c > 1] * np.log2(c[c > 1])))  # noqa: E731
        assert abs(S1h[g] - f(c1)) < 1e-9 * max(1.0, f(c1))
        assert abs(S2h[g] - f(c2)) < 1e-9 * max(1.0, f(c2))


def test_sort_pairs_narrow_equal_histogram(dev):
    """The occupied-key kernel on narrow pairs gives the histogram kernels'
    TE (both exact fixed-point sums of the same counts)."""
    n, T = 30, 2500
    d = _wide_bins(n, T, np.random.default_rng(3).integers(2, 60, n), 3)
    b8, b16, npv, b_max, _ = core.upload_bins_any(d, dev)
    assert b16 is None
    prep = core.prepare(b8, npv, T, 1)
    m = core.te_matrix_internal(prep)
    pairs = O.pairfree_pairs(n)
    out = torch.zeros((n, n), dtype=torch.float32, device=dev)
    core.te_sort_pairs(prep, pairs, out, n)
    assert torch.abs(out - m).max().item() < 1e-6
    vec = torch.zeros(len(pairs), dtype=torch.float32, device=dev)
    core.te_sort_pairs(prep, pairs, vec, 0)
    pt = torch.from_numpy(pairs).to(dev)
    assert torch.equal(vec, out[pt[:, 0], pt[:, 1]])


def test_wide_through_plugin_and_pair_list(dev):
    """The kernel plugin API (compute_single_gpu) on int16-valued bins with
    an explicit pair list mixing wide and narrow pairs."""
    from tenex_amd.kernels.lds_hist import AdaptiveSMEMKernel
    n, T = 16, 1500
    levels = np.array([300, 5, 7, 200, 3, 9, 40, 2, 130, 6, 4, 8, 12, 2, 3, 5])
    d = _wide_bins(n, T, levels, 9)
    k = AdaptiveSMEMKernel()
    assert k.supports(int(levels.max()), True, 160 * 1024, 0, None, False)
    rng = np.random.default_rng(1)
    pairs = rng.integers(0, n, size=(200, 2))
    got = k.compute_single_gpu(d[:, :, None].astype(np.int16), pairs, int(levels.max()),
                               levels, 1, 1024, dev)
    v = pairs[:, 0] != pairs[:, 1]
    exp = np.zeros(len(pairs))
    exp[v] = O.te_pairs_sparse(d, pairs[v])
    assert np.abs(got - exp).max() < TOL
    with pytest.raises(NotImplementedError):
        b8, b16, npv, _, _ = core.upload_bins_any(d, dev)
        core.te_pairfree(core.prepare(b8, npv, T, 1, bins16=b16))


def test_large_pairs_routed_to_sort_match_dense(dev):
    """Short series with many levels: the matrix path sends pairs above
    sort_threshold(L) cells to the occupied-key kernel and skips them in the
    histogram launch (cap_sort); the result equals the all-dense launch and
    the FP64 oracle."""
    from tenex_amd import _lib
    n, T = 40, 1200
    levels = np.random.default_rng(5).integers(2, 110, n)
    d = _wide_bins(n, T, levels, 5)
    b8, b16, npv, _, _ = core.upload_bins_any(d, dev)
    prep = core.prepare(b8, npv, T, 1)
    cap = core.sort_threshold(prep.L)
    npvh = prep.npv_host.astype(np.int64)
    pairs = O.pairfree_pairs(n)
    cells = npvh[pairs[:, 0]] ** 2 * npvh[pairs[:, 1]]
    assert cap > 0 and (cells > cap).sum() > 200
    m = core.te_matrix_internal(prep).cpu().numpy()
    plan = core.get_plan(prep)
    dense = torch.zeros((n, n), dtype=torch.float32, device=dev)
    core.te_run(prep, plan, dense, _lib.TNX_OUT_MATRIX, ld_out=n)
    dense = dense.cpu().numpy()
    assert np.abs(m - dense).max() < 1e-6
    te64 = O.te_pairs_sparse(d, pairs)
    assert np.abs(m[pairs[:, 0], pairs[:, 1]].astype(np.float64) - te64).max() < TOL
    # the histogram launch with cap_sort leaves exactly the large pairs unwritten
    skip = torch.full((n, n), -7.0, dtype=torch.float32, device=dev)
    args = core.te_args(prep, plan, skip, _lib.TNX_OUT_MATRIX, ld_out=n, cap_sort=cap)
    core.te_run(prep, plan, skip, _lib.TNX_OUT_MATRIX, args=args)
    sk = skip.cpu().numpy()[pairs[:, 0], pairs[:, 1]]
    assert np.all((sk == -7.0) == (cells > cap))
