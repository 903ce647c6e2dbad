"""GPU parity of GEMM-B2 (INT8 and FP4 MFMA contractions, both exact) against
the oracle."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tenex_amd import _lib, core, synthetic
from tenex_amd.kernels.gemm_b2 import te_gemm_b2_device

pytestmark = pytest.mark.gpu
TOL = 1e-6


def _binary_prep(x, dev):
    d, nb, npv, bm = O.discretize(x)
    assert bm <= 2
    bins, npv_d, b_max, T = core.upload_bins(d, dev)
    # keep the remap's level counts (constant genes have one level)
    return d, npv, core.prepare(bins, torch.from_numpy(npv).to(dev), T, 1)


# T = 9000: L + 1 > 8192, so the epilogue reads the c*log2(c) table from
# memory (double-float sums) instead of LDS
@pytest.mark.parametrize("n,T,seed", [(150, 300, 0), (97, 1001, 1), (260, 64, 2), (33, 4097, 3),
                                      (40, 9000, 4)])
@pytest.mark.parametrize("precision", [_lib.TNX_B2_INT8, _lib.TNX_B2_FP4])
def test_gemm_b2_vs_oracle(dev, n, T, seed, precision):
    x, planted = synthetic.bernoulli_planted(n, T, n_planted=min(20, n // 3), seed=seed)
    x[0] = 0.0  # a constant gene
    d, npv, prep = _binary_prep(x, dev)
    te, cnt = te_gemm_b2_device(prep, counts_dbg=True, precision=precision)
    te = te.cpu().numpy()
    cnt = cnt.cpu().numpy()
    ref, C, D, E = O.te_b2(d, counts=True)
    assert np.array_equal(cnt[0], C) and np.array_equal(cnt[1], D) and np.array_equal(cnt[2], E)
    assert np.abs(te - ref).max() < TOL
    pairs = O.pairfree_pairs(n)
    te_s = O.te_pairs(d, npv, pairs)
    assert np.abs(te[pairs[:, 0], pairs[:, 1]] - te_s).max() < TOL
    assert np.all(np.diag(te) == 0)


def test_gemm_b2_planted_closed_form(dev):
    """y[t+1] = x[t] exactly: TE(x -> y) = H(Y_{t+1} | Y_t) from y's own counts."""
    n, T = 64, 2000
    x, planted = synthetic.bernoulli_planted(n, T, n_planted=10, flip=0.0, seed=5)
    d, npv, prep = _binary_prep(x, dev)
    te = te_gemm_b2_device(prep).cpu().numpy()
    s1, s2a = O.gene_entropies(d, npv)
    L = T - 1
    for s, t in planted:
        h_cond = (s1[t] - s2a[t]) / L  # H(Y_{t+1}|Y_t) = (S1 - S2a)/L
        assert abs(te[t, s] - h_cond) < TOL


@pytest.mark.parametrize("precision", [_lib.TNX_B2_INT8, _lib.TNX_B2_FP4])
def test_gemm_b2_c1_scale_counts(dev, precision):
    """2,000 genes x 5,000 cells with 200 planted pairs: counts of a row
    block bit-exact, marginal identities on the full matrix."""
    n, T = 2000, 5000
    x, planted = synthetic.bernoulli_planted(n, T, n_planted=200, seed=7)
    d, npv, prep = _binary_prep(x, dev)
    te, cnt = te_gemm_b2_device(prep, counts_dbg=True, precision=precision)
    cnt = cnt.cpu().numpy()
    p = d[:, :-1].astype(np.int64)
    f = d[:, 1:].astype(np.int64)
    rows = np.arange(0, n, 97)
    Cr = (f[rows] * p[rows]) @ p.T
    Dr = p[rows] @ p.T
    Er = f[rows] @ p.T
    assert np.array_equal(cnt[0][rows], Cr)
    assert np.array_equal(cnt[1][rows], Dr)
    assert np.array_equal(cnt[2][rows], Er)
    assert np.array_equal(cnt[1], cnt[1].T)
    assert np.array_equal(np.diag(cnt[1]), p.sum(1))
    te = te.cpu().numpy()
    pairs = np.stack([np.repeat(rows, 40), np.tile(np.arange(0, n, 50), len(rows))], 1)
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    te_s = O.te_pairs(d, npv, pairs)
    assert np.abs(te[pairs[:, 0], pairs[:, 1]] - te_s).max() < TOL
    # planted edges dominate their rows
    for s, t in planted[:20]:
        assert te[t, s] > 0.05


def test_gemm_b2_fp4_equals_int8(dev):
    """The default (FP4 below 2^24 samples) and the INT8 contractions give
    the same counts, so the same TE matrix bit for bit."""
    x, planted = synthetic.bernoulli_planted(300, 2500, n_planted=40, seed=9)
    d, npv, prep = _binary_prep(x, dev)
    te4, c4 = te_gemm_b2_device(prep, counts_dbg=True)
    te8, c8 = te_gemm_b2_device(prep, counts_dbg=True, precision=_lib.TNX_B2_INT8)
    assert torch.equal(c4, c8)
    assert torch.equal(te4, te8)
