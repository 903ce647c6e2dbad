"""GPU DPI trimming (tnx_dpi_maxmin) against the restated _trim of
tenex/inference/fdr.py:61-133 and against the CPU device path: identical
edges and weights (comparisons only, so bit for bit)."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tenex_amd.inference.fdr import trim_dpi

pytestmark = pytest.mark.gpu


def _net(rng, n, dens, neg=0.02):
    A = np.where(rng.random((n, n)) < dens, rng.random((n, n)), 0).astype(np.float32)
    A[rng.random((n, n)) < neg] *= -1
    np.fill_diagonal(A, 0)
    return A


@pytest.mark.parametrize("n,dens,bs,thr", [(1, 0.5, 512, 0.0), (63, 0.1, 512, 0.0),
                                           (130, 0.05, 16, 0.0), (257, 0.03, 512, 0.02),
                                           (300, 0.2, 100, 0.0)])
def test_dpi_gpu_vs_oracle(dev, n, dens, bs, thr):
    rng = np.random.default_rng(n)
    A = _net(rng, n, dens)
    A[: n // 5] = np.where(A[: n // 5] > 0, 0, A[: n // 5])
    p = np.argwhere(A != 0)
    names = np.array([f"g{i}" for i in range(n)])
    got = trim_dpi(A, names, p, A[p[:, 0], p[:, 1]], dev, batch_size=bs, trim_threshold=thr)
    ref = O.trim_dpi(A, batch_size=bs, trim_threshold=thr)
    if ref is None:
        assert np.array_equal(got.pairs, p)
        return
    assert np.array_equal(got.pairs, ref[0]) and np.array_equal(got.te, ref[1])


def test_dpi_gpu_vs_cpu_path_large(dev):
    """1,500 genes, 2% density with 300 source-only rows: GPU kernel and the
    CPU device path keep the same edges."""
    rng = np.random.default_rng(5)
    n = 1500
    A = _net(rng, n, 0.02, neg=0.0)
    p = np.argwhere(A != 0)
    names = np.array([f"g{i}" for i in range(n)])
    g = trim_dpi(A, names, p, A[p[:, 0], p[:, 1]], dev)
    c = trim_dpi(A, names, p, A[p[:, 0], p[:, 1]], torch.device("cpu"))
    assert np.array_equal(g.pairs, c.pairs) and np.array_equal(g.te, c.te)
    assert 0 < len(g) < len(p)


def test_fdr_infer_with_trimming(dev):
    import tenex_amd as tnx
    rng = np.random.default_rng(2)
    n = 200
    M = rng.random((n, n)).astype(np.float32) * 0.01
    M[rng.random((n, n)) < 0.03] += 0.5
    np.fill_diagonal(M, 0)
    names = np.array([f"g{i}" for i in range(n)])
    nw = tnx.NetWeaver((M, names), fdr=0.05)
    grn, trimmed = nw.infer("fdr", device=dev)
    trimmed_c = trim_dpi(M, names, grn.pairs, grn.te, torch.device("cpu"))
    assert np.array_equal(trimmed.pairs, trimmed_c.pairs)
    assert np.array_equal(trimmed.te, trimmed_c.te)
    assert len(trimmed) < len(grn)
