"""Ingest and binarization: load_scrna / align_data against the reference's
output on committed files, the binarizers' threshold arithmetic against
numpy (the reference's numpy backend), and the HIP binarizer and the
binarized engine against the reference-written fixture."""
import os
import shutil

import numpy as np
import pytest

from oracle import oracle as O

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = 1e-6


def load(name):
    return dict(np.load(os.path.join(G, name), allow_pickle=False))


def _q32(row, q):
    """numpy's linear quantile in float32 restated: v = (n - 1) q,
    g = v - floor v, lerp of the two order statistics from the nearer end."""
    n = len(row)
    s = np.sort(row)
    v = np.float32(n - 1) * np.float32(q)
    lo = int(np.floor(v))
    hi = min(lo + 1, n - 1)
    g = np.float32(v - np.float32(lo))
    a, b = s[lo], s[hi]
    d = np.float32(b - a)
    if g >= 0.5:
        return np.float32(b - np.float32(d * np.float32(np.float32(1) - g)))
    return np.float32(a + np.float32(d * g))


def _med32(row):
    s = np.sort(row)
    n = len(s)
    return s[n // 2] if n % 2 else np.float32(np.float32(s[n // 2 - 1] + s[n // 2]) / np.float32(2))


def test_threshold_restatements_match_numpy():
    """The float32 quantile and median formulas the HIP kernel evaluates give
    np.quantile / np.median bit for bit (random rows with zeros, ties and
    negatives, odd and even lengths)."""
    rng = np.random.default_rng(0)
    for it in range(2000):
        n = int(rng.integers(1, 400))
        row = (rng.lognormal(0, 2, n) * (rng.random(n) < 0.7)).astype(np.float32)
        if it % 3 == 0:
            row = np.round(row, 1) - np.float32(1.0)
        for q in (0.5, 0.1, 0.8, 1 / 3, 0.0, 1.0, 0.999):
            assert _q32(row, q) == np.quantile(row, q), (it, q)
        assert _med32(row) == np.median(row), it


def test_load_scrna_matches_reference(tmp_path):
    """load_scrna on the committed files (copied to a scratch folder) gives
    the reference's aligned data, names, sources and .npy cache files; a
    second call reads the caches."""
    import tenex_amd as tnx
    g = load("scrna_small.npz")
    src = os.path.join(G, "scrna_small")
    for fn in os.listdir(src):
        shutil.copy(os.path.join(src, fn), tmp_path)
    p = lambda f: str(tmp_path / f)  # noqa: E731
    for rep in range(2):
        r = tnx.load_scrna(p("expression.csv"), p("pseudotime.txt"), p("branch.txt"),
                           branch_id=1, sources=p("sources.txt"))
        assert r.data.dtype == np.float32 and np.array_equal(r.data, g["data"])
        assert list(r.gene_names) == list(g["gene_names"])
        assert list(r.sources) == list(g["sources"])
        rb = tnx.load_scrna(p("expression.csv"), p("pseudotime.txt"), p("branch.txt"),
                            branch_id=2, make_binary=True)
        assert np.array_equal(rb.data, g["data_b2"])
        assert sorted(f for f in os.listdir(tmp_path) if f.endswith(".npy")) == \
            list(g["caches"])


def test_align_data_and_errors():
    from tenex_amd.preprocess import align_data
    rng = np.random.default_rng(1)
    x = rng.random((3, 30)).astype(np.float32)
    pt = rng.random(30).astype(np.float32)
    br = rng.integers(0, 3, 30)
    out = align_data(x, pt, br, branch_id=2)
    sel = np.nonzero(br == 2)[0]
    assert out.shape == (3, len(sel))
    assert np.all(np.diff(pt[sel][np.argsort(pt[sel])]) >= 0)
    assert sorted(map(tuple, out.T.tolist())) == sorted(map(tuple, x[:, sel].T.tolist()))
    with pytest.raises(ValueError):
        align_data(x, pt[:-1], br)
    with pytest.raises(ValueError):
        align_data(x, pt, br, branch_id=7)
    with pytest.raises(ValueError):
        import tenex_amd as tnx
        tnx.load_scrna(x, pt, br)          # arrays need gene_names


def test_binarizer_names_and_errors():
    import tenex_amd as tnx
    from tenex_amd.binarize import binarizer_doc
    assert tnx.available_binarizers() == ["mean", "median", "nonzero", "quantile"]
    assert "median" in binarizer_doc("median")
    with pytest.raises(ValueError):
        tnx.binarize(np.zeros((2, 5), np.float32), method="otsu")
    with pytest.raises(ValueError):
        tnx.binarize(np.zeros((2, 5), np.float32), backend="opencl")


@pytest.mark.gpu
def test_hip_binarize_thresholds_bit_exact(dev):
    """tnx_binarize's thresholds equal numpy's mean / median / quantile bit
    for bit and its bins the numpy backend's, with zero row padding."""
    import torch
    from tenex_amd.binarize import binarize_gpu
    rng = np.random.default_rng(2)
    for T in (1, 2, 7, 400, 401, 5000, 70001):
        x = (rng.lognormal(0, 2, (24, T)) * (rng.random((24, T)) < 0.6)).astype(np.float32)
        x[3] = np.round(x[3], 1) - 1.0
        x[4] = 2.5
        xd = torch.from_numpy(x).to(dev)
        for method, ref in (("nonzero", lambda r: np.float32(0)),
                            ("mean", lambda r: r.mean()),
                            ("median", lambda r: np.median(r)),
                            ("quantile", lambda r: np.quantile(r, 0.8))):
            bins, thr = binarize_gpu(xd, method, q=0.8)
            thr = thr.cpu().numpy()
            b = bins.cpu().numpy()
            for g_ in range(24):
                assert thr[g_] == ref(x[g_]), (T, method, g_)
            assert np.array_equal(b[:, :T], (x > thr[:, None]).astype(np.int8)), (T, method)
            assert not b[:, T:].any()


@pytest.mark.gpu
def test_golden_binarized_engine(dev):
    """engine.compute(binarization=...) on the GPU against the reference's
    CPU run: bins and n_per_var identical, TE within 1e-6 (GEMM-B2 for the
    full matrix, the histogram kernels with a source filter)."""
    import tenex_amd as tnx
    g = load("binarize_60x400.npz")
    names = np.array([f"g{i}" for i in range(60)])
    for tag, method, kw in (("median", "median", None), ("q80", "quantile", {"q": 0.8}),
                            ("mean", "mean", None), ("nonzero", "nonzero", None)):
        eng = tnx.TransferEntropyEngine(g["data"], names)
        r = eng.compute(accelerator="gpu", binarization=method, binarization_kwargs=kw)
        assert np.array_equal(r.bin_arrs[:, :, 0], g[f"bins_{tag}"]), tag
        assert np.array_equal(r.n_per_var, g[f"npv_{tag}"]), tag
        assert np.abs(r.matrix - g[f"matrix_{tag}"]).max() < TOL, tag
        assert r.kernel == "GEMM-B2"
        d = g[f"bins_{tag}"].astype(np.int32)
        assert np.abs(r.matrix - O.te_matrix(d, d.max(axis=1) + 1)).max() < TOL
        rs = tnx.TransferEntropyEngine(g["data"], names, sources=names[:5]).compute(
            accelerator="gpu", binarization=method, binarization_kwargs=kw)
        assert rs.kernel != "GEMM-B2"
        assert np.abs(rs.matrix[:5] - g[f"matrix_{tag}"][:5]).max() < TOL
