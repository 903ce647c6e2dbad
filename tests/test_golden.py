"""Parity against fixtures written by the reference implementation itself
(tests/golden/make_golden.py, CPU scatter_add path and CPU surrogate loop)."""
import os

import numpy as np
import pytest
import torch

from oracle import oracle as O
from oracle import scatter_add_ref as S

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return dict(np.load(os.path.join(G, name), allow_pickle=False))


def test_golden_c0_bins_and_te_oracle():
    g = load("te_c0_100x500.npz")
    d, nb, npv, bm = O.discretize(g["data"])
    assert np.array_equal(d.astype(np.int8), g["bins"])
    assert np.array_equal(npv, g["n_per_var"]) and bm == int(g["b_max"])
    M = O.te_matrix(d, npv)
    assert np.abs(M - g["matrix"]).max() < 1e-6
    assert str(g["kernel"]) == "scatter_add"


def test_golden_c0_scatter_add_restatement():
    """The torch-CPU restatement reproduces the reference's matrix.  torch's
    CPU float32 sums split work by thread count and SIMD width, so the
    comparison is bit-exact only on the host that wrote the fixture; across
    hosts it holds to float32 rounding."""
    g = load("te_c0_100x500.npz")
    d = g["bins"].astype(np.int32)
    pairs = O.build_pairs_engine(100)
    te = S.te_pairs(d, int(g["b_max"]), pairs)
    M = np.zeros((100, 100), np.float32)
    M[pairs[:, 0], pairs[:, 1]] = te
    assert np.abs(M.T - g["matrix"]).max() < 1e-6


def test_golden_surrogate_oracle():
    g = load("surrogate_40x400.npz")
    d = g["bins"].astype(np.int32)
    npv = g["n_per_var"]
    n = d.shape[0]
    obs = g["observed"]
    ns, bl, seed = int(g["n_surrogates"]), int(g["block_length"]), int(g["seed"])
    perm, eff, nbk = O.block_perm(n, d.shape[1], "block", bl, 0, ns, seed, device="cpu")
    pairs = O.pairfree_pairs(n)
    s, s2, cg = O.surrogate_stats(d, npv, perm, eff, pairs, obs[pairs[:, 1], pairs[:, 0]])
    Sm = np.zeros((n, n)); S2 = np.zeros((n, n)); CG = np.zeros((n, n), np.int32)
    Sm[pairs[:, 1], pairs[:, 0]] = s
    S2[pairs[:, 1], pairs[:, 0]] = s2
    CG[pairs[:, 1], pairs[:, 0]] = cg
    dec = O.surrogate_decision(obs, Sm, S2, CG, ns, fdr=float(g["fdr"]))
    off = ~np.eye(n, dtype=bool)
    assert np.abs(dec["mean"] - g["mean"])[off].max() < 1e-6
    assert np.abs(dec["std"] - g["std"])[off].max() < 1e-5
    assert np.abs(dec["p"] - g["p"])[off].max() < 1e-3
    got = set(map(tuple, np.argwhere(dec["sig"]).tolist()))
    exp = set(map(tuple, g["edges"].tolist()))
    assert got == exp


@pytest.mark.gpu
def test_golden_c0_gpu_engine():
    import tenex_amd as tnx
    g = load("te_c0_100x500.npz")
    names = np.array([f"g{i}" for i in range(100)])
    res = tnx.TransferEntropyEngine(g["data"], names).compute(accelerator="gpu")
    assert np.array_equal(res.bin_arrs[:, :, 0], g["bins"])
    assert np.array_equal(res.n_per_var, g["n_per_var"])
    assert np.abs(res.matrix - g["matrix"]).max() < 1e-6


@pytest.mark.gpu
def test_golden_surrogate_gpu():
    import tenex_amd as tnx
    g = load("surrogate_40x400.npz")
    n = g["bins"].shape[0]
    names = np.array([f"g{i}" for i in range(n)])
    res = tnx.TransferEntropyEngine(g["data"], names).compute(accelerator="gpu")
    assert np.abs(res.matrix - g["observed"]).max() < 1e-6
    st = tnx.NetWeaver(res, fdr=float(g["fdr"])).infer(
        method="surrogate_test", n_surrogates=int(g["n_surrogates"]),
        block_length=int(g["block_length"]), seed=int(g["seed"]), rng_device="cpu")
    off = ~np.eye(n, dtype=bool)
    assert np.abs(st.mean_surrogate_te - g["mean"])[off].max() < 1e-6
    assert np.abs(st.std_surrogate_te - g["std"])[off].max() < 1e-5
    assert np.abs(st.p_values - g["p"])[off].max() < 1e-3
    got = set(map(tuple, st.grn.pairs.tolist()))
    exp = set(map(tuple, g["edges"].tolist()))
    assert got == exp


def _fdr_inputs():
    from tenex_amd.inference import build_pairs
    g = load("fdr_300x600.npz")
    names = np.array([f"g{i}" for i in range(g["matrix"].shape[0])])
    pairs = build_pairs(names)
    return g, names, pairs, g["matrix"][pairs[:, 0], pairs[:, 1]]


def test_golden_fdr_oracle():
    """The oracle's threshold and DPI trim reproduce the reference's
    fdr_threshold_gpu and _trim outputs bit for bit (46 edges at FDR 0.01,
    8 indirect edges of planted chains removed)."""
    g, names, pairs, te = _fdr_inputs()
    tc, pc = O.fdr_threshold(te, pairs, float(g["fdr"]))
    assert np.array_equal(pc, g["pairs_cut"]) and np.array_equal(tc, g["te_cut"])
    A = np.zeros_like(g["matrix"])
    A[pc[:, 0], pc[:, 1]] = tc
    tp, tv = O.trim_dpi(A)
    assert np.array_equal(tp, g["trimmed_pairs"]) and np.array_equal(tv, g["trimmed_te"])
    assert len(g["pairs_cut"]) - len(g["trimmed_pairs"]) == 8


def test_golden_fdr_host_on_cpu_tensors():
    """The package's device FDR arithmetic and DPI trim, run on CPU tensors,
    give the reference's edges and weights bit for bit."""
    from tenex_amd.inference import fdr_threshold_gpu
    from tenex_amd.inference.fdr import trim_dpi
    g, names, pairs, te = _fdr_inputs()
    tc, pc = fdr_threshold_gpu(torch.from_numpy(te), torch.from_numpy(pairs).long(),
                               float(g["fdr"]))
    assert np.array_equal(pc.numpy(), g["pairs_cut"]) and np.array_equal(tc.numpy(), g["te_cut"])
    tr = trim_dpi(g["matrix"], names, pc.numpy(), tc.numpy(), "cpu")
    assert np.array_equal(tr.pairs, g["trimmed_pairs"])
    assert np.array_equal(tr.te, g["trimmed_te"])


@pytest.mark.gpu
def test_golden_fdr_netweaver_gpu():
    """NetWeaver((matrix, names)).infer("fdr") on the GPU (device z score,
    erfc, BH; the HIP max-min DPI kernel): thresholded and trimmed edge
    lists and weights identical to the reference's."""
    import tenex_amd as tnx
    g, names, pairs, te = _fdr_inputs()
    grn, trimmed = tnx.NetWeaver((g["matrix"], names), fdr=float(g["fdr"])).infer("fdr")
    assert np.array_equal(np.asarray(grn.pairs, np.int64), g["pairs_cut"])
    assert np.array_equal(grn.te, g["te_cut"])
    assert np.array_equal(np.asarray(trimmed.pairs, np.int64), g["trimmed_pairs"])
    assert np.array_equal(trimmed.te, g["trimmed_te"])
    assert list(trimmed.source) == [f"g{i}" for i in g["trimmed_pairs"][:, 0]]


def test_golden_surrogate_120_oracle():
    """120 genes x 1,000 cells, 50 block shuffles of length 40: the oracle's
    shuffle-and-count and decision reproduce the reference's statistics and
    its 19 edges."""
    g = load("surrogate_120x1000.npz")
    d = g["bins"].astype(np.int32)
    n = d.shape[0]
    d2, nb, npv, bm = O.discretize(g["data"])
    assert np.array_equal(d2, d) and np.array_equal(npv, g["n_per_var"])
    obs = O.te_matrix(d, npv)                         # public [source, target]
    assert np.abs(obs - g["observed"]).max() < 1e-6
    obs = g["observed"]
    ns, bl, seed = int(g["n_surrogates"]), int(g["block_length"]), int(g["seed"])
    perm, eff, nbk = O.block_perm(n, d.shape[1], "block", bl, 0, ns, seed, device="cpu")
    pairs = O.pairfree_pairs(n)
    s, s2, cg = O.surrogate_stats(d, npv, perm, eff, pairs, obs[pairs[:, 1], pairs[:, 0]])
    Sm = np.zeros((n, n)); S2 = np.zeros((n, n)); CG = np.zeros((n, n), np.int32)
    Sm[pairs[:, 1], pairs[:, 0]] = s
    S2[pairs[:, 1], pairs[:, 0]] = s2
    CG[pairs[:, 1], pairs[:, 0]] = cg
    dec = O.surrogate_decision(obs, Sm, S2, CG, ns, fdr=float(g["fdr"]))
    off = ~np.eye(n, dtype=bool)
    assert np.abs(dec["mean"] - g["mean"])[off].max() < 1e-6
    assert np.abs(dec["std"] - g["std"])[off].max() < 1e-5
    assert np.abs(dec["p"] - g["p"])[off].max() < 1e-3
    got = set(map(tuple, np.argwhere(dec["sig"]).tolist()))
    exp = set(map(tuple, g["edges"].tolist()))
    assert got == exp and len(exp) == 19


@pytest.mark.gpu
def test_golden_surrogate_120_gpu():
    import tenex_amd as tnx
    g = load("surrogate_120x1000.npz")
    n = g["bins"].shape[0]
    names = np.array([f"g{i}" for i in range(n)])
    res = tnx.TransferEntropyEngine(g["data"], names).compute(accelerator="gpu")
    assert np.abs(res.matrix - g["observed"]).max() < 1e-6
    for fused in (False, True):
        st = tnx.NetWeaver(res, fdr=float(g["fdr"])).infer(
            method="surrogate_test", n_surrogates=int(g["n_surrogates"]),
            block_length=int(g["block_length"]), seed=int(g["seed"]), rng_device="cpu",
            fused=fused)
        off = ~np.eye(n, dtype=bool)
        assert np.abs(st.mean_surrogate_te - g["mean"])[off].max() < 1e-6
        assert np.abs(st.std_surrogate_te - g["std"])[off].max() < 1e-5
        assert np.abs(st.p_values - g["p"])[off].max() < 1e-3
        got = set(map(tuple, st.grn.pairs.tolist()))
        exp = set(map(tuple, g["edges"].tolist()))
        assert got == exp, fused
