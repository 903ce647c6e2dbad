"""The tnx-style API on the GPU: engine, kernel selection, NetWeaver, Pipeline."""
import numpy as np
import pytest

import tenex_amd as tnx
from oracle import oracle as O
from tenex_amd import synthetic

pytestmark = pytest.mark.gpu
TOL = 1e-6


def test_engine_c0_matches_oracle():
    x = synthetic.zero_inflated_lognormal(100, 500, seed=0)
    names = np.array([f"g{i}" for i in range(100)])
    res = tnx.TransferEntropyEngine(x, names).compute(accelerator="gpu", profile=True)
    d, nb, npv, bm = O.discretize(x)
    ref = O.te_matrix(d, npv)
    assert res.matrix.dtype == np.float32 and res.shape == (100, 100)
    assert np.abs(res.matrix - ref).max() < TOL
    assert res.kernel == "Full-SMEM" and res.b_max == bm
    assert np.array_equal(res.bin_arrs[:, :, 0].astype(np.int32), d)
    assert np.array_equal(res.n_per_var, npv)
    assert set(res.timings) >= {"total", "discretize", "compute"}


def test_engine_sources_and_kernels():
    x = synthetic.zero_inflated_lognormal(30, 300, seed=1)
    names = np.array([f"g{i}" for i in range(30)])
    src = names[[2, 5, 11]]
    res = tnx.TransferEntropyEngine(x, names, sources=src).compute(accelerator="gpu")
    d, nb, npv, bm = O.discretize(x)
    ref = O.te_matrix(d, npv)
    rows = [2, 5, 11]
    assert np.abs(res.matrix[rows] - ref[rows]).max() < TOL
    other = np.setdiff1d(np.arange(30), rows)
    assert np.all(res.matrix[other] == 0)
    r2 = tnx.TransferEntropyEngine(x, names).compute(accelerator="gpu", kernel="Adaptive-SMEM")
    assert r2.kernel == "Adaptive-SMEM" and np.abs(r2.matrix - ref).max() < TOL
    with pytest.raises(ValueError):
        tnx.TransferEntropyEngine(x, names).compute(accelerator="gpu", kernel="GEMM-B2")


def test_engine_autotune_matches_oracle():
    """compute(autotune=True) times the launch candidates (LDS histogram
    sizes, key-0 diversion) on a sample of targets, records its pick, and
    whatever it picks the matrix stays within 1e-6 of the FP64 oracle."""
    from tenex_amd import core
    b = synthetic.mixed_cardinality_bins(150, 2000, seed=13)
    names = np.array([f"g{i}" for i in range(150)])
    eng = tnx.TransferEntropyEngine(b.astype(np.float32), names)
    res = eng.compute(accelerator="gpu", autotune=True)
    info = eng._last_run_info["autotune"]
    assert (info["cap_wave"], info["cap_block"]) in core.TUNE_CAPS
    assert info["ms"] <= info["default_ms"]
    d, nb, npv, bm = O.discretize(b.astype(np.float32))
    ref = O.te_matrix(d, npv)
    assert np.abs(res.matrix - ref).max() < TOL


def test_engine_binary_routes_to_gemm_b2():
    x, planted = synthetic.bernoulli_planted(200, 800, n_planted=20, seed=3)
    names = np.array([f"g{i}" for i in range(200)])
    res = tnx.TransferEntropyEngine(x, names).compute(accelerator="gpu")
    assert res.kernel == "GEMM-B2" and res.b_max == 2
    d, nb, npv, bm = O.discretize(x)
    assert np.abs(res.matrix - O.te_matrix(d, npv)).max() < TOL
    r2 = tnx.TransferEntropyEngine(x, names).compute(accelerator="gpu", kernel="Full-SMEM")
    assert np.abs(r2.matrix - res.matrix).max() < TOL


def test_netweaver_fdr_and_pipeline():
    x = synthetic.zero_inflated_lognormal(50, 400, seed=4)
    names = np.array([f"g{i}" for i in range(50)])
    res = tnx.TransferEntropyEngine(x, names).compute(accelerator="gpu")
    grn, trimmed = tnx.NetWeaver(res, fdr=0.05).infer(method="fdr")
    assert len(trimmed) <= len(grn)
    pipe = tnx.Pipeline(fdr=0.05).configure(tau=1)
    pr = pipe.run(data=x, variable_names=names, methods=["fdr", "surrogate_test"],
                  method_kwargs={"surrogate_test": {"n_surrogates": 10}})
    assert np.array_equal(pr.matrix, res.matrix)
    assert pr.surrogate_test.observed_te.shape == (50, 50)
    ml = pipe.infer("surrogate_test", taus=[1, 2], n_surrogates=5)
    assert ml.p_values.shape == (50, 50)


@pytest.mark.parametrize("smem_optin,limit", [(160 * 1024, 33), (16 * 1024, 15)])
def test_adaptive_coarsening_matches_oracle(dev, smem_optin, limit):
    """Adaptive-SMEM prepare(use_coarsening=True) coarsens exactly as the
    reference's coarsen_bins (tenex/preprocess.py:462-551) with the limit of
    compute_smem_bin_limit: bins and n_per_var bit-exact, TE of the coarsened
    data within 1e-6 of the FP64 oracle; None/False keep the exact bins."""
    from tenex_amd import core
    from tenex_amd.kernels.lds_hist import AdaptiveSMEMKernel
    from tenex_amd.preprocess import coarsen_bins
    d = synthetic.mixed_cardinality_bins(40, 2000, seed=6).astype(np.int32)
    bins, npv, b_max, T = core.upload_bins(d, dev)
    prep = core.prepare(bins, npv, T, 1)
    assert b_max > limit
    k = AdaptiveSMEMKernel()
    for flag in (None, False):
        p0, n0, b0 = k.prepare(prep, npv, b_max, use_coarsening=flag, smem_optin=smem_optin)
        assert p0 is prep and b0 == b_max
    p2, npv2, b2 = k.prepare(prep, npv, b_max, use_coarsening=True, smem_optin=smem_optin)
    assert O.smem_bin_limit(smem_optin, 32) == limit
    dc, npc, bmc = O.coarsen_bins(d, prep.npv_host, limit)
    assert b2 == bmc == limit
    assert np.array_equal(p2.bins[:, :T].cpu().numpy().astype(np.int32), dc)
    assert np.array_equal(npv2.cpu().numpy(), npc)
    assert np.array_equal(prep.bins[:, :T].cpu().numpy().astype(np.int32), d)  # input intact
    m = core.te_matrix_internal(p2).cpu().numpy()
    pairs = O.pairfree_pairs(40)
    te64 = O.te_pairs(dc, npc, pairs)
    assert np.abs(m[pairs[:, 0], pairs[:, 1]].astype(np.float64) - te64).max() < TOL
    # the public preprocess.coarsen_bins on a dense array
    b3, n3, bm3 = coarsen_bins(d[:, :, None], prep.npv_host, limit)
    assert np.array_equal(b3[:, :, 0].cpu().numpy().astype(np.int32), dc) and bm3 == limit
    assert np.array_equal(n3.cpu().numpy(), npc)


def test_engine_coarsening_flag():
    """compute(coarsening=True) on data within the LDS bin limit leaves the
    bins exact and reports coarsened=False (the reference's behaviour below
    the limit)."""
    x = synthetic.zero_inflated_lognormal(30, 600, seed=2)
    names = np.array([f"g{i}" for i in range(30)])
    eng = tnx.TransferEntropyEngine(x, names)
    res = eng.compute(accelerator="gpu", kernel="Adaptive-SMEM", coarsening=True)
    d, nb, npv, bm = O.discretize(x)
    assert np.abs(res.matrix - O.te_matrix(d, npv)).max() < TOL
    assert eng._last_run_info["coarsened"] is False


def test_compute_single_gpu_sparse_pair_list(dev):
    """An explicit sparse pair list (duplicates, target == source, few of the
    variables) costs its own pairs: the pair-list plan holds exactly the
    distinct valid pairs, and every TE is within 1e-6 of the FP64 oracle."""
    from tenex_amd import core
    from tenex_amd.kernels.lds_hist import AdaptiveSMEMKernel
    n, T = 300, 1500
    d = synthetic.mixed_cardinality_bins(n, T, seed=12).astype(np.int32)
    rng = np.random.default_rng(12)
    few = rng.choice(n, 60, replace=False)
    pairs = few[rng.integers(0, 60, size=(900, 2))]
    pairs = np.concatenate([pairs, pairs[:40], [[few[0], few[0]]]])
    k = AdaptiveSMEMKernel()
    npv = d.max(axis=1) + 1
    got = k.compute_single_gpu(d[:, :, None], pairs, int(npv.max()), npv, 1, 1024, dev)
    exp = np.zeros(len(pairs))
    v = pairs[:, 0] != pairs[:, 1]
    exp[v] = O.te_pairs(d, npv, pairs[v])
    assert got.dtype == np.float32 and np.abs(got - exp).max() < TOL
    # all variables touched: the direct path over the full prepared data
    full = np.stack([rng.integers(0, n, 5000), rng.integers(0, n, 5000)], 1)
    full[:n, 0] = np.arange(n)
    got2 = k.compute_single_gpu(d[:, :, None], full, int(npv.max()), npv, 1, 1024, dev)
    v2 = full[:, 0] != full[:, 1]
    exp2 = np.zeros(len(full))
    exp2[v2] = O.te_pairs(d, npv, full[v2])
    assert np.abs(got2 - exp2).max() < TOL
    # the plan costs the distinct valid pairs, not the targets x sources grid
    involved = np.unique(pairs[v])
    loc = np.searchsorted(involved, pairs[v])
    plan = core.te_pairlist_plan(npv[involved].astype(np.int32), np.unique(loc, axis=0), dev,
                                 core.WAVE_CAP)
    x, y = core.plan_pairs(plan)
    assert len(x) == len(np.unique(loc, axis=0))
