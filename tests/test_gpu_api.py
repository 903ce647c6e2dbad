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
