"""Multi-lag surrogate test (tenex/pipeline.py:252-409): per-lag statistics
with every lag in each launch, combined per pair.  The oracle's per-lag
statistics through the package's combine reproduce the reference's CPU
pipeline; on the GPU the batched lags equal the lag-by-lag run and the
reference's result."""
import os

import numpy as np
import pytest
import torch

from oracle import oracle as O

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = 1e-6


def load(name):
    return dict(np.load(os.path.join(G, name), allow_pickle=False))


class _R:
    """Per-lag result fields the combine reads."""

    def __init__(self, **kw):
        self.__dict__.update(kw)


def _oracle_lag(d, npv, tau, ns, seed, fdr):
    n, T = d.shape
    bl = max(2, int(np.sqrt(T - tau)))
    obs = O.te_matrix(d, npv, tau)                       # public [source, target]
    perm, eff, nbk = O.block_perm(n, T, "block", bl, 0, ns, seed, device="cpu")
    pairs = O.pairfree_pairs(n)
    s, s2, cg = O.surrogate_stats(d, npv, perm, eff, pairs, obs[pairs[:, 1], pairs[:, 0]],
                                  tau=tau)
    Sm = np.zeros((n, n)); S2 = np.zeros((n, n)); CG = np.zeros((n, n), np.int32)
    Sm[pairs[:, 1], pairs[:, 0]] = s
    S2[pairs[:, 1], pairs[:, 0]] = s2
    CG[pairs[:, 1], pairs[:, 0]] = cg
    dec = O.surrogate_decision(obs, Sm, S2, CG, ns, fdr=fdr)
    return _R(observed_te=obs.astype(np.float32), effective_te=dec["eff"],
              mean_surrogate_te=dec["mean"], std_surrogate_te=dec["std"], p_values=dec["p"],
              fdr=fdr, n_surrogates=ns, shuffle_method="block", block_length=bl,
              p_method="parametric")


def test_golden_multilag_oracle_combine():
    """Oracle per-lag statistics (each lag's own block length and orders)
    combined by the package's combine_surrogate_multi_lag give the
    reference's combined statistics and edges for 'max' and
    'adaptive_select'."""
    from tenex_amd.pipeline import combine_surrogate_multi_lag
    g = load("multilag_50x500.npz")
    d, nb, npv, bm = O.discretize(g["data"])
    d = d.astype(np.int32)
    ns, seed, fdr = int(g["n_surrogates"]), int(g["seed"]), float(g["fdr"])
    taus = [int(t) for t in g["taus"]]
    per = [_oracle_lag(d, npv, t, ns, seed, fdr) for t in taus]
    names = np.array([f"g{i}" for i in range(d.shape[0])])
    r = combine_surrogate_multi_lag(per, taus, "max", names, torch.device("cpu"))
    off = ~np.eye(d.shape[0], dtype=bool)
    assert np.abs(r.observed_te - g["observed"])[off].max() < TOL
    assert np.abs(r.mean_surrogate_te - g["mean"])[off].max() < TOL
    assert np.abs(r.std_surrogate_te - g["std"])[off].max() < 1e-5
    assert np.abs(r.p_values - g["p"])[off].max() < 1e-3
    assert set(map(tuple, r.grn.pairs.tolist())) == set(map(tuple, g["edges"].tolist()))
    ra = combine_surrogate_multi_lag(per, taus, "adaptive_select", names, torch.device("cpu"))
    assert set(map(tuple, ra.grn.pairs.tolist())) == \
        set(map(tuple, g["edges_adaptive"].tolist()))


@pytest.mark.gpu
def test_batched_lags_equal_lag_by_lag(dev):
    """Pipeline.infer(taus=...) with every lag in shared launches gives the
    lag-by-lag run's combined statistics bit for bit (same tables: T - tau
    within one power of two) and its edges, for the three combines."""
    import tenex_amd as tnx
    from tenex_amd import synthetic
    x = synthetic.zero_inflated_lognormal(70, 900, seed=8)
    rng = np.random.default_rng(8)
    for s_, t_, lag in ((1, 2, 2), (3, 4, 3), (5, 6, 1)):
        x[t_, lag:] = x[s_, :-lag] * (0.5 + rng.random(900 - lag).astype(np.float32))
    names = np.array([f"g{i}" for i in range(70)])
    eng = tnx.TransferEntropyEngine(x, names)
    for combine in ("max", "adaptive_select", "weighted_sum"):
        kw = dict(taus=[1, 2, 3], lag_combine=combine, n_surrogates=24, seed=7)
        a = tnx.Pipeline(eng, fdr=0.05).infer("surrogate_test", **kw)
        b = tnx.Pipeline(eng, fdr=0.05).infer("surrogate_test", batched_lags=False, **kw)
        for f in ("observed_te", "mean_surrogate_te", "std_surrogate_te", "effective_te",
                  "p_values"):
            assert np.array_equal(getattr(a, f), getattr(b, f)), (combine, f)
        assert np.array_equal(a.grn.pairs, b.grn.pairs), combine
        assert len(a.grn) >= 3


@pytest.mark.gpu
def test_multilag_observed_vs_oracle(dev):
    """The observed TE of every lag from one launch (per-variable series
    lengths) against the FP64 oracle, including lags whose T - tau sit on
    either side of a power of two (different fixed-point tables)."""
    from tenex_amd import core, synthetic
    for n, T, taus in ((40, 1026, [1, 3, 9]), (30, 600, [2, 5])):
        b = synthetic.mixed_cardinality_bins(n, T, seed=T).astype(np.int32)
        bins, npv, _, _ = core.upload_bins(b, dev)
        prep = core.prepare(bins, npv, T, min(taus))
        plan = core.get_plan(prep)
        obs = core.multilag_observed(prep, plan, taus)
        assert obs is not None
        pairs = O.pairfree_pairs(n)
        for tau, m in zip(taus, obs):
            te64 = O.te_pairs(b, prep.npv_host, pairs, tau=tau)
            got = m.cpu().numpy()[pairs[:, 0], pairs[:, 1]].astype(np.float64)
            assert np.abs(got - te64).max() < TOL, (T, tau)
            single = core.te_matrix_internal(core.prepare(bins, npv, T, tau)).cpu().numpy()
            assert np.abs(single - m.cpu().numpy()).max() < 1e-6


@pytest.mark.gpu
def test_golden_multilag_gpu(dev):
    """The GPU pipeline against the reference's CPU multi-lag result."""
    import tenex_amd as tnx
    g = load("multilag_50x500.npz")
    names = np.array([f"g{i}" for i in range(50)])
    eng = tnx.TransferEntropyEngine(g["data"], names)
    off = ~np.eye(50, dtype=bool)
    for combine, key in (("max", "edges"), ("adaptive_select", "edges_adaptive")):
        r = tnx.Pipeline(eng, fdr=float(g["fdr"])).infer(
            "surrogate_test", taus=[int(t) for t in g["taus"]], lag_combine=combine,
            n_surrogates=int(g["n_surrogates"]), seed=int(g["seed"]), rng_device="cpu")
        if combine == "max":
            assert np.abs(r.observed_te - g["observed"])[off].max() < TOL
            assert np.abs(r.mean_surrogate_te - g["mean"])[off].max() < TOL
            assert np.abs(r.std_surrogate_te - g["std"])[off].max() < 1e-5
            assert np.abs(r.p_values - g["p"])[off].max() < 1e-3
        assert set(map(tuple, r.grn.pairs.tolist())) == set(map(tuple, g[key].tolist()))
