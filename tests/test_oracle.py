"""Pin the CPU oracle (runs without a GPU).

Each check is chosen so that a plausible mistake in the oracle (a dropped
term, a wrong sign or index, a transposed operand) fails it: numpy's own
arithmetic, two independent FP64 forms, closed forms, invariances and a
pure-Python brute force on tiny inputs.
"""
import math
from collections import Counter

import numpy as np
import pytest
import torch

from oracle import oracle as O
from oracle import scatter_add_ref as S
from tenex_amd import synthetic


# ── FSBW-L and remap against numpy itself ─────────────────────────────────

@pytest.mark.parametrize("T", [2, 7, 8, 9, 127, 128, 129, 500, 3000, 8191, 8192, 8193, 9999, 16500])
def test_fsbw_l_matches_numpy(T):
    x = synthetic.zero_inflated_lognormal(25, T, seed=T)
    got = O.fsbw_l(x)
    exp = O.fsbw_l_numpy(x)
    for g, e in zip(got, exp):
        assert np.array_equal(g, e)


def test_fsbw_l_constant_and_kp():
    x = synthetic.zero_inflated_lognormal(6, 300, seed=1)
    x[0] = 3.0          # std == 0 -> 1
    x[1, :] = 0.0
    for kp in (0.5, 0.3, 1.0):
        got = O.fsbw_l(x, kp)
        exp = O.fsbw_l_numpy(x, kp)
        for g, e in zip(got, exp):
            assert np.array_equal(g, e)
    assert np.all(O.fsbw_l(x)[0][0] == 0)


def test_np_rowsum_matches_numpy_sum():
    rng = np.random.default_rng(0)
    for n in (1, 5, 8, 100, 129, 1000, 8192, 8193, 20000):
        a = rng.standard_normal(n).astype(np.float32) * 1e3
        assert O.np_rowsum_f32(a) == a.reshape(1, -1).sum(axis=1)[0]


def test_remap_matches_unique():
    rng = np.random.default_rng(2)
    raw = rng.integers(0, 40, size=(30, 200)).astype(np.int32)
    raw[3] = 17
    a = O.remap(raw)
    b = O.remap_numpy(raw)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]


# ── TE: brute force, two forms, closed forms ──────────────────────────────

def _brute_te(x, y, tau=1):
    """Pure-Python TE(y -> x) from dictionaries of counts (plug-in, bits)."""
    L = len(x) - tau
    c3 = Counter((x[t + tau], x[t], y[t]) for t in range(L))
    c2a = Counter((x[t + tau], x[t]) for t in range(L))
    c2b = Counter((x[t], y[t]) for t in range(L))
    c1 = Counter(x[t] for t in range(L))
    te = 0.0
    for (a, b, c), n in c3.items():
        te += n / L * math.log2(n * c1[b] / (c2a[(a, b)] * c2b[(b, c)]))
    return te


def test_brute_force_tiny():
    rng = np.random.default_rng(3)
    for trial in range(20):
        n, T = 3, int(rng.integers(3, 9))
        raw = rng.integers(0, 3, size=(n, T)).astype(np.int32)
        d, npv, _ = O.remap(raw)
        pairs = O.pairfree_pairs(n)
        for tau in (1, 2):
            if tau >= T:
                continue
            te_s, te_d = O.te_pairs(d, npv, pairs, tau=tau, direct=True)
            for p, (t, s) in enumerate(pairs):
                bt = _brute_te(list(d[t]), list(d[s]), tau)
                assert abs(te_s[p] - bt) < 1e-12
                assert abs(te_d[p] - bt) < 1e-12


def test_counts_sum_and_marginals():
    b = synthetic.mixed_cardinality_bins(8, 400, seed=4).astype(np.int32)
    npv = b.max(axis=1) + 1
    c = O.joint_counts(b, npv, 2, 5)
    assert c.sum() == 399
    # marginal over (a, c) is the histogram of x_t
    assert np.array_equal(c.sum(axis=(0, 2)), np.bincount(b[2, :-1], minlength=npv[2]))
    assert np.array_equal(c.sum(axis=1).sum(axis=0), np.bincount(b[5, :-1], minlength=npv[5]))


def test_two_forms_and_bounds_c0():
    x = synthetic.zero_inflated_lognormal(60, 500, seed=0)
    d, nb, npv, bm = O.discretize(x)
    pairs = O.pairfree_pairs(60)
    te_s, te_d = O.te_pairs(d, npv, pairs, direct=True)
    assert np.abs(te_s - te_d).max() < 1e-12
    assert te_s.min() > -1e-12
    assert np.all(te_s <= np.log2(npv[pairs[:, 0]]) + 1e-12)


def test_constant_source_and_invariances():
    rng = np.random.default_rng(5)
    d = rng.integers(0, 4, size=(5, 300)).astype(np.int32)
    d[4] = 0
    d, npv, _ = O.remap(d)
    pairs = O.pairfree_pairs(5)
    te = O.te_pairs(d, npv, pairs)
    assert np.all(np.abs(te[pairs[:, 1] == 4]) < 1e-14)
    # injective relabelling of one variable leaves every TE unchanged
    d2 = d.copy()
    d2[1] = (npv[1] - 1) - d[1]
    assert np.abs(O.te_pairs(d2, npv, pairs) - te).max() < 1e-12
    # gene permutation permutes the matrix
    perm = np.array([3, 0, 4, 1, 2])
    M = O.te_matrix(d, npv)
    Mp = O.te_matrix(d[perm], npv[perm])
    assert np.abs(Mp - M[np.ix_(perm, perm)]).max() < 1e-12


def test_planted_copy_closed_form():
    rng = np.random.default_rng(6)
    T = 2000
    x = rng.integers(0, 3, size=T)
    y = np.empty(T, np.int64)
    y[0] = 0
    y[1:] = x[:-1]
    d = np.stack([x, y]).astype(np.int32)
    d, npv, _ = O.remap(d)
    te = O.te_pairs(d, npv, np.array([[1, 0]]))[0]  # TE(x -> y)
    s1, s2a = O.gene_entropies(d, npv)
    assert abs(te - (s1[1] - s2a[1]) / (T - 1)) < 1e-12


def test_gemm_b2_closed_form_matches_counts_form():
    x, planted = synthetic.bernoulli_planted(40, 600, n_planted=5, seed=8)
    d, nb, npv, bm = O.discretize(x)
    assert bm == 2
    te, C, D, E = O.te_b2(d, counts=True)
    p = d[:, :-1].astype(np.int64)
    f = d[:, 1:].astype(np.int64)
    assert np.array_equal(C, (f * p) @ p.T)
    assert np.array_equal(D, p @ p.T)
    assert np.array_equal(E, f @ p.T)
    M = O.te_matrix(d, npv)           # public layout: M[j, i] = TE(j -> i)
    assert np.abs(te - M.T).max() < 1e-12


def test_scatter_add_restatement_close_to_fp64():
    x = synthetic.zero_inflated_lognormal(40, 500, seed=0)
    d, nb, npv, bm = O.discretize(x)
    pairs = O.build_pairs_engine(40)
    te32 = S.te_pairs(d, bm, pairs)
    te64 = O.te_pairs(d, npv, pairs)
    assert te32.dtype == np.float32
    assert np.abs(te32 - te64).max() < 1e-6


# ── surrogates, BH ────────────────────────────────────────────────────────

def test_block_perm_schedule():
    perm, eff, nb = O.block_perm(7, 300, "block", 50, 0, 3, 42)
    assert eff == 50 and nb == 6 and perm.shape == (3, 7, 6)
    for k in range(3):
        for v in range(7):
            assert sorted(perm[k, v]) == list(range(6))
    g = torch.Generator().manual_seed(43)
    exp = torch.argsort(torch.rand(7, 6, generator=g), dim=1).numpy()
    assert np.array_equal(perm[1], exp)
    r, eff, nb = O.block_perm(4, 30, "block", 40, 0, 1, 1)
    assert eff == 1 and nb == 30


def test_surrogate_oracle_identity_and_stats():
    x = synthetic.zero_inflated_lognormal(10, 200, seed=9)
    d, nb, npv, bm = O.discretize(x)
    pairs = O.pairfree_pairs(10)
    obs = O.te_pairs(d, npv, pairs).astype(np.float32)
    ident = np.zeros((4, 10, 1), np.int32)
    s, s2, cg, tk = O.surrogate_stats(d, npv, ident, 200, pairs, obs, return_te_k=True)
    assert np.abs(tk - obs[:, None]).max() < 1e-7
    perm, eff, nbk = O.block_perm(10, 200, "block", 20, 0, 5, 42)
    s, s2, cg, tk = O.surrogate_stats(d, npv, perm, eff, pairs, obs, return_te_k=True)
    assert np.allclose(s, tk.sum(1)) and np.allclose(s2, (tk ** 2).sum(1))
    assert np.array_equal(cg, (tk.astype(np.float32) >= obs[:, None]).sum(1))
    # the shuffled target's marginal histogram is unchanged over the head
    assert np.all(tk > -1e-12)


def test_bh_restatement_matches_torch():
    from tenex_amd.inference import bh_adjust_np, bh_fdr_gpu
    rng = np.random.default_rng(10)
    p = rng.random(5000).astype(np.float32) ** 3
    p[:50] = p[0]  # ties
    a = O.bh_fdr(p)
    b = bh_fdr_gpu(torch.from_numpy(p), 0.05).numpy()
    assert np.array_equal(a, b)
    c = bh_adjust_np(p.astype(np.float64))
    assert np.abs(c - a).max() < 1e-5


# ── MLP oracle: finite differences ────────────────────────────────────────

def test_mlp_oracle_gradient():
    D, H, Oo = 6, 64, 2
    rng = np.random.default_rng(11)
    P = O.mlp_n_params(D, H, Oo)
    params = rng.standard_normal(P) * 0.2
    X = rng.standard_normal((17, D))
    Tg = rng.standard_normal((17, Oo))
    w = np.array([1.0, 0.5])
    loss, g = O.mlp_loss_grad(params, D, H, Oo, X, Tg, w)
    Y = O.mlp_forward(params, D, H, Oo, X)
    assert abs(loss - np.mean(0.5 * w * (Y - Tg) ** 2)) < 1e-12
    idx = rng.choice(P, 40, replace=False)
    eps = 1e-6
    for i in idx:
        pp = params.copy(); pp[i] += eps
        lp, _ = O.mlp_loss_grad(pp, D, H, Oo, X, Tg, w)
        pm = params.copy(); pm[i] -= eps
        lm, _ = O.mlp_loss_grad(pm, D, H, Oo, X, Tg, w)
        assert abs((lp - lm) / (2 * eps) - g[i]) < 1e-6 * max(1.0, abs(g[i]))


def test_coarsen_restatement_properties():
    """Coarsening maps the levels of a variable with ng > limit onto every
    level 0..limit-1 (the surjectivity the reference's docstring proves),
    monotonically, merging ceil/floor(ng / limit) old levels per new level;
    variables at or below the limit are unchanged."""
    rng = np.random.default_rng(0)
    for limit in (2, 3, 7, 12, 33):
        ng = np.arange(2, 90)
        dense = np.tile(np.arange(90), (len(ng), 1)) % ng[:, None]
        d2, npv2, bm = O.coarsen_bins(dense, ng, limit)
        for g, b in enumerate(ng):
            old = np.arange(b)
            new = d2[g, :b]
            if b <= limit:
                assert np.array_equal(new, old) and npv2[g] == b
                continue
            assert npv2[g] == limit
            assert np.array_equal(np.unique(new), np.arange(limit))
            assert np.all(np.diff(new) >= 0)
            sizes = np.bincount(new, minlength=limit)
            assert sizes.min() >= b // limit and sizes.max() <= -(-b // limit)
        assert bm == max(min(limit, int(ng.max())), 2)
    # the joint counts of coarsened data are the sums of the original counts
    # over the merged levels (brute force on a small case)
    d = np.stack([rng.integers(0, 9, 400), rng.integers(0, 5, 400)]).astype(np.int32)
    npv = np.array([9, 5], np.int32)
    c = O.joint_counts(d, npv, 0, 1)
    dc, npc, _ = O.coarsen_bins(d, npv, 4)
    cc = O.joint_counts(dc, npc, 0, 1)
    m, ms = np.arange(9) * 4 // 9, np.arange(5) * 4 // 5
    agg = np.zeros_like(cc)
    for a in range(9):
        for b in range(9):
            for s in range(5):
                agg[m[a], m[b], ms[s]] += c[a, b, s]
    assert np.array_equal(agg, cc)


def test_smem_bin_limit_brute_force():
    for optin in (48 * 1024, 64 * 1024, 99 * 1024, 160 * 1024, 228 * 1024):
        for warps in (4, 16, 32):
            best = max(B for B in range(2, 200)
                       if (B ** 3 + 2 * B ** 2 + B + warps) * 4 <= optin)
            assert O.smem_bin_limit(optin, warps) == best
    assert O.smem_bin_limit(160 * 1024, 32) == 33


def test_sparse_te_oracle_equals_dense():
    """te_pairs_sparse (occupied states) equals te_pairs (dense counts)."""
    rng = np.random.default_rng(4)
    b = np.stack([rng.integers(0, k, 700) for k in (2, 5, 9, 17, 3, 30)]).astype(np.int32)
    npv = (b.max(axis=1) + 1).astype(np.int32)
    pairs = O.pairfree_pairs(6)
    for tau in (1, 3):
        d = O.te_pairs(b, npv, pairs, tau=tau)
        sp = O.te_pairs_sparse(b, pairs, tau=tau)
        assert np.abs(d - sp).max() < 1e-12
