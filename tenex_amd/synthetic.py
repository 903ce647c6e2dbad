"""Seeded synthetic inputs for the benchmark / test workloads.

zero_inflated_lognormal  continuous expression: 30% zeros, LogNormal(1, 1)
bernoulli_planted        binary genes Bernoulli(p_g), p_g ~ U(0.1, 0.9), plus
                         planted lagged pairs y[t+1] = x[t] xor Bernoulli(flip)
mixed_cardinality_bins   dense bins with per-gene level counts drawn as 80% in
                         [2, 8], 18% in [9, 32], 2% in [33, 64]; skewed
                         categorical levels, every level present
"""
from __future__ import annotations

import numpy as np


def zero_inflated_lognormal(n: int, T: int, seed: int = 0, zero_frac: float = 0.3,
                            mu: float = 1.0, sigma: float = 1.0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    x = rng.lognormal(mu, sigma, size=(n, T)).astype(np.float32)
    x[rng.random((n, T)) < zero_frac] = 0.0
    return x


def bernoulli_planted(n: int, T: int, n_planted: int = 200, flip: float = 0.2,
                      seed: int = 0):
    """Returns (x float32 (n, T) in {0, 1}, planted int64 (n_planted, 2)
    rows (source x, target y) with y[t+1] = x[t] xor flip)."""
    rng = np.random.default_rng(seed)
    p = rng.uniform(0.1, 0.9, size=n)
    x = (rng.random((n, T)) < p[:, None]).astype(np.float32)
    perm = rng.permutation(n)
    planted = np.stack([perm[0:2 * n_planted:2], perm[1:2 * n_planted:2]], axis=1)
    for s, t in planted:
        fl = rng.random(T - 1) < flip
        x[t, 1:] = np.logical_xor(x[s, :-1] > 0, fl).astype(np.float32)
    return x, planted.astype(np.int64)


def mixed_cardinality_levels(n: int, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    u = rng.random(n)
    b = np.where(u < 0.80, rng.integers(2, 9, size=n),
                 np.where(u < 0.98, rng.integers(9, 33, size=n), rng.integers(33, 65, size=n)))
    return b.astype(np.int32)


def mixed_cardinality_bins(n: int, T: int, seed: int = 0) -> np.ndarray:
    """Dense int8 bins (n, T); gene g has exactly levels[g] levels, drawn
    from a skewed categorical (Dirichlet(0.7) weights), each level present."""
    rng = np.random.default_rng(seed)
    levels = mixed_cardinality_levels(n, seed)
    out = np.empty((n, T), np.int8)
    for g in range(n):
        b = int(levels[g])
        w = rng.dirichlet(np.full(b, 0.7))
        row = rng.choice(b, size=T, p=w).astype(np.int8)
        pos = rng.choice(T, size=b, replace=False)
        row[pos] = np.arange(b, dtype=np.int8)
        out[g] = row
    return out
