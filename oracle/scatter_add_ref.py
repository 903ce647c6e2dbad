"""torch-CPU restatement of the reference's scatter_add TE path: the CPU baseline.

TEST AND CPU-BASELINE INFRASTRUCTURE ONLY (bench.py cpu_baseline leg and tests).

Follows tenex/kernels/scatter_add.py:33-119 (_compute_te_batch) op for op:
int64 flat index xt1*B^2 + xt*B + yt + pair*B^3 with a global B = b_max,
int32 scatter_add_, float32 marginals, clamp(min=1) of the denominator,
log2, nan_to_num, sum / N; batches of 2048 pairs on the CPU
(scatter_add.py:309-311), pairs in _build_pairs order
(transferentropy.py:456-498).
"""
from __future__ import annotations

import time

import numpy as np
import torch

CPU_BATCH = 2048


def te_batch(bin_arrs: torch.Tensor, t_genes: torch.Tensor, s_genes: torch.Tensor,
             B: int, tau: int = 1) -> torch.Tensor:
    """One batch of pairs: bin_arrs (n, T, K) int32; returns (batch,) float32."""
    batch = t_genes.shape[0]
    T, K = bin_arrs.shape[1], bin_arrs.shape[2]
    L = T - tau
    N = L * K
    B2, B3 = B * B, B * B * B
    target = bin_arrs[t_genes]
    xt1 = target[:, tau:, :].to(torch.int64)
    xt = target[:, :-tau, :].to(torch.int64)
    del target
    xt1.mul_(B2)
    xt1.add_(xt.mul_(B))
    del xt
    source = bin_arrs[s_genes]
    yt = source[:, :-tau, :].to(torch.int64)
    del source
    xt1.add_(yt)
    del yt
    pair_off = torch.arange(batch, dtype=torch.int64)
    xt1.add_(pair_off.view(-1, 1, 1) * B3)
    idx = xt1.reshape(-1)
    del xt1
    ones = torch.ones(batch * N, dtype=torch.int32)
    cnt3_flat = torch.zeros(batch * B3, dtype=torch.int32)
    cnt3_flat.scatter_add_(0, idx, ones)
    del idx, ones
    c3f = cnt3_flat.view(batch, B, B, B).float()
    del cnt3_flat
    cnt2a = c3f.sum(dim=3)
    cnt2b = c3f.sum(dim=1)
    cnt1 = c3f.sum(dim=(1, 3))
    numer = c3f * cnt1[:, None, :, None]
    del cnt1
    denom = cnt2a[:, :, :, None] * cnt2b[:, None, :, :]
    del cnt2a, cnt2b
    denom.clamp_(min=1)
    numer.div_(denom)
    del denom
    numer.log2_()
    numer.mul_(c3f)
    numer.nan_to_num_(0.0)
    te = numer.sum(dim=(1, 2, 3)).div_(N)
    del numer, c3f
    return te


def te_pairs(bins: np.ndarray, b_max: int, pairs: np.ndarray, tau: int = 1,
             batch_size: int = CPU_BATCH) -> np.ndarray:
    """compute_te_scatter_add (scatter_add.py:325-372) on the CPU for an
    explicit [target, source] pair list.  Returns float32 (n_pairs,)."""
    b = torch.from_numpy(np.ascontiguousarray(bins, dtype=np.int32))
    if b.dim() == 2:
        b = b[:, :, None]
    p = torch.from_numpy(np.ascontiguousarray(pairs, dtype=np.int64))
    out = np.empty(len(pairs), np.float32)
    for s in range(0, len(pairs), batch_size):
        e = min(s + batch_size, len(pairs))
        out[s:e] = te_batch(b, p[s:e, 0], p[s:e, 1], b_max, tau).numpy()
    return out


def time_pairs(bins: np.ndarray, b_max: int, pairs: np.ndarray, tau: int = 1,
               repeats: int = 3, warmup: int = 1):
    """Median wall time of te_pairs over `repeats` runs after `warmup` runs
    (benchmarks/run_benchmark.py:67-83 protocol).  Returns (median_s, te)."""
    te = None
    for _ in range(warmup):
        te = te_pairs(bins, b_max, pairs, tau)
    ts = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        te = te_pairs(bins, b_max, pairs, tau)
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), te
