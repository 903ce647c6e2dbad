"""Link inference from a TE matrix: method registry and shared utilities
(tenex/inference/__init__.py:32-199).

Implemented methods: 'fdr' (z-score -> BH -> optional DPI trim),
'surrogate_test' (block-shuffle surrogates on the gfx950 kernels, optional
learned null) and 'trace' (thresholded network, OutTE / InTE).  The
reference's 'clr', 'nd' and 'point' are outside this build's hot path; they
are registered so available_methods() keeps the reference's list, and raise
NotImplementedError when called.
"""
from __future__ import annotations

from abc import ABC, abstractmethod

import numpy as np
import torch

from .grn import GRN


class InferenceMethod(ABC):
    """Abstract base class for link inference algorithms."""

    @property
    @abstractmethod
    def name(self) -> str:
        """Short method name."""

    @abstractmethod
    def infer(self, te_matrix: np.ndarray, variable_names: np.ndarray, device: torch.device,
              **kwargs):
        """Run inference on a (n, n) TE matrix M[i, j] = TE(i -> j)."""


_REGISTRY: dict[str, InferenceMethod] = {}


def register(method: InferenceMethod):
    _REGISTRY[method.name] = method


def get_method(name: str) -> InferenceMethod:
    if name not in _REGISTRY:
        available = ", ".join(sorted(_REGISTRY.keys()))
        raise ValueError(f"Unknown inference method: {name!r}. Available: {available}")
    return _REGISTRY[name]


def available_methods() -> list[str]:
    return sorted(_REGISTRY.keys())


def build_pairs(variable_names: np.ndarray, sources=None) -> np.ndarray:
    """Directed pairs [source, target] (tenex/inference/__init__.py:86-103)."""
    n = len(variable_names)
    if sources is not None:
        _, src_idx, _ = np.intersect1d(variable_names, sources, return_indices=True)
        all_idx = np.arange(n)
        src_rep = np.repeat(src_idx, n)
        tgt_rep = np.tile(all_idx, len(src_idx))
        mask = src_rep != tgt_rep
        return np.stack((src_rep[mask], tgt_rep[mask]), axis=1)
    idx = np.arange(n)
    src = np.repeat(idx, n - 1)
    base = np.tile(np.arange(n - 1), n)
    tgt = base + (base >= src).astype(np.int64)
    return np.stack((src, tgt), axis=1).astype(np.int32)


def make_grn(variable_names: np.ndarray, pairs: np.ndarray, te: np.ndarray) -> GRN:
    return GRN(source=variable_names[pairs[:, 0]], target=variable_names[pairs[:, 1]],
               te=te.astype(np.float32), pairs=pairs)


def bh_fdr_gpu(pval: torch.Tensor, alpha: float) -> torch.Tensor:
    """Benjamini-Hochberg adjusted p-values in the p-value dtype
    (tenex/inference/__init__.py:115-126)."""
    m = pval.shape[0]
    if m == 0:
        return pval
    sorted_pval, sort_idx = pval.sort()
    rank = torch.arange(1, m + 1, device=pval.device, dtype=pval.dtype)
    adjusted = (sorted_pval * m / rank).clamp(max=1.0)
    adjusted = adjusted.flip(0).cummin(0).values.flip(0)
    result = torch.empty_like(adjusted)
    result[sort_idx] = adjusted
    return result


def bh_reject_gpu(pval: torch.Tensor, alpha: float) -> torch.Tensor:
    """bh_fdr_gpu(pval, alpha) < alpha without the reverse running minimum:
    with q_(i) = min(p_(i) m / i, 1) in the p-value dtype (the same
    arithmetic), adj_(i) = min_{j >= i} q_(j) < alpha exactly when some
    j >= i has q_(j) < alpha, i.e. i <= k, the last rank with q < alpha."""
    m = pval.shape[0]
    if m == 0:
        return torch.zeros(0, dtype=torch.bool, device=pval.device)
    sorted_pval, sort_idx = pval.sort()
    rank = torch.arange(1, m + 1, device=pval.device, dtype=pval.dtype)
    q = (sorted_pval * m / rank).clamp(max=1.0)
    pos = torch.arange(1, m + 1, device=pval.device)
    k = torch.where(q < alpha, pos, torch.zeros_like(pos)).max()
    result = torch.empty(m, dtype=torch.bool, device=pval.device)
    result[sort_idx] = pos <= k
    return result


def bh_adjust_np(p: np.ndarray) -> np.ndarray:
    """BH adjusted p-values in float64 (the statsmodels fdr_bh arithmetic the
    reference's CPU path calls, tenex/inference/__init__.py:150)."""
    p = np.asarray(p, dtype=np.float64)
    m = p.shape[0]
    if m == 0:
        return p
    order = np.argsort(p)
    ranked = p[order] * m / np.arange(1, m + 1)
    adj = np.minimum.accumulate(ranked[::-1])[::-1]
    adj = np.minimum(adj, 1.0)
    out = np.empty_like(adj)
    out[order] = adj
    return out


def fdr_threshold_gpu(te: torch.Tensor, pairs: torch.Tensor, fdr: float):
    """z-score -> p -> BH on the device (tenex/inference/__init__.py:129-138)."""
    std = te.std()
    if std == 0:
        return torch.empty(0, device=te.device), pairs[:0]
    z = (te - te.mean()) / std
    pval = 0.5 * torch.erfc(z / 1.4142135623730951)
    fdr_pval = bh_fdr_gpu(pval, 0.05)
    mask = fdr_pval < fdr
    return te[mask], pairs[mask]


def fdr_threshold_cpu(te: np.ndarray, pairs: np.ndarray, fdr: float):
    """z-score (ddof=0) -> 1 - Phi(z) -> BH on the host
    (tenex/inference/__init__.py:141-152)."""
    import scipy.stats
    std = te.std()
    if std == 0:
        return te[:0], pairs[:0]
    z = (te - te.mean()) / std
    pval = 1.0 - scipy.stats.norm.cdf(z)
    fdr_pval = bh_adjust_np(pval)
    mask = fdr_pval < fdr
    return te[mask], pairs[mask]


def threshold_pairs(te, pairs, device, fdr=0.01, links=0, sources=None):
    """Top-K or FDR thresholding (tenex/inference/__init__.py:155-178)."""
    if isinstance(te, torch.Tensor):
        if links > 0:
            k = min(links, len(te))
            _, idx = te.topk(k)
            idx = idx[te[idx].argsort(descending=True)]
            return te[idx].cpu().numpy(), pairs[idx].cpu().numpy()
        if te.device.type == "cuda":
            te_c, p_c = fdr_threshold_gpu(te, pairs, fdr)
            return te_c.cpu().numpy(), p_c.cpu().numpy()
        return fdr_threshold_cpu(te.cpu().numpy(), pairs.cpu().numpy(), fdr)
    if links > 0:
        k = min(links, len(te))
        idx = np.argpartition(te, -k)[-k:]
        idx = idx[np.argsort(te[idx])[::-1]]
        return te[idx], pairs[idx]
    return fdr_threshold_cpu(te, pairs, fdr)


class _OutOfScope(InferenceMethod):
    def __init__(self, name):
        self._name = name

    @property
    def name(self) -> str:
        return self._name

    def infer(self, te_matrix, variable_names, device, **kwargs):
        raise NotImplementedError(f"inference method {self._name!r} is outside this build")


def _register_all():
    from .fdr import FDRMethod
    from .surrogate_test import SurrogateTestMethod
    from .trace import TRACEMethod
    register(FDRMethod())
    register(SurrogateTestMethod())
    register(TRACEMethod())
    for nm in ("clr", "nd", "point"):
        register(_OutOfScope(nm))


_register_all()

from .netweaver import NetWeaver  # noqa: E402
