"""Binarization into {0, 1} bins (tenex/binarize.py:1-295): a per-variable
threshold, then 1[x > thr].  b_max = 2 routes the engine to GEMM-B2.

Every method runs as one HIP launch pair (tnx_binarize): the threshold per
variable (an exact radix select for median / quantile, numpy's pairwise
float32 sum for the mean), then the comparison into padded int8 rows.  The
thresholds follow the reference's numpy backend bit for bit: 'median' is
numpy's median (the mean of the two middle values for an even length),
'quantile' numpy's linear quantile in float32.
"""
from __future__ import annotations

import numpy as np
import torch

from . import core

_METHODS = {
    "nonzero": (0, "active where nonzero (x > 0); scBoolSeq zero-inflated rule"),
    "mean": (1, "active above the per-variable mean (x > mean_g)"),
    "median": (2, "active above the per-variable median (x > median_g)"),
    "quantile": (3, "active above a per-variable quantile (x > Q_g(q), q default 0.5)"),
}
_BACKENDS = ("numpy", "torch", "cupy", "cuda")


def available_binarizers() -> list:
    """Sorted binarization method names."""
    return sorted(_METHODS)


def binarizer_doc(method: str) -> str:
    """One-line description of a binarization method."""
    return _METHODS[method.lower()][1]


def cupy_available() -> bool:
    """CuPy is not part of this build."""
    return False


def _check_q(q) -> float:
    if not 0.0 <= float(q) <= 1.0:
        raise ValueError(f"quantile q must be in [0, 1] (got {q})")
    return float(q)


def binarize_gpu(x: torch.Tensor, method: str = "nonzero", q: float = 0.5):
    """Device float32 (n, T) -> (bins int8 (n, ld) padded rows, thresholds
    float32 (n,)) through tnx_binarize."""
    key = method.lower()
    if key not in _METHODS:
        raise ValueError(f"unknown binarization '{method}'. available: {available_binarizers()}")
    code = _METHODS[key][0]
    q = _check_q(q) if key == "quantile" else 0.5
    x = x.to(torch.float32).contiguous()
    n, T = x.shape
    ld = core.round_up(T, core.ROW_ALIGN)
    bins = torch.empty((n, ld), dtype=torch.int8, device=x.device)
    thr = torch.empty(n, dtype=torch.float32, device=x.device)
    s = core._stream(x.device)
    for r0 in range(0, n, 65535):
        r1 = min(n, r0 + 65535)
        core.call("tnx_binarize", core.ptr(x[r0:r1]), r1 - r0, T, T, code, float(q),
                  core.ptr(bins[r0:r1]), ld, core.ptr(thr[r0:r1]), s)
    return bins, thr


def binarize(arr, method: str = "nonzero", device: torch.device = None,
             use_numpy: bool = True, backend: str = None, **params) -> tuple:
    """Binarize an aligned (n, T) float matrix (tenex/binarize.py:225-295).

    Returns (bin_arr (n, T, 1) int8 tensor on the device, n_bins (n,) int32,
    b_max, n_per_var (n,) int32) with n_per_var = max - min + 1 of each row
    (1 for a constant row), as the reference.  backend None, 'numpy',
    'torch' and 'cuda' all run the HIP kernels; 'cupy' is not available.
    use_numpy is accepted for compatibility (the thresholds follow numpy's
    arithmetic either way)."""
    key = method.lower()
    if key not in _METHODS:
        raise ValueError(f"unknown binarization '{method}'. available: {available_binarizers()}")
    if backend is not None:
        b = backend.lower()
        if b not in _BACKENDS:
            raise ValueError(f"unknown backend '{backend}'. available: {list(_BACKENDS)}")
        if b == "cupy":
            raise ValueError("the cupy backend is not available in this build")
    if device is None:
        device = torch.device("cuda" if torch.cuda.is_available() else "cpu")
    dev = core.require_gpu(torch.device(device) if torch.device(device).type == "cuda" else None)
    a = arr if isinstance(arr, torch.Tensor) else torch.from_numpy(
        np.ascontiguousarray(np.asarray(arr), dtype=np.float32))
    if a.dim() != 2:
        raise ValueError(f"binarize expects a (n, T) matrix (got shape {tuple(a.shape)})")
    bins, _ = binarize_gpu(a.to(dev), key, params.get("q", 0.5))
    T = a.shape[1]
    bin_arr = bins[:, :T].contiguous().unsqueeze(2)
    flat = bin_arr[:, :, 0]
    n_per_var = (flat.amax(1) - flat.amin(1) + 1).to(torch.int32)
    b_max = int(n_per_var.max().item()) if n_per_var.numel() else 0
    return bin_arr, n_per_var.clone(), b_max, n_per_var
