"""Cell alignment (host) and discretisation on the GPU (FSBW-L, dense remap,
coarsening).

discretize() restates tenex/preprocess.py:554-688 for method 'FSBW-L':
numpy-exact statistics (the reference's bin_fsbw_l_numpy, :61-88) computed by
a HIP kernel, clamp(min=0), per-variable dense remap (remap_bins,
:401-459) and int8 bins (_compact_dtype, :335-354).  The GPU statistics
reproduce numpy's float32 pairwise sums bit for bit, so use_numpy_bins=True
and False give identical bins here.
"""
from __future__ import annotations

import numpy as np
import torch

from . import core

BINNING_METHODS = ("fsbw-l", "default")


def align_data(exp_data: np.ndarray, trajectory: np.ndarray, branch: np.ndarray,
               branch_id: int = 1) -> np.ndarray:
    """Cells of one branch in pseudotime order (tenex/preprocess.py:16-45):
    (n_genes, n_cells) -> (n_genes, n_selected).  The order is numpy's
    default argsort of the selected pseudotimes, as in the reference, so
    tied pseudotimes keep the reference's arrangement."""
    n_cells = exp_data.shape[1]
    if len(trajectory) != n_cells or len(branch) != n_cells:
        raise ValueError(f"Dimension mismatch: exp_data has {n_cells} cells, trajectory has "
                         f"{len(trajectory)}, branch has {len(branch)}")
    mask = np.asarray(branch) == branch_id
    if mask.sum() == 0:
        raise ValueError(f"No cells found with branch_id={branch_id}")
    order = np.argsort(np.asarray(trajectory)[mask])
    return np.asarray(exp_data)[:, mask][:, order]


def discretize(arr, method: str = "fsbw-l", kp: float = 0.5, device=None,
               use_numpy_bins: bool = True):
    """Returns (bins int8 (n, ld) device, n_bins int32 (n,) device, b_max,
    n_per_var int32 (n,) device, T)."""
    if method.lower() not in BINNING_METHODS:
        raise ValueError(f"Unknown binning method '{method}'. This build implements "
                         f"{list(BINNING_METHODS)}")
    dev = core.require_gpu(device)
    x = torch.as_tensor(np.ascontiguousarray(arr, dtype=np.float32)).to(dev)
    n, T = x.shape
    if T < 2:
        raise ValueError("need at least 2 time points")
    raw, n_bins = core.fsbw_l_gpu(x, kp)
    bins, npv, b_max = core.remap_gpu(raw)
    return bins, n_bins, b_max, npv, T


def discretize_any(arr, method: str = "fsbw-l", kp: float = 0.5, device=None,
                   use_numpy_bins: bool = True):
    """discretize() for any level count up to 32,767: returns (bins int8,
    bins16 int16 (or None when every variable has at most 127 levels),
    n_bins, b_max, n_per_var, T).  The reference keeps int16 bins above 127
    levels (tenex/preprocess.py:335-354)."""
    if method.lower() not in BINNING_METHODS:
        raise ValueError(f"Unknown binning method '{method}'. This build implements "
                         f"{list(BINNING_METHODS)}")
    dev = core.require_gpu(device)
    x = torch.as_tensor(np.ascontiguousarray(arr, dtype=np.float32)).to(dev)
    n, T = x.shape
    if T < 2:
        raise ValueError("need at least 2 time points")
    raw, n_bins = core.fsbw_l_gpu(x, kp)
    bins, bins16, npv, b_max = core.remap_any_gpu(raw)
    return bins, bins16, n_bins, b_max, npv, T


def remap_bins(bin_arr):
    """Per-variable dense relabelling of an (n, T[, 1]) integer tensor on the
    GPU (tenex/preprocess.py:401-459).  Returns (bins (n, T, 1) int8 device,
    n_per_var int32 device, b_max)."""
    t = torch.as_tensor(bin_arr)
    if t.dim() == 3:
        t = t[:, :, 0]
    dev = core.require_gpu(t.device if t.is_cuda else None)
    raw = t.to(device=dev, dtype=torch.int32).clamp(min=0)
    bins, npv, b_max = core.remap_gpu(raw)
    return bins[:, : t.shape[1]].unsqueeze(-1), npv, b_max


def coarsen_bins(bin_arr, n_per_var, smem_bin_limit: int):
    """new = old * limit // n_per_var for variables above the limit
    (tenex/preprocess.py:462-551) via the HIP kernel."""
    t = torch.as_tensor(bin_arr)
    if t.dim() == 3:
        t = t[:, :, 0]
    dev = core.require_gpu(t.device if t.is_cuda else None)
    bins, npv, b_max, T = core.upload_bins(t, dev)
    npv_in = torch.as_tensor(n_per_var).to(device=dev, dtype=torch.int32).contiguous()
    core.coarsen_gpu(bins, npv_in, T, int(smem_bin_limit))
    return bins[:, :T].unsqueeze(-1), npv_in, int(npv_in.max().item())
