"""TE kernel registry: the reference's plugin API (tenex/kernels/__init__.py:85-519).

Kernels implement TEKernel; auto_select() returns the first registered
kernel whose supports() accepts the configuration.  Priority order:

  1. GEMM-B2       binary data (b_max == 2), no source filter: INT8 MFMA
  2. Full-SMEM     every pair's b_t^2 b_s histogram fits one LDS pass
  3. Adaptive-SMEM any b_max up to 127 (multi-pass LDS histograms)

All three run hand-written gfx950 HIP kernels from libtenex_hip.so.  The
reference's fourth kernel, the torch scatter_add fallback, is a CPU/torch
path; this build does not ship it (its restatement lives in oracle/ as the
CPU baseline), so on a machine without a ROCm GPU auto_select raises.
"""
from __future__ import annotations

from abc import ABC, abstractmethod
from typing import Optional

import numpy as np
import torch

from .._log import vprint

LDS_BYTES_GFX950 = 160 * 1024


def _next_pow2(x: int) -> int:
    return 1 << max(0, (x - 1).bit_length())


def query_smem_optin(device_ids=0) -> int:
    """Max LDS bytes per workgroup (min over the listed devices).
    Mirrors query_smem_optin (tenex/kernels/__init__.py:44-71); returns the
    gfx950 value, 160 KiB, when no device can be queried."""
    ids = list(device_ids) if isinstance(device_ids, (list, tuple)) else [int(device_ids)]
    if not ids:
        return LDS_BYTES_GFX950
    try:
        from ..core import device_props, require_gpu
        vals = [int(device_props(require_gpu(f"cuda:{d}")).lds_bytes_per_block) for d in ids]
        return min(vals)
    except Exception:
        return LDS_BYTES_GFX950


def compute_smem_bytes(b_max: int, n_time: int, n_kernels: int, tau: int) -> int:
    """LDS bytes of the largest single-pass histogram, b_max^3 int32 cells plus
    the b_max^2 (x_t, y_t) marginal (the role of compute_smem_bytes,
    tenex/kernels/__init__.py:74-80, for the gfx950 layout)."""
    return (b_max ** 3 + b_max ** 2) * 4


class TEKernel(ABC):
    """Abstract base class for TE kernels (tenex/kernels/__init__.py:85-442)."""

    @property
    @abstractmethod
    def name(self) -> str:
        """Kernel name (e.g. 'Full-SMEM')."""

    @property
    def is_matrix_kernel(self) -> bool:
        return False

    @property
    def supports_pairfree(self) -> bool:
        return False

    @abstractmethod
    def supports(self, b_max: int, on_cuda: bool, smem_optin: int, smem_bytes: int,
                 n_per_var: Optional[torch.Tensor], source_filter: bool) -> bool:
        """True if this kernel can handle the configuration."""

    def peak_batch_bytes(self, b_max, T, K, batch_size, tau=1) -> int:
        return self.peak_bytes_per_pair(b_max, T, K, tau) * batch_size

    def peak_bytes_per_pair(self, b_max, T, K, tau=1) -> int:
        return 4

    def heuristic_batch_size(self, bin_arrs, b_max, tau=1) -> int:
        """Pairs per launch.  The LDS kernels take whole plans, so the batch
        only bounds host-side chunking."""
        return 1 << 22

    def autotune_batch_size(self, bin_arrs, b_max, n_per_var, tau, device, **kw) -> int:
        return self.heuristic_batch_size(bin_arrs, b_max, tau)

    @abstractmethod
    def compute_single_gpu(self, bin_arrs, pairs, b_max, n_per_var, tau, batch_size, device,
                           **kwargs) -> np.ndarray:
        """TE for explicit [target, source] pairs; returns float32 (n_pairs,)."""

    def compute_multi_gpu(self, bin_arrs_cpu, pairs_np, b_max, n_per_var_cpu, tau, batch_size,
                          device_ids, **kwargs) -> np.ndarray:
        """One GPU per process in this build: runs on the first device."""
        dev = torch.device(f"cuda:{device_ids[0]}")
        return self.compute_single_gpu(bin_arrs_cpu, pairs_np, b_max, n_per_var_cpu, tau,
                                       batch_size, dev, **kwargs)

    def compute_matrix(self, data, n_vars, tau, device, **kwargs) -> np.ndarray:
        raise NotImplementedError(f"{self.name} is a pair-based kernel")

    def compute_matrix_multi_gpu(self, data, n_vars, tau, device_ids, **kwargs):
        raise NotImplementedError(f"{self.name} is a pair-based kernel")

    def prepare(self, bin_arrs, n_per_var, b_max, **kwargs):
        return bin_arrs, n_per_var, b_max

    def log_description(self, **kwargs) -> str:
        b_max = kwargs.get("b_max")
        return f"{self.name} (b_max={b_max})" if b_max is not None else self.name


_REGISTRY: list[TEKernel] = []
_REGISTERED = False


def register(kernel: TEKernel) -> TEKernel:
    _REGISTRY.append(kernel)
    return kernel


def get_kernel(name: str) -> TEKernel:
    """Case-insensitive lookup; KeyError if absent (kernels/__init__.py:457-469)."""
    needle = name.lower()
    for k in _REGISTRY:
        if k.name.lower() == needle:
            return k
    raise KeyError(f"Kernel '{name}' not found. Available: {[k.name for k in _REGISTRY]}")


def auto_select(b_max: int, on_cuda: bool, smem_optin: int, smem_bytes: int,
                n_per_var: Optional[torch.Tensor], source_filter: bool,
                exclude: Optional[set] = None) -> TEKernel:
    """First registered kernel whose supports() is true (kernels/__init__.py:472-491)."""
    for kernel in _REGISTRY:
        if exclude and kernel.name in exclude:
            continue
        if kernel.supports(b_max, on_cuda, smem_optin, smem_bytes, n_per_var, source_filter):
            return kernel
    raise RuntimeError(f"No suitable TE kernel found for b_max={b_max}, on_cuda={on_cuda}, "
                       f"smem_optin={smem_optin}")


def registered_kernels() -> list[TEKernel]:
    return list(_REGISTRY)


def _register_all():
    global _REGISTERED
    if _REGISTERED:
        return
    _REGISTERED = True
    from .gemm_b2 import GEMMB2Kernel
    from .lds_hist import AdaptiveSMEMKernel, FullSMEMKernel
    register(GEMMB2Kernel())
    register(FullSMEMKernel())
    register(AdaptiveSMEMKernel())


_register_all()
