"""Full-SMEM and Adaptive-SMEM on gfx950: LDS-resident per-pair histograms.

Both names run the same HIP kernels (te_hist.hip): a wave-per-pair kernel for
pairs whose b_t^2 b_s histogram fits a wavefront's LDS slice and a
workgroup-per-pair multi-pass kernel for the rest, so every pair is counted
exactly in LDS whatever b_max is.  They differ only in supports() and in
prepare() (Adaptive-SMEM may coarsen on request), as in the reference
(tenex/kernels/full_smem.py:670-750, tenex/kernels/adaptive_smem.py:1212-1482).
"""
from __future__ import annotations

import numpy as np
import torch

from .. import core
from .._log import vprint
from . import TEKernel, LDS_BYTES_GFX950


def compute_smem_bin_limit(smem_optin: int, warps: int = 32) -> int:
    """Largest B with (B^3 + 2B^2 + B + warps) * 4 <= smem_optin
    (tenex/kernels/adaptive_smem.py:701-726)."""
    B = max(2, int((smem_optin / 4) ** (1.0 / 3.0)))
    while B > 2 and (B ** 3 + 2 * B ** 2 + B + warps) * 4 > smem_optin:
        B -= 1
    return B


def to_prepared(bin_arrs, n_per_var, tau, device) -> core.Prepared:
    """Accept a core.Prepared or a dense (n, T[, 1]) bin array."""
    if isinstance(bin_arrs, core.Prepared):
        if bin_arrs.tau != tau:
            return core.prepare(bin_arrs.bins, bin_arrs.npv, bin_arrs.T, tau)
        return bin_arrs
    dev = core.require_gpu(device)
    bins, bins16, npv, _, T = core.upload_bins_any(bin_arrs, dev)
    if n_per_var is not None:
        npv_in = torch.as_tensor(np.asarray(n_per_var.cpu() if isinstance(n_per_var, torch.Tensor)
                                            else n_per_var), dtype=torch.int32).to(dev)
        if npv_in.shape[0] == npv.shape[0]:
            if bool((npv_in < npv).any()):
                raise ValueError("n_per_var is smaller than the levels present in bin_arrs")
            npv = npv_in.contiguous()
    return core.prepare(bins, npv, T, tau, bins16=bins16)


class _LDSKernel(TEKernel):
    cap_wave = 0   # 0 = library default
    cap_block = 0

    @property
    def supports_pairfree(self) -> bool:
        return True

    def peak_bytes_per_pair(self, b_max, T, K, tau=1):
        return 4

    def compute_pairfree(self, bin_arrs, n_vars, b_max, n_per_var, tau, device,
                         pair_offset: int = 0, n_pairs_local=None, as_tensor: bool = False):
        prep = to_prepared(bin_arrs, n_per_var, tau, device)
        if prep.n_vars != n_vars:
            raise ValueError(f"n_vars={n_vars} does not match bin_arrs ({prep.n_vars})")
        out = core.te_pairfree(prep, pair_offset, n_pairs_local, self.cap_wave, self.cap_block)
        return out if as_tensor else out.cpu().numpy()

    def compute_single_gpu(self, bin_arrs, pairs, b_max, n_per_var, tau, batch_size, device,
                           **kwargs):
        prep = to_prepared(bin_arrs, n_per_var, tau, device)
        p = pairs.cpu().numpy() if isinstance(pairs, torch.Tensor) else np.asarray(pairs)
        p = p.astype(np.int64).reshape(-1, 2)
        if len(p) == 0:
            return np.empty(0, np.float32)
        return core.te_pair_list(prep, p, self.cap_wave).cpu().numpy()

    def log_description(self, **kwargs):
        return f"{self.name} (gfx950 LDS histograms, b_max={kwargs.get('b_max')})"


class FullSMEMKernel(_LDSKernel):
    """Pairs whose whole b^3 histogram fits one LDS pass."""

    @property
    def name(self) -> str:
        return "Full-SMEM"

    def supports(self, b_max, on_cuda, smem_optin, smem_bytes, n_per_var, source_filter):
        if not on_cuda:
            return False
        return b_max ** 3 <= 65536 and smem_bytes <= (smem_optin or LDS_BYTES_GFX950)


class AdaptiveSMEMKernel(_LDSKernel):
    """Per-pair b_t^2 b_s histograms, multi-pass in LDS for large pairs."""

    @property
    def name(self) -> str:
        return "Adaptive-SMEM"

    def supports(self, b_max, on_cuda, smem_optin, smem_bytes, n_per_var, source_filter):
        # any level count (tenex/kernels/adaptive_smem.py:1227-1229): pairs
        # of variables above 127 levels run on the occupied-key kernel
        return bool(on_cuda) and b_max <= core.MAX_LEVELS_INT16

    @staticmethod
    def bin_limit(L: int, smem_optin: int) -> int:
        """The reference's coarsening limit for a series of L samples: the
        LDS bin limit of a next_pow2(L)-thread block clamped to [128, 1024]
        (adaptive_smem.py:1289-1290)."""
        warps = max(128, min(1024, 1 << (max(1, int(L)) - 1).bit_length())) // 32
        return compute_smem_bin_limit(int(smem_optin or LDS_BYTES_GFX950), warps=warps)

    def prepare(self, bin_arrs, n_per_var, b_max, **kwargs):
        """coarsening=True applies the reference's lossy coarsening
        (tenex/kernels/adaptive_smem.py:1231-1300) with the gfx950 LDS bin
        limit; None/False keep exact counts (the multi-pass kernel needs no
        coarsening)."""
        use_coarsening = kwargs.get("use_coarsening")
        if use_coarsening is not True:
            return bin_arrs, n_per_var, b_max
        smem_optin = kwargs.get("smem_optin") or LDS_BYTES_GFX950
        if not isinstance(bin_arrs, core.Prepared):
            raise TypeError("prepare() expects core.Prepared data")
        limit = self.bin_limit(bin_arrs.L, smem_optin)
        if b_max <= limit:
            return bin_arrs, n_per_var, b_max
        vprint(f"[TENEX] bin coarsening: b_max={b_max}->{limit}")
        npv = bin_arrs.npv.clone()
        if bin_arrs.bins16 is not None:     # levels above 127: from the int16 rows
            bins = torch.empty_like(bin_arrs.bins)
            core.call("tnx_coarsen_bins16", core.ptr(bin_arrs.bins16), bin_arrs.n_vars,
                      bin_arrs.T, bin_arrs.ld, core.ptr(npv), int(limit), core.ptr(bins),
                      core._stream(bin_arrs.device))
        else:
            bins = bin_arrs.bins.clone()
            core.coarsen_gpu(bins, npv, bin_arrs.T, limit)
        prep = core.prepare(bins, npv, bin_arrs.T, bin_arrs.tau)
        return prep, npv, int(prep.npv_host.max())
