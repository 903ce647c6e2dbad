"""GEMM-B2: binary data (b_max == 2) on exact MFMA contractions (gemm_b2.hip).

Reference: tenex/kernels/gemm_b2.py:307-425 (three FP32 torch.mm
contractions C = fp.p^T, D = p.p^T, E = f.p^T, then the 8-state epilogue
:210-285 in Triton).  Here one fused gfx950 kernel contracts the 0/1
operands (FP4 e2m1 with FP32 sums below 2^24 samples, INT8 with INT32
sums beyond: counts exact either way) and applies the 8-state epilogue in
registers.
"""
from __future__ import annotations

import numpy as np
import torch

from .. import _lib, core
from ..core import call, ptr
from . import TEKernel


def te_gemm_b2_device(prep: core.Prepared, counts_dbg: bool = False,
                      precision: int = _lib.TNX_B2_AUTO):
    """Internal-layout (n, n) TE matrix te[i, j] = TE(j -> i) on the device.
    With counts_dbg=True also returns int32 (3, n, n) counts C, D, E.
    precision: TNX_B2_AUTO (FP4 0/1 operands with FP32 sums below 2^24
    samples, else INT8), TNX_B2_INT8 or TNX_B2_FP4; all exact."""
    n, T, tau = prep.n_vars, prep.T, prep.tau
    if int(prep.npv_host.max(initial=1)) > 2:
        raise ValueError("GEMM-B2 requires binary data (b_max <= 2)")
    ws_bytes = int(_lib.load().tnx_gemm_b2_workspace(n, T, tau))
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=prep.device)
    out = torch.empty((n, n), dtype=torch.float32, device=prep.device)
    cnt = torch.zeros((3, n, n), dtype=torch.int32, device=prep.device) if counts_dbg else None
    call("tnx_te_gemm_b2", ptr(prep.bins), n, T, prep.ld, int(tau), ptr(prep.S1), ptr(prep.S2a),
         ptr(prep.xlogx), ptr(ws), ws_bytes, ptr(out), n, ptr(cnt), int(precision),
         torch.cuda.current_stream(prep.device).cuda_stream)
    return (out, cnt) if counts_dbg else out


class GEMMB2Kernel(TEKernel):
    """Exact TE for binary data from INT8 MFMA contractions."""

    @property
    def name(self) -> str:
        return "GEMM-B2"

    @property
    def is_matrix_kernel(self) -> bool:
        return True

    def supports(self, b_max, on_cuda, smem_optin, smem_bytes, n_per_var, source_filter):
        return bool(on_cuda) and b_max == 2 and not source_filter

    def peak_bytes_per_pair(self, b_max, T, K, tau=1):
        return 4

    def _prep(self, data, tau, device):
        from .lds_hist import to_prepared
        return to_prepared(data, None, tau, device)

    def compute_matrix(self, data, n_vars, tau, device, as_tensor: bool = False, **kwargs):
        """(n, n) internal layout: result[i, j] = TE(j -> i), diagonal 0
        (compute_te_gemm_b2, tenex/kernels/gemm_b2.py:307-336)."""
        prep = self._prep(data, tau, device)
        out = te_gemm_b2_device(prep)
        return out if as_tensor else out.cpu().numpy()

    def compute_single_gpu(self, bin_arrs, pairs, b_max, n_per_var, tau, batch_size, device,
                           **kwargs):
        m = self.compute_matrix(bin_arrs, None, tau, device, as_tensor=True)
        p = pairs.cpu().numpy() if isinstance(pairs, torch.Tensor) else np.asarray(pairs)
        pt = torch.from_numpy(p.astype(np.int64).reshape(-1, 2)).to(m.device)
        return m[pt[:, 0], pt[:, 1]].cpu().numpy()

    def log_description(self, **kwargs):
        return "GEMM-B2 (FP4/INT8 MFMA on 0/1 operands, exact counts, fused 8-state epilogue)"
