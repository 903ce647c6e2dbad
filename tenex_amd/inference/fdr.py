"""FDR inference: z-score -> Benjamini-Hochberg -> optional DPI trimming
(tenex/inference/fdr.py:22-133)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from .._log import vprint
from . import InferenceMethod, build_pairs, make_grn, threshold_pairs


class FDRMethod(InferenceMethod):

    @property
    def name(self) -> str:
        return "fdr"

    def infer(self, te_matrix, variable_names, device, **kwargs):
        fdr = kwargs.get("fdr", 0.01)
        links = kwargs.get("links", 0)
        sources = kwargs.get("sources", None)
        is_trimming = kwargs.get("is_trimming", True)
        trim_threshold = kwargs.get("trim_threshold", 0.0)
        batch_size = kwargs.get("batch_size", 512)
        pairs_np = build_pairs(variable_names, sources)
        n_pairs = len(pairs_np)
        te_matrix = np.asarray(te_matrix, dtype=np.float32)
        use_gpu = device.type == "cuda" and n_pairs <= 50_000_000
        if use_gpu:
            pairs_t = torch.from_numpy(pairs_np).to(device, dtype=torch.long)
            M = torch.from_numpy(te_matrix).to(device)
            te_t = M[pairs_t[:, 0], pairs_t[:, 1]]
            te_cut, pairs_cut = threshold_pairs(te_t, pairs_t, device, fdr=fdr, links=links)
        else:
            te_flat = te_matrix[pairs_np[:, 0], pairs_np[:, 1]]
            te_cut, pairs_cut = threshold_pairs(te_flat, pairs_np, torch.device("cpu"), fdr=fdr,
                                                links=links)
        grn = make_grn(variable_names, pairs_cut, te_cut)
        vprint(f"[TENEX] inference(fdr): {len(grn)} links after thresholding")
        if not is_trimming:
            return grn
        trimmed = trim_dpi(te_matrix, variable_names, pairs_cut, te_cut, device=device,
                           batch_size=batch_size, trim_threshold=trim_threshold)
        vprint(f"[TENEX] inference(fdr): {len(trimmed)} links after trimming")
        return grn, trimmed


def trim_dpi(te_matrix, variable_names, pairs, te_values, device, batch_size=512,
             trim_threshold=0.0):
    """Data-processing-inequality trimming (tenex/inference/fdr.py:61-133) of
    the thresholded network A (edges `pairs`, weights `te_values`): edge
    i -> j survives iff A[i,j] >= max(0, max_k min(A[i,k], A[k,j])) +
    trim_threshold, k over the mediators (rows of A with positive sum).  As
    in the reference, a block of `batch_size` rows none of whose edges
    reach a mediator keeps every edge, and a network without mediators is
    returned unchanged.  On a GPU the max-min product is the HIP kernel
    tnx_dpi_maxmin over the gathered rows x mediators; on the CPU device it
    is evaluated with torch in mediator chunks.  Output edges are in
    row-major order."""
    n = te_matrix.shape[0]
    dev = torch.device(device)
    A = torch.zeros((n, n), dtype=torch.float32, device=dev)
    pairs = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    if len(pairs):
        pd = torch.from_numpy(pairs).to(dev)
        A[pd[:, 0], pd[:, 1]] = torch.from_numpy(np.asarray(te_values, np.float32)).to(dev)
    mediator = A.sum(dim=1) > 0
    if not bool(mediator.any()):
        return make_grn(variable_names, pairs, np.asarray(te_values))
    nz = A != 0
    rows = torch.nonzero(nz.any(dim=1)).flatten()
    med = torch.nonzero(mediator).flatten()
    # blocks of rows whose edges reach no mediator are kept whole
    reach = (nz & mediator[None, :]).any(dim=1).to(torch.int32)
    blk = torch.arange(n, device=dev) // batch_size
    blk_active = torch.zeros(int((n + batch_size - 1) // batch_size), dtype=torch.int32,
                             device=dev).index_add_(0, blk, reach) > 0
    E = A.index_select(0, rows).contiguous()
    A_ik = E.index_select(1, med).contiguous()
    B_kn = A.index_select(0, med).contiguous()
    out = _maxmin_keep(A_ik, B_kn, E, float(trim_threshold), dev)
    whole = ~blk_active[blk[rows]]
    out[whole] = E[whole]
    r, c = torch.nonzero(out, as_tuple=True)
    kept = torch.stack([rows[r], c], 1).cpu().numpy()
    return make_grn(variable_names, kept, out[r, c].cpu().numpy())


def _maxmin_keep(A_ik, B_kn, E, thr, dev):
    """out = E where E >= max(0, max_k min(A_ik[:, k], B_kn[k, :])) + thr."""
    m, k = A_ik.shape
    n = B_kn.shape[1]
    out = torch.empty_like(E)
    if dev.type == "cuda":
        from .. import core
        core.call("tnx_dpi_maxmin", core.ptr(A_ik), m, k, A_ik.stride(0) if k else 1,
                  core.ptr(B_kn), n, n, core.ptr(E), n, C.c_float(thr), core.ptr(out), n,
                  core._stream(dev))
        return out
    acc = torch.zeros_like(E)
    kc = max(1, (256 << 20) // max(1, 4 * m * n))
    for k0 in range(0, k, kc):
        paths = torch.minimum(A_ik[:, k0:k0 + kc, None], B_kn[None, k0:k0 + kc, :])
        acc = torch.maximum(acc, paths.amax(dim=1))
    return torch.where(E >= acc + np.float32(thr), E, torch.zeros_like(E))
