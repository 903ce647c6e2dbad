"""FDR inference: z-score -> Benjamini-Hochberg -> optional DPI trimming
(tenex/inference/fdr.py:22-133)."""
from __future__ import annotations

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
    """Data-processing-inequality trimming (tenex/inference/fdr.py:61-133):
    an edge i->j survives iff A[i, j] >= max_k min(A[i, k], A[k, j]) +
    trim_threshold over the thresholded network A."""
    n = te_matrix.shape[0]
    arr_np = np.zeros((n, n), dtype=np.float32)
    arr_np[pairs[:, 0], pairs[:, 1]] = np.asarray(te_values, dtype=np.float32)
    arr = torch.from_numpy(arr_np).to(device)
    mediators = (arr.sum(dim=1) > 0).nonzero(as_tuple=True)[0]
    if len(mediators) == 0:
        return make_grn(variable_names, pairs, np.asarray(te_values))
    if device.type == "cuda":
        free_mem, _ = torch.cuda.mem_get_info(device)
        budget = int(free_mem * 0.4)
    else:
        budget = 2 * 1024 ** 3
    out_pairs, out_te = [], []
    for r0 in range(0, n, batch_size):
        r1 = min(r0 + batch_size, n)
        batch = arr[r0:r1]
        if not bool((batch != 0).any()):
            continue
        cols = batch.nonzero(as_tuple=True)[1].unique()
        active = cols[torch.isin(cols, mediators)]
        if len(active) == 0:
            rows, cc = batch.nonzero(as_tuple=True)
            out_pairs.append(torch.stack((rows + r0, cc), 1).cpu().numpy())
            out_te.append(batch[rows, cc].cpu().numpy())
            continue
        bs = r1 - r0
        kc = max(1, min(budget // (bs * n * 4), len(active)))
        indirect = torch.zeros_like(batch)
        for k0 in range(0, len(active), kc):
            kidx = active[k0:k0 + kc]
            paths = torch.min(batch[:, kidx].unsqueeze(2), arr[kidx].unsqueeze(0))
            indirect = torch.max(indirect, paths.max(dim=1).values)
        if trim_threshold != 0.0:
            indirect += trim_threshold
        trimmed = torch.where(batch >= indirect, batch, torch.zeros_like(batch))
        rows, cc = torch.nonzero(trimmed, as_tuple=True)
        if len(rows) == 0:
            continue
        out_pairs.append(torch.stack((rows + r0, cc), 1).cpu().numpy())
        out_te.append(trimmed[rows, cc].cpu().numpy())
    if not out_pairs:
        return make_grn(variable_names, np.empty((0, 2), np.int32), np.empty(0, np.float32))
    return make_grn(variable_names, np.concatenate(out_pairs), np.concatenate(out_te))
