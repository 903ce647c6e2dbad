"""NetWeaver: dispatcher from a TE result to an inference method
(tenex/inference/netweaver.py:21-165)."""
from __future__ import annotations

import numpy as np
import torch

from .._log import set_verbose
from . import get_method


class NetWeaver:
    """Infer network links from a TE matrix."""

    def __init__(self, result, sources=None, fdr: float = 0.01, links: int = 0,
                 is_trimming: bool = True, trim_threshold: float = 0.0):
        from ..result import TransferEntropyResult
        if isinstance(result, TransferEntropyResult):
            self.result_matrix = result.matrix
            self.variable_names = result.variable_names
            self.bin_arrs = result.bin_arrs
            self.n_per_var = result.n_per_var
            self.b_max = result.b_max
            self.kernel = result.kernel
            self.tau = result.tau
        elif isinstance(result, tuple) and len(result) == 2:
            self.result_matrix = np.asarray(result[0], dtype=np.float32)
            self.variable_names = np.asarray(result[1])
            self.bin_arrs = None
            self.n_per_var = None
            self.b_max = None
            self.kernel = None
            self.tau = 1
        else:
            raise TypeError("result must be a TransferEntropyResult or "
                            "(matrix, variable_names) tuple")
        self.sources = sources if sources is not None else getattr(result, "sources", None)
        self.fdr = fdr
        self.links = links
        self.is_trimming = is_trimming
        self.trim_threshold = trim_threshold

    def infer(self, method: str = "fdr", device=None, verbose: bool = False, **kwargs):
        set_verbose(verbose)
        if device is None:
            device = torch.device("cuda:0" if torch.cuda.is_available() else "cpu")
        else:
            device = torch.device(device)
        method_obj = get_method(method)
        params = dict(fdr=self.fdr, links=self.links, sources=self.sources,
                      is_trimming=self.is_trimming, trim_threshold=self.trim_threshold)
        if method == "trace":
            if "bin_data" not in kwargs and self.bin_arrs is not None:
                ba = np.asarray(self.bin_arrs)
                if ba.ndim == 3 and ba.shape[2] != 1:
                    raise ValueError(
                        f"TRACE requires K=1 bin arrays; the result has shape {ba.shape}. "
                        f"Re-run engine.compute() with a K=1 binning method (e.g. FSBW-L) "
                        f"before passing to TRACE.")
                params["bin_data"] = self.bin_arrs
            if "tau" not in kwargs:
                params["tau"] = self.tau
            for drop in ("fdr", "links", "sources", "is_trimming", "trim_threshold"):
                params.pop(drop, None)
        if method == "surrogate_test":
            if "bin_data" not in kwargs and self.bin_arrs is not None:
                params["bin_data"] = self.bin_arrs
            if "n_per_var" not in kwargs and self.n_per_var is not None:
                params["n_per_var"] = self.n_per_var
            if "b_max" not in kwargs and self.b_max is not None:
                params["b_max"] = self.b_max
            if "kernel" not in kwargs and self.kernel:
                params["kernel"] = self.kernel
            if "tau" not in kwargs:
                params["tau"] = self.tau
            for drop in ("links", "is_trimming", "trim_threshold"):
                params.pop(drop, None)
        params.update(kwargs)
        return method_obj.infer(self.result_matrix, self.variable_names, device, **params)
