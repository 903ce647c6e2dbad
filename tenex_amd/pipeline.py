"""Pipeline: compute TE once, run several inference methods
(tenex/pipeline.py:46-458)."""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any

import numpy as np
import torch

from ._log import set_verbose, vprint
from .inference import get_method
from .inference.surrogate_test import surrogate_test_lags
from .result import TransferEntropyResult
from .transferentropy import TransferEntropyEngine

_MATRIX_METHODS = frozenset({"fdr", "clr", "nd"})
_FUSED_METHODS = frozenset({"surrogate_test", "point"})


@dataclass
class PipelineResult:
    """te_result (shared TE matrix) and inference_results (method -> result)."""
    te_result: TransferEntropyResult
    inference_results: dict = field(default_factory=dict)

    @property
    def matrix(self) -> np.ndarray:
        return self.te_result.matrix

    def get(self, method: str) -> Any:
        return self.inference_results[method]

    def __getattr__(self, name: str) -> Any:
        inf = self.__dict__.get("inference_results", {})
        if name in inf:
            return inf[name]
        raise AttributeError(f"{type(self).__name__!r} has no attribute {name!r}")


class Pipeline:
    """End-to-end TE + link inference with a cached TE result."""

    def __init__(self, engine: TransferEntropyEngine | None = None, **defaults):
        self.engine = engine
        self._defaults = dict(defaults)
        self._te_result: TransferEntropyResult | None = None
        self._compute_kwargs: dict = {}
        self._data_fingerprint = None
        if engine is not None:
            self._data_fingerprint = self._fingerprint(engine._data, engine._variable_names,
                                                       engine._sources)

    @staticmethod
    def _fingerprint(data, variable_names, sources) -> tuple:
        def _key(arr):
            if arr is None:
                return None
            a = np.asarray(arr)
            return (a.ctypes.data, a.shape, str(a.dtype))
        return (_key(data), _key(variable_names), _key(sources))

    def configure(self, **kwargs) -> "Pipeline":
        changed = any(k not in self._compute_kwargs or self._compute_kwargs[k] != v
                      for k, v in kwargs.items())
        self._compute_kwargs.update(kwargs)
        if changed and self._te_result is not None:
            self._te_result = None
        return self

    def _ensure_engine(self, data, variable_names, sources) -> None:
        if data is None:
            if self.engine is None:
                raise ValueError("Pipeline has no engine; pass data and variable_names to run() "
                                 "(or construct Pipeline with a pre-built engine).")
            return
        if variable_names is None:
            raise ValueError("variable_names is required when data is provided.")
        fp = self._fingerprint(data, variable_names, sources)
        if self.engine is not None and fp == self._data_fingerprint:
            return
        self.engine = TransferEntropyEngine(data=data, variable_names=variable_names,
                                            sources=sources)
        self._data_fingerprint = fp
        self._te_result = None

    @property
    def te_result(self) -> TransferEntropyResult:
        if self._te_result is None:
            self._te_result = self.engine.compute(**self._compute_kwargs)
        return self._te_result

    def clear_cache(self) -> None:
        self._te_result = None

    def infer(self, method: str, taus=None, lag_combine: str = "max", **kwargs) -> Any:
        if taus is not None:
            return self._infer_multi_lag(taus=taus, combine=lag_combine,
                                         inference_method=method, **kwargs)
        params = {**self._defaults, **kwargs}
        method_obj = get_method(method)
        te = self.te_result
        device = params.pop("device", None)
        if device is None:
            device = torch.device("cuda:0" if torch.cuda.is_available() else "cpu")
        elif not isinstance(device, torch.device):
            device = torch.device(device)
        if method in _FUSED_METHODS:
            params.setdefault("bin_data", te.bin_arrs)
            params.setdefault("n_per_var", te.n_per_var)
            params.setdefault("b_max", te.b_max)
            params.setdefault("kernel", te.kernel)
            params.setdefault("tau", te.tau)
        src = getattr(self.engine, "_sources", None)
        if src is not None:
            params.setdefault("sources", src)
        return method_obj.infer(te.matrix, te.variable_names, device, **params)

    def _infer_multi_lag(self, taus, combine="max", inference_method="surrogate_test", **kwargs):
        """Per-lag surrogate tests combined per pair (tenex/pipeline.py:252-409):
        'max' (largest effective TE), 'adaptive_select' (smallest p) or
        'weighted_sum'; Bonferroni across the K lags, BH on off-diagonal.
        By default every lag shares each launch (surrogate_test_lags: the
        observed TE of all lags in one launch, each batch of surrogates of
        all lags in one); batched_lags=False, or options that path does not
        cover, run the lags one by one as the reference does."""
        if not taus:
            raise ValueError("taus must be a non-empty list of time-lags")
        if combine not in ("adaptive_select", "weighted_sum", "max"):
            raise ValueError(f"combine must be 'adaptive_select', 'weighted_sum', or 'max' "
                             f"(got {combine!r})")
        if inference_method != "surrogate_test":
            raise NotImplementedError(f"multi_lag combine for {inference_method!r} is not "
                                      f"supported; only 'surrogate_test' is implemented")
        results = None
        if kwargs.get("batched_lags", True):
            # every lag in each launch (observed TE, then each surrogate batch)
            te = self.te_result
            params = {**self._defaults, **kwargs}
            src = getattr(self.engine, "_sources", None)
            if src is not None:
                params.setdefault("sources", src)
            params.pop("batched_lags", None)
            device = params.pop("device", None)
            vprint(f"[TENEX] multi_lag: taus={list(taus)} in shared launches")
            results = surrogate_test_lags(te.bin_arrs, te.n_per_var, te.variable_names, taus,
                                          device, **params)
        if results is None:
            kwargs.pop("batched_lags", None)
            results = []
            original_tau = self._compute_kwargs.get("tau", 1)
            for tau in taus:
                self.configure(tau=tau)
                self.clear_cache()
                vprint(f"[TENEX] multi_lag: computing tau={tau}")
                results.append(self.infer(inference_method, **kwargs))
            self.configure(tau=original_tau)
        device = torch.device("cuda:0" if torch.cuda.is_available() else "cpu")
        return combine_surrogate_multi_lag(results, taus, combine, self.engine.variable_names,
                                           device)

    def run(self, data=None, variable_names=None, sources=None, methods=None,
            method_kwargs=None, verbose: bool = False) -> PipelineResult:
        set_verbose(verbose)
        if methods is None:
            raise ValueError("methods is required")
        self._ensure_engine(data, variable_names, sources)
        method_kwargs = method_kwargs or {}
        self._compute_kwargs["verbose"] = verbose
        _ = self.te_result
        results = {m: self.infer(m, **method_kwargs.get(m, {})) for m in methods}
        return PipelineResult(te_result=self._te_result, inference_results=results)


def combine_surrogate_multi_lag(results, taus, combine, variable_names, device):
    """Combine per-tau SurrogateTestResults (tenex/pipeline.py:313-409)."""
    from .inference import bh_fdr_gpu, make_grn
    from .inference.surrogate_test import SurrogateTestResult
    n = results[0].observed_te.shape[0]
    variable_names = np.asarray(variable_names)
    K = len(taus)
    obs = np.stack([r.observed_te for r in results])
    eff = np.stack([r.effective_te for r in results])
    mean = np.stack([r.mean_surrogate_te for r in results])
    std = np.stack([r.std_surrogate_te for r in results])
    pv = np.stack([r.p_values for r in results])
    if combine in ("max", "adaptive_select"):
        sel = (np.argmax(eff, axis=0, keepdims=True) if combine == "max"
               else np.argmin(pv, axis=0, keepdims=True))
        effective = np.take_along_axis(eff, sel, axis=0)[0]
    else:
        w = np.where(pv < 0.05, -np.log(pv + 1e-12), 0.0).astype(np.float32)
        if not np.any(w > 0):
            raise ValueError(f"weighted_sum: no pair passed p<0.05 at any lag in taus={taus}; "
                             f"the combined result would be identically zero. Loosen "
                             f"n_surrogates, add more lags, or switch to combine='max'.")
        effective = np.sum(eff * w, axis=0)
        sel = np.argmin(pv, axis=0, keepdims=True)
    observed = np.take_along_axis(obs, sel, axis=0)[0]
    mean_s = np.take_along_axis(mean, sel, axis=0)[0]
    std_s = np.take_along_axis(std, sel, axis=0)[0]
    p_values = np.minimum(1.0, pv.min(axis=0) * K).astype(np.float32)
    np.fill_diagonal(p_values, 1.0)
    fdr = results[0].fdr
    off = ~np.eye(n, dtype=bool)
    adj = bh_fdr_gpu(torch.from_numpy(p_values[off]).to(device), alpha=fdr).cpu().numpy()
    sig = np.zeros((n, n), dtype=bool)
    sig[off] = adj < fdr
    sig &= effective > 0
    pairs = np.argwhere(sig).astype(np.int64)
    grn = make_grn(variable_names, pairs, effective[pairs[:, 0], pairs[:, 1]])
    first = results[0]
    return SurrogateTestResult(
        observed_te=observed.astype(np.float32), mean_surrogate_te=mean_s.astype(np.float32),
        std_surrogate_te=std_s.astype(np.float32), effective_te=effective.astype(np.float32),
        p_values=p_values, grn=grn, n_surrogates=first.n_surrogates,
        shuffle_method=first.shuffle_method, block_length=first.block_length,
        p_method=first.p_method, fdr=fdr)
