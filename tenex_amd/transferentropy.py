"""TransferEntropyEngine: the drop-in entry point (tenex/transferentropy.py:38-1436).

compute(accelerator="gpu") discretises on the GPU (numpy-exact FSBW-L, dense
remap), selects a kernel through the registry (auto_select), and computes
every ordered pair with the gfx950 HIP kernels; the result matrix has the
reference's public layout matrix[i, j] = TE(var_i -> var_j).

This build runs on one ROCm GPU per process.  accelerator="cpu" (the
reference's torch scatter_add path) is not provided: it raises instead of
silently falling back.
"""
from __future__ import annotations

import time
import warnings
from typing import Optional, Union

import numpy as np
import torch

from . import core
from ._log import set_verbose, vprint
from .kernels import compute_smem_bytes, get_kernel, query_smem_optin
from .result import TransferEntropyResult


def get_device_list() -> list[int]:
    return list(range(torch.cuda.device_count())) if torch.cuda.is_available() else []


class TransferEntropyEngine:
    """Pairwise TE for all variable pairs of a multivariate time series.

    data (n_vars, T) float32 ordered along time; variable_names (n_vars,)
    unique; sources optional subset of variable_names used as TE sources.
    """

    def __init__(self, data: np.ndarray, variable_names: np.ndarray,
                 sources: Optional[np.ndarray] = None):
        if data is None or variable_names is None:
            raise ValueError("data and variable_names are required.")
        self._data = np.asarray(data, dtype=np.float32)
        self._variable_names = np.asarray(variable_names)
        if self._data.ndim != 2 or self._data.size == 0:
            raise ValueError(f"data must be a non-empty 2-D array of shape (n_vars, T); "
                             f"got shape {self._data.shape}")
        if not np.isfinite(self._data).all():
            raise ValueError("data contains NaN or infinite values")
        if self._variable_names.ndim != 1:
            raise ValueError(f"variable_names must be 1-D; got shape "
                             f"{self._variable_names.shape}")
        n_vars = self._data.shape[0]
        if self._variable_names.shape[0] != n_vars:
            raise ValueError(f"variable_names length ({self._variable_names.shape[0]}) must "
                             f"match the number of variables ({n_vars})")
        if len(np.unique(self._variable_names)) != n_vars:
            raise ValueError("variable_names must be unique")
        if sources is not None:
            src = np.atleast_1d(np.asarray(sources))
            if src.ndim != 1:
                raise ValueError(f"sources must be 1-D; got shape {src.shape}")
            missing = np.setdiff1d(src, self._variable_names)
            if missing.size > 0:
                raise ValueError(f"source names not found in variable_names: {list(missing)}")
            if len(np.unique(src)) != src.shape[0]:
                raise ValueError("sources must not contain duplicate names")
            self._sources = src.copy()
        else:
            self._sources = None
        self._result_matrix = None
        self._bin_arrs_cache: dict = {}
        self._pairs_cache: dict = {}
        self._last_run_info: Optional[dict] = None
        self._last_prep: Optional[core.Prepared] = None

    @property
    def variable_names(self) -> np.ndarray:
        return self._variable_names

    def clear_cache(self):
        """Free cached device tensors and plans."""
        self._bin_arrs_cache.clear()
        self._pairs_cache.clear()
        self._last_prep = None
        if torch.cuda.is_available():
            torch.cuda.empty_cache()

    def save_result_matrix(self, spath: str):
        """Tab-separated matrix with a header row/column of names
        (tenex/transferentropy.py:124-135)."""
        if self._result_matrix is None:
            raise ValueError("compute() has not been called yet")
        tmp = np.concatenate([self._variable_names[:, None],
                              self._result_matrix.astype(str)], axis=1)
        header = np.concatenate((["TE"], self._variable_names))
        tmp = np.concatenate([header[None, :], tmp])
        np.savetxt(spath, tmp, delimiter="\t", fmt="%s")
        vprint(f"Saved result matrix -> {spath}")

    @staticmethod
    def _resolve_accelerator_and_devices(accelerator="auto", devices="auto") -> list[int]:
        """Same contract as the reference (transferentropy.py:139-214): [] means
        CPU; 'gpu' without a device raises; unknown accelerators raise."""
        accel = accelerator.lower()
        if accel == "cpu":
            return []
        if accel == "auto":
            if not torch.cuda.is_available():
                return []
        elif accel == "gpu":
            if not torch.cuda.is_available():
                raise RuntimeError("accelerator='gpu' but no GPU device is available")
        else:
            raise ValueError(f"Unknown accelerator '{accelerator}'. "
                             f"Choose from: 'auto', 'cpu', 'gpu'")
        all_gpus = get_device_list()
        if devices is None or (isinstance(devices, str) and devices == "auto"):
            return all_gpus
        if isinstance(devices, int):
            if devices == -1:
                return all_gpus
            if devices <= 0:
                raise ValueError(f"devices as int must be >= 1 or -1 (got {devices})")
            if devices > len(all_gpus):
                raise ValueError(f"Requested {devices} GPUs but only {len(all_gpus)} available")
            return all_gpus[:devices]
        if isinstance(devices, (list, tuple)):
            for d in devices:
                if d not in all_gpus:
                    raise ValueError(f"GPU {d} not available. Available: {all_gpus}")
            return list(devices)
        raise TypeError(f"devices must be 'auto', int, list[int], or None "
                        f"(got {type(devices).__name__})")

    def _discretize(self, binning_method, kp, device, use_numpy_bins) -> core.Prepared:
        key = (binning_method.upper(), float(kp), str(device))
        hit = self._bin_arrs_cache.get(key)
        if hit is not None:
            return hit
        from .preprocess import discretize_any
        bins, bins16, n_bins, b_max, npv, T = discretize_any(self._data, binning_method, kp,
                                                             device, use_numpy_bins)
        self._bin_arrs_cache[key] = (bins, bins16, n_bins, b_max, npv, T)
        return self._bin_arrs_cache[key]

    def _binarize(self, method, device, backend, params):
        """{0, 1} bins of the data (tnx_binarize), cached per method, backend,
        parameters and device: (bins (n, ld) int8, kernel levels (n,) =
        max + 1, b_max, reported levels (n,) = max - min + 1, T)."""
        from .binarize import binarize_gpu, _BACKENDS, _METHODS
        if method.lower() not in _METHODS:
            raise ValueError(f"unknown binarization '{method}'. available: {sorted(_METHODS)}")
        if backend is not None and backend.lower() not in _BACKENDS:
            raise ValueError(f"unknown backend '{backend}'. available: {list(_BACKENDS)}")
        if backend is not None and backend.lower() == "cupy":
            raise ValueError("the cupy backend is not available in this build")
        key = ("BIN:" + method.upper(), tuple(sorted(params.items())), str(device))
        hit = self._bin_arrs_cache.get(key)
        if hit is not None:
            return hit
        x = torch.from_numpy(self._data).to(device)
        bins, _ = binarize_gpu(x, method, params.get("q", 0.5))
        T = self._data.shape[1]
        rows = bins[:, :T]
        hi = rows.amax(dim=1).to(torch.int32)
        npv = (hi + 1).contiguous()                   # levels 0..max for the kernels
        report = (hi - rows.amin(dim=1).to(torch.int32) + 1).contiguous()
        b_max = int(report.max().item())
        self._bin_arrs_cache[key] = (bins, npv, b_max, report, T)
        return self._bin_arrs_cache[key]

    def _source_indices(self) -> Optional[np.ndarray]:
        if self._sources is None:
            return None
        _, inds, _ = np.intersect1d(self._variable_names, self._sources, return_indices=True)
        if len(inds) == 0:
            raise ValueError("No source names found in variable_names.")
        return inds.astype(np.int32)

    def compute(self, accelerator: str = "auto", devices: Union[int, list, str, None] = "auto",
                binning_method: str = "FSBW-L", binarization: Optional[str] = None,
                binarization_backend: Optional[str] = None,
                binarization_kwargs: Optional[dict] = None, kp: float = 0.5, tau: int = 1,
                batch_size: Optional[int] = None, autotune: bool = False,
                kernel: Optional[str] = None, coarsening: Optional[bool] = None,
                use_numpy_bins: Optional[bool] = None, profile: bool = False,
                verbose: bool = False) -> TransferEntropyResult:
        """Pairwise TE (signature of tenex/transferentropy.py:537-562).

        Returns a TransferEntropyResult with matrix[i, j] = TE(var_i -> var_j).
        batch_size is accepted for API compatibility (the HIP kernels take
        whole plans); autotune=True times the TE kernels' LDS histogram sizes
        and the key-0 diversion on a sample of targets and runs the fastest
        (core.autotune).  coarsening=True applies the reference's
        lossy coarsening; the default keeps counts exact at any b_max <= 127.
        binarization ('nonzero', 'mean', 'median', 'quantile' with
        binarization_kwargs={'q': ...}) replaces binning by {0, 1} bins
        (tnx_binarize), which route to GEMM-B2 without a source filter.
        """
        set_verbose(verbose)
        t0 = time.time()
        device_ids = self._resolve_accelerator_and_devices(accelerator, devices)
        if not device_ids:
            raise RuntimeError(
                "tenex_amd computes on a ROCm GPU only; accelerator='cpu' (the reference's "
                "torch scatter_add path) is not provided by this build")
        if len(device_ids) > 1:
            vprint(f"[TENEX] one GPU per process: using cuda:{device_ids[0]}")
        dev = core.require_gpu(f"cuda:{device_ids[0]}")
        if binarization is None and binning_method.upper() not in ("FSBW-L", "DEFAULT"):
            raise ValueError(f"binning_method {binning_method!r} is not implemented (FSBW-L only)")
        vprint(f"[TENEX] accelerator: gpu:[{dev.index}]")
        ev = None
        stream = torch.cuda.current_stream(dev)
        if profile:
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(8)]
            ev[0].record(stream)
        arr = self._data
        n_vars, T = arr.shape
        if isinstance(tau, bool) or not isinstance(tau, (int, np.integer)) or tau < 1:
            raise ValueError(f"tau must be an integer >= 1; got {tau!r}")
        if T <= tau:
            raise ValueError(f"Time series length T={T} must be greater than tau={tau}")
        if n_vars < 2:
            self._result_matrix = np.zeros((n_vars, n_vars), dtype=np.float32)
            self._last_run_info = {}
            self._last_prep = None
            return self._wrap_result()
        if profile:
            ev[1].record(stream)
        report_npv = None
        bins16 = None
        if binarization is not None:
            bins, npv, b_max, report_npv, T = self._binarize(
                binarization, dev, binarization_backend, binarization_kwargs or {})
        else:
            bins, bins16, n_bins, b_max, npv, T = self._discretize(binning_method, kp, dev,
                                                                   use_numpy_bins)
        b_max_before = b_max
        if bins16 is not None and coarsening is True:
            # the reference's lossy coarsening, from int16 levels into int8 rows
            from .kernels.lds_hist import AdaptiveSMEMKernel
            lim = AdaptiveSMEMKernel.bin_limit(T - int(tau), query_smem_optin([dev.index]))
            bins = torch.empty_like(bins)
            npv = npv.clone()
            core.call("tnx_coarsen_bins16", core.ptr(bins16), n_vars, T, bins16.shape[1],
                      core.ptr(npv), int(lim), core.ptr(bins), core._stream(dev))
            bins16 = None
            b_max = int(npv.max().item())
        prep = core.prepare(bins, npv, T, int(tau), bins16=bins16)
        if profile:
            ev[2].record(stream)
        smem_bytes = compute_smem_bytes(b_max, T, 1, tau)
        smem_optin = query_smem_optin([dev.index])
        source_filter = self._sources is not None
        kernel_obj = self._select_kernel(kernel, b_max, True, smem_optin, smem_bytes, npv,
                                         source_filter)
        run_info = {"n_vars": n_vars, "T": T, "K": 1, "tau": tau, "n_pairs": n_vars * (n_vars - 1),
                    "device_ids": [dev.index], "n_gpus": 1,
                    "binning_method": binning_method if binarization is None else binarization,
                    "kp": kp, "use_numpy_bins": use_numpy_bins, "b_max_before_coarsen": b_max_before,
                    "b_max": b_max, "kernel": kernel_obj.name, "smem_optin": smem_optin,
                    "smem_bytes": smem_bytes, "source_filter": source_filter,
                    "binarization": binarization,
                    "bin_dtype": "torch.int16" if bins16 is not None else "torch.int8",
                    "bin_bytes": n_vars * T * (2 if bins16 is not None else 1)}
        if profile:
            ev[3].record(stream)
        src_idx = self._source_indices()
        if kernel_obj.is_matrix_kernel:
            run_info.update(mode="matrix", pair_free=False, coarsened=False)
            if profile:
                ev[4].record(stream)
            internal = kernel_obj.compute_matrix(prep, n_vars, tau, dev, as_tensor=True)
            if profile:
                ev[5].record(stream)
        else:
            prep2, npv2, b_max2 = kernel_obj.prepare(prep, npv, b_max, use_coarsening=coarsening,
                                                     on_cuda=True, n_vars=n_vars, tau=tau,
                                                     smem_optin=smem_optin)
            run_info.update(b_max=b_max2, coarsened=b_max2 != b_max_before,
                            mode="pair-free" if src_idx is None else "pair-based",
                            pair_free=src_idx is None)
            prep = prep2
            if profile:
                ev[4].record(stream)
                ev[5].record(stream)
            cw, cb = kernel_obj.cap_wave, kernel_obj.cap_block
            if autotune and not len(prep.wide):
                tuned = core.autotune(prep, None, src_idx)
                cw, cb = tuned["cap_wave"], tuned["cap_block"]
                run_info["autotune"] = tuned
            internal = core.te_matrix_internal(prep, None, src_idx, cw, cb)
        if profile:
            ev[6].record(stream)
        # pinned host memory: NetWeaver's tests upload it again without staging
        self._result_matrix = core.d2h_pinned(internal.t().contiguous())
        if profile:
            ev[7].record(stream)
            torch.cuda.synchronize(dev)
            ms = lambda a, b: ev[a].elapsed_time(ev[b]) / 1000.0  # noqa: E731
            run_info["timings"] = {"total": time.time() - t0, "data_align": ms(0, 1),
                                   "discretize": ms(1, 2), "kernel_select": ms(2, 3),
                                   "prepare": ms(3, 4), "compute": ms(5, 6),
                                   "assemble": ms(6, 7)}
        self._last_run_info = run_info
        self._last_prep = prep
        self._last_report_npv = None if report_npv is None else report_npv.cpu().numpy()
        el = time.time() - t0
        vprint(f"[TENEX] done -- {el:.3f}s  ({n_vars * (n_vars - 1) / el:,.0f} pairs/s)")
        return self._wrap_result()

    @staticmethod
    def _select_kernel(kernel_name, b_max, on_cuda, smem_optin, smem_bytes, n_per_var,
                       source_filter):
        """Explicit name (validated with supports()) or auto_select()
        (tenex/transferentropy.py:1325-1357)."""
        from .kernels import auto_select
        if kernel_name is not None:
            k = get_kernel(kernel_name)
            if k.name == "GEMM-B2" and b_max != 2:
                raise ValueError(f"kernel='GEMM-B2' requires binary data (b_max=2), but this "
                                 f"data has b_max={b_max}. Use kernel=None for auto-selection.")
            if not k.supports(b_max, on_cuda, smem_optin, smem_bytes, n_per_var, source_filter):
                raise ValueError(f"kernel={kernel_name!r} does not support this configuration "
                                 f"(b_max={b_max}, on_cuda={on_cuda}, "
                                 f"source_filter={source_filter}). Use kernel=None for "
                                 f"auto-selection.")
            return k
        return auto_select(b_max, on_cuda, smem_optin, smem_bytes, n_per_var, source_filter)

    def _wrap_result(self) -> TransferEntropyResult:
        info = self._last_run_info or {}
        prep = self._last_prep
        bin_arrs = n_per_var = None
        if prep is not None:
            src = prep.bins16 if prep.bins16 is not None else prep.bins
            # pinned host memory: an inference step uploads these bins again
            bin_arrs = core.d2h_pinned(src[:, : prep.T].contiguous())[:, :, None]
            n_per_var = prep.npv_host.copy()
            if getattr(self, "_last_report_npv", None) is not None:
                n_per_var = self._last_report_npv.copy()   # binarized: max - min + 1
        return TransferEntropyResult(matrix=self._result_matrix,
                                     variable_names=self._variable_names, bin_arrs=bin_arrs,
                                     n_per_var=n_per_var, b_max=info.get("b_max", 0),
                                     tau=info.get("tau", 1), kernel=info.get("kernel", ""),
                                     timings=info.get("timings"), sources=self._sources)
