"""TRACE: thresholded causal network plus OutTE / InTE per variable
(tenex/inference/trace.py:243-760).

The network keeps the pairs whose TE exceeds mean + significance * std of
n_surrogates random off-diagonal entries (build_causal_network).  OutTE(X)
is the transfer entropy from X into the joint state of its network targets,
InTE(X) the one from the joint state of its network sources into X; both are
four joint entropies of 64-bit hashed states, evaluated on the GPU by
tnx_trace_entropy (one workgroup per variable and entropy: keys built in
registers, sorted in LDS, runs summed exactly in fixed point).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from .. import core
from .._log import vprint
from . import GRN, InferenceMethod, make_grn


@dataclass
class TRACEResult:
    """Result of TRACE analysis (tenex/inference/trace.py:243-261)."""
    outte: np.ndarray          # (n,) float32 -- OutTE of each variable
    inte: np.ndarray           # (n,) float32 -- InTE of each variable
    variable_names: np.ndarray
    network: dict              # target index -> source indices
    grn: GRN                   # the thresholded network's edges and TE

    def top_drivers(self, k: int = 10) -> list[tuple[str, float]]:
        """Top-k variables by OutTE."""
        idx = np.argsort(self.outte)[::-1][:k]
        return [(self.variable_names[i], float(self.outte[i])) for i in idx]

    def top_receivers(self, k: int = 10) -> list[tuple[str, float]]:
        """Top-k variables by InTE."""
        idx = np.argsort(self.inte)[::-1][:k]
        return [(self.variable_names[i], float(self.inte[i])) for i in idx]


def causal_threshold(te_matrix: np.ndarray, n_surrogates: int = 100,
                     significance: float = 2.0, seed: int = 42) -> float:
    """mean + significance * std (float32, population std) of n_surrogates
    off-diagonal entries drawn with default_rng(seed) (trace.py:561-572)."""
    n = te_matrix.shape[0]
    rng = np.random.default_rng(seed)
    rows = rng.integers(0, n, size=n_surrogates)
    cols = rng.integers(0, n, size=n_surrogates)
    same = rows == cols
    cols[same] = (cols[same] + 1) % n
    sv = np.asarray(te_matrix, dtype=np.float32)[rows, cols]
    return float(sv.mean() + significance * sv.std())


def build_causal_network(te_matrix: np.ndarray, n_surrogates: int = 100,
                         significance: float = 2.0, seed: int = 42,
                         device: torch.device | None = None) -> dict[int, list[int]]:
    """Directed network {target: sources} of the pairs with TE above the
    causal threshold (trace.py:550-612).  On a GPU device the diagonal is
    zeroed before the comparison, on the host it is excluded after it (the
    reference's two branches).  Sources of a target are in ascending
    order."""
    te = np.asarray(te_matrix, dtype=np.float32)
    n = te.shape[0]
    thr = causal_threshold(te, n_surrogates, significance, seed)
    if device is not None and torch.device(device).type == "cuda":
        t = torch.from_numpy(te).to(device)
        t.fill_diagonal_(0.0)
        e = torch.nonzero(t > thr, as_tuple=False).cpu().numpy()
        src, tgt = e[:, 0], e[:, 1]
    else:
        edges = te > thr
        np.fill_diagonal(edges, False)
        src, tgt = np.nonzero(edges)
    network: dict = {j: [] for j in range(n)}
    if len(tgt):
        order = np.argsort(tgt, kind="stable")
        ts = tgt[order]
        ss = src[order].astype(np.int64)
        splits = np.searchsorted(ts, np.arange(n + 1))
        for j in range(n):
            a, b = int(splits[j]), int(splits[j + 1])
            if a < b:
                network[j] = ss[a:b]
    return network


def _outgoing(network: dict, n: int) -> dict:
    """Reverse index source -> ascending targets (trace.py:697-708)."""
    src = [np.asarray(network[j], dtype=np.int64) for j in range(n)]
    all_src = np.concatenate(src) if n else np.zeros(0, np.int64)
    all_tgt = np.concatenate([np.full(len(s), j, np.int64) for j, s in enumerate(src)]) \
        if n else np.zeros(0, np.int64)
    out: dict = {i: [] for i in range(n)}
    if len(all_src):
        order = np.argsort(all_src, kind="stable")
        s_sorted, t_sorted = all_src[order], all_tgt[order]
        splits = np.searchsorted(s_sorted, np.arange(n + 1))
        for i in range(n):
            a, b = int(splits[i]), int(splits[i + 1])
            if a < b:
                out[i] = t_sorted[a:b]
    return out, all_src, all_tgt


def trace_scores(bins: torch.Tensor, T: int, tau: int, network: dict, outgoing: dict):
    """(OutTE, InTE) float32 (n,) of every variable on device bins (n, ld)."""
    n = bins.shape[0]
    outte = torch.zeros(n, dtype=torch.float32, device=bins.device)
    inte = torch.zeros(n, dtype=torch.float32, device=bins.device)
    for mode, conn, dst in ((0, outgoing, outte), (1, network, inte)):
        ids = [i for i in range(n) if len(conn[i]) > 0]
        if ids:
            te = core.trace_te(bins, T, tau, mode, np.asarray(ids, np.int32),
                               [conn[i] for i in ids])
            dst[core.h2d(np.asarray(ids, np.int64), bins.device)] = te
    return outte, inte


class TRACEMethod(InferenceMethod):
    """method='trace' (tenex/inference/trace.py:615-760)."""

    @property
    def name(self) -> str:
        return "trace"

    def infer(self, te_matrix, variable_names, device, **kwargs) -> TRACEResult:
        """kwargs: bin_data (n, T) or (n, T, 1) discretised series (required),
        tau (1), n_surrogates (100), significance (2.0), seed (42).  fdr,
        links and sources are refused, as in the reference."""
        bin_data = kwargs.get("bin_data", None)
        if bin_data is None:
            raise ValueError(
                "TRACE requires 'bin_data' (discretized time series). "
                "Pass the bin arrays used for TE computation.")
        for ignored in ("fdr", "links", "sources"):
            if ignored in kwargs:
                raise TypeError(
                    f"TRACE does not accept {ignored!r}: it prunes the network via a "
                    f"marginal TE threshold (n_surrogates, significance). When invoking "
                    f"via NetWeaver these are dropped automatically -- remove the explicit "
                    f"kwarg.")
        tau = int(kwargs.get("tau", 1))
        n_surrogates = int(kwargs.get("n_surrogates", 100))
        significance = float(kwargs.get("significance", 2.0))
        seed = int(kwargs.get("seed", 42))
        te_matrix = np.asarray(te_matrix, dtype=np.float32)
        b = bin_data.cpu().numpy() if isinstance(bin_data, torch.Tensor) else np.asarray(bin_data)
        if b.ndim == 3:
            if b.shape[2] != 1:
                raise ValueError(
                    f"TRACE requires K=1 bin arrays; got shape {b.shape}. Multi-kernel "
                    f"binning (e.g. FSBW-B/FSBW-T) is not supported by the TRACE descriptors.")
            b = b[:, :, 0]
        n = te_matrix.shape[0]
        if b.shape[0] != n:
            raise ValueError(f"bin_data has {b.shape[0]} variables, te_matrix {n}")
        dev = core.require_gpu(device)
        network = build_causal_network(te_matrix, n_surrogates, significance, seed, dev)
        outgoing, all_src, all_tgt = _outgoing(network, n)
        if len(all_src):
            grn = make_grn(np.asarray(variable_names), np.column_stack([all_src, all_tgt]),
                           te_matrix[all_src, all_tgt])
        else:
            grn = make_grn(np.asarray(variable_names), np.empty((0, 2), dtype=np.int64),
                           np.empty(0, dtype=np.float32))
        vprint(f"[TENEX] TRACE:{len(all_src)} causal edges")
        bins, _, _, T = core.upload_bins(b, dev)
        if T <= tau:
            raise ValueError(f"Time series length T={T} must be greater than tau={tau}")
        outte, inte = trace_scores(bins, T, tau, network, outgoing)
        outte, inte = outte.cpu().numpy(), inte.cpu().numpy()
        vprint(f"[TENEX] TRACE:OutTE non-zero: {(outte > 0).sum()}/{n}, "
               f"InTE non-zero: {(inte > 0).sum()}/{n}")
        return TRACEResult(outte=outte, inte=inte, variable_names=variable_names,
                           network=network, grn=grn)
