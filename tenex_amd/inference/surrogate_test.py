"""Time-axis shuffle surrogate test (tenex/inference/surrogate_test.py:586-948).

For every ordered pair, n_surrogates block-shuffled copies of the data give a
null distribution of TE; effective TE = observed - mean(surrogate TE); a
per-pair Gaussian z-test (parametric) or the Monte Carlo rank (mc) gives
p-values; BH-FDR over the tested pairs and effective TE > 0 give the GRN.

Paths:
  default          per surrogate, tnx_block_shuffle writes a block-shuffled
                   copy of the bins to HBM, tnx_gene_prep rebuilds its
                   states and entropies, and the TE kernels (tnx_te_run)
                   add each pair's float32 TE_k to FP64 statistics in their
                   epilogue (TNX_OUT_STATS).  When the tested pairs touch
                   only some variables (targets / sources given), only those
                   rows are shuffled and re-prepared, with the same block
                   orders (rows of the full draw).
  fused=True       the fused kernel surrogate_wave_kernel shuffles on the fly
                   and counts every (pair, surrogate) histogram itself (the
                   role of the reference's fused Full/Adaptive-SMEM
                   surrogate kernels); it can also return every TE_k.
  fused=False      loop path: per surrogate, shuffle the bins on the device
                   (torch gather) and run the TE kernels on the shuffled data,
                   accumulating in FP64 (the reference's loop path,
                   surrogate_test.py:322-468).  Same numbers, kept as an
                   independent cross-check.
  learned='mlp'    learned null (no reference counterpart): see
                   tenex_amd/inference/learned_null.py.

Permutations follow the reference schedule exactly: surrogate k uses
torch.Generator(rng_device).manual_seed(seed + k), rand(n_vars, n_blocks),
argsort (surrogate_test.py:165-207); rng_device defaults to the compute
device as in the reference's GPU path.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from .. import core
from .._log import vprint
from ..kernels import get_kernel
from . import GRN, InferenceMethod, bh_fdr_gpu, bh_reject_gpu, build_pairs, make_grn


_MATRIX_FIELDS = frozenset(("observed_te", "mean_surrogate_te", "std_surrogate_te",
                            "effective_te", "p_values"))


@dataclass
class SurrogateTestResult:
    """observed_te, mean_surrogate_te, std_surrogate_te, effective_te,
    p_values: (n, n) float32 in the public layout [i, j] = TE(i -> j);
    grn; n_surrogates; shuffle_method; block_length; p_method; fdr
    (tenex/inference/surrogate_test.py:74-104).

    The matrices may be handed over as device tensors: they stay in device
    memory until first read, which copies them to numpy once."""
    observed_te: np.ndarray
    mean_surrogate_te: np.ndarray
    std_surrogate_te: np.ndarray
    effective_te: np.ndarray
    p_values: np.ndarray
    grn: GRN
    n_surrogates: int
    shuffle_method: str
    block_length: int | None
    p_method: str
    fdr: float

    def __getattribute__(self, name):
        v = object.__getattribute__(self, name)
        if name in _MATRIX_FIELDS and isinstance(v, torch.Tensor):
            v = v.cpu().numpy()
            object.__setattr__(self, name, v)
        return v


def shuffle_bins(bins: torch.Tensor, T: int, perm_k: torch.Tensor, eff_bl: int) -> torch.Tensor:
    """Shuffled copy of dense bins (n, T) by per-variable block order perm_k
    (n, n_blocks) (_block_shuffle_gpu / _random_shuffle_gpu,
    surrogate_test.py:107-162): head blocks reordered, tail untouched."""
    n = bins.shape[0]
    nb = perm_k.shape[1]
    head = nb * eff_bl
    src = bins[:, :T]
    h = src[:, :head].reshape(n, nb, eff_bl)
    idx = perm_k.to(torch.int64)[:, :, None].expand(-1, -1, eff_bl)
    out = torch.gather(h, 1, idx).reshape(n, head)
    if T > head:
        out = torch.cat([out, src[:, head:]], dim=1)
    return out


def surrogate_accumulate(prep: core.Prepared, observed_pub: torch.Tensor, n_surrogates: int,
                         shuffle_method: str, block_length: int, seed: int, rng_device,
                         fused: bool | None = None, targets=None, sources=None,
                         perm_batch: int = 32, k_start: int = 0,
                         te_k_dbg: torch.Tensor | None = None):
    """Device (n, n) public-layout sum, sumsq (float64) and count_ge (int32)
    over surrogates [k_start, k_start + n_surrogates).  observed_pub None
    (default path only): count_ge is not counted (left zero).

    fused=None (default): per surrogate a block-shuffled copy of the bins in
    HBM through the TE kernels with their statistics epilogue.  fused=True:
    the fused surrogate kernel (shuffles on the fly; te_k_dbg available).
    fused=False: the reference's loop (torch gather, TE matrix, torch sums).
    The first two leave the diagonal untouched."""
    n, T = prep.n_vars, prep.T
    dev = prep.device
    s = torch.zeros((n, n), dtype=torch.float64, device=dev)
    s2 = torch.zeros((n, n), dtype=torch.float64, device=dev)
    cg = torch.zeros((n, n), dtype=torch.int32, device=dev)
    k_end = k_start + n_surrogates
    if fused is None and te_k_dbg is None:
        every = np.arange(n, dtype=np.int64)
        tg = every if targets is None else np.asarray(targets, dtype=np.int64)
        sr = every if sources is None else np.asarray(sources, dtype=np.int64)
        involved = np.union1d(tg, sr)
        if len(involved) < n:
            # shuffle and re-prepare only the variables the pairs touch
            plan = core.get_te_plan(core.var_subset(prep, involved),
                                    np.searchsorted(involved, tg).astype(np.int32),
                                    np.searchsorted(involved, sr).astype(np.int32))
            ss, ss2, scg = core.surrogate_stats_subset(prep, involved, plan, observed_pub,
                                                       n_surrogates, shuffle_method,
                                                       block_length, seed, rng_device, k_start,
                                                       perm_batch)
            vd = core.var_subset(prep, involved).plans["rows"]
            s[vd[:, None], vd[None, :]] = ss
            s2[vd[:, None], vd[None, :]] = ss2
            cg[vd[:, None], vd[None, :]] = scg
            return s, s2, cg
        plan = core.get_te_plan(prep, targets, sources)
        work = {}
        kb = max(perm_batch, core.copies_per_launch(prep, n_surrogates, n * n))
        for k0 in range(k_start, k_end, kb):
            k1 = min(k0 + kb, k_end)
            perm, eff_bl = core.copy_perms(n, T, shuffle_method, block_length, k0, k1, k_start,
                                           k_end, seed, rng_device, dev, perm_batch=perm_batch)
            core.surrogate_stats_materialized(prep, plan, perm, eff_bl, observed_pub, s, s2,
                                              None if observed_pub is None else cg, work)
        return s, s2, cg
    if fused is None or fused:
        plan = core.surrogate_plan(prep, targets, sources)
        for k0 in range(k_start, k_end, perm_batch):
            k1 = min(k0 + perm_batch, k_end)
            perm, eff_bl, nb = core.block_perm_device(n, T, shuffle_method, block_length, k0, k1,
                                                      seed, rng_device, dev)
            dbg = None if te_k_dbg is None else te_k_dbg[:, k0 - k_start:k1 - k_start]
            if dbg is not None:
                tmp = torch.zeros((n * n, k1 - k0), dtype=torch.float32, device=dev)
                core.surrogate_run(prep, plan, perm, eff_bl, observed_pub, s, s2, cg, tmp)
                dbg.copy_(tmp)
            else:
                core.surrogate_run(prep, plan, perm, eff_bl, observed_pub, s, s2, cg)
        return s, s2, cg
    # loop path: shuffle then the TE kernels
    for k in range(k_start, k_end):
        perm, eff_bl, nb = core.block_perm_device(n, T, shuffle_method, block_length, k, k + 1,
                                                  seed, rng_device, dev)
        shuf = shuffle_bins(prep.bins, T, perm[0], eff_bl)
        sb = torch.zeros_like(prep.bins)
        sb[:, :T] = shuf
        sp = core.prepare(sb, prep.npv, T, prep.tau)
        te_int = core.te_matrix_internal(sp, targets, sources)
        te_k = te_int.t()  # public layout
        te_d = te_k.to(torch.float64)
        s += te_d
        s2 += te_d * te_d
        cg += (te_k >= observed_pub).to(torch.int32)
    return s, s2, cg


def surrogate_decision(observed_te, sum_s, sumsq_s, count_ge, n_surr, p_method, fdr, mask,
                       device, host: bool = True):
    """Statistics -> p-values -> BH -> edges with the reference's arithmetic
    (surrogate_test.py:890-929): float32 mean/std from FP64 sums, z in
    float32, erfc in float64, BH in float32.  Inputs may be numpy arrays or
    tensors; the arithmetic runs on `device`.  Returns (mean, std, eff, p,
    sig) as numpy arrays, or as tensors on `device` when host=False."""
    dev = torch.device(device)
    t = lambda a, dt: (a if isinstance(a, torch.Tensor) else torch.from_numpy(np.asarray(a))).to(
        device=dev, dtype=dt)
    obs = t(observed_te, torch.float32)
    S = t(sum_s, torch.float64)
    S2 = t(sumsq_s, torch.float64)
    mean64 = S / n_surr
    mean = mean64.to(torch.float32)
    var = S2 / n_surr - mean64 * mean64
    std = torch.sqrt(torch.clamp(var, min=0.0)).to(torch.float32)
    eff = obs - mean
    if p_method == "parametric":
        z = (obs - mean) / torch.clamp(std, min=1e-12)
        p = (0.5 * torch.special.erfc(z.to(torch.float64) / np.sqrt(2.0))).to(torch.float32)
    else:
        cg = t(count_ge, torch.float64)
        p = ((1.0 + cg) / (1.0 + n_surr)).to(torch.float32)
    p.fill_diagonal_(1.0)
    m = t(mask, torch.bool)
    sig = torch.zeros_like(m)
    sig[m] = bh_reject_gpu(p[m], fdr)
    sig &= eff > 0
    out = (mean, std, eff, p, sig)
    return tuple(x.cpu().numpy() for x in out) if host else out


def surrogate_decision_pairs(observed_te, pairs, sum_s, sumsq_s, n_surr, fdr):
    """surrogate_decision (parametric, on the device) for sums given only at
    the tested pairs: pairs (P, 2) [source, target] device int64 (distinct,
    off the diagonal), sum_s / sumsq_s (P,) float64.  The dense decision's
    arithmetic on vectors: untested entries have zero sums, so their mean
    and std are 0, eff = obs - 0 and z = eff / clamp(0, 1e-12), computed
    over the dense matrix as there; the tested entries are computed on the
    P vectors and scattered.  BH runs over the tested p-values (its
    decisions do not depend on their order).  Returns the same (mean, std,
    eff, p, sig) tensors as surrogate_decision(host=False) with the sums
    scattered to dense matrices and mask = the pairs."""
    obs = observed_te.to(torch.float32)
    dev = obs.device
    src, tgt = pairs[:, 0], pairs[:, 1]
    mean = torch.zeros_like(obs)
    std = torch.zeros_like(obs)
    eff = obs - mean
    tiny = torch.clamp(torch.zeros(1, dtype=torch.float32, device=dev), min=1e-12)
    z = eff / tiny
    p = (0.5 * torch.special.erfc(z.to(torch.float64) / np.sqrt(2.0))).to(torch.float32)
    S = sum_s.to(torch.float64)
    S2 = sumsq_s.to(torch.float64)
    ov = obs[src, tgt]
    mean64 = S / n_surr
    mv = mean64.to(torch.float32)
    var = S2 / n_surr - mean64 * mean64
    sv = torch.sqrt(torch.clamp(var, min=0.0)).to(torch.float32)
    ev = ov - mv
    zv = (ov - mv) / torch.clamp(sv, min=1e-12)
    pv = (0.5 * torch.special.erfc(zv.to(torch.float64) / np.sqrt(2.0))).to(torch.float32)
    mean[src, tgt] = mv
    std[src, tgt] = sv
    eff[src, tgt] = ev
    p[src, tgt] = pv
    p.fill_diagonal_(1.0)
    sig = torch.zeros(obs.shape, dtype=torch.bool, device=dev)
    sig[src, tgt] = bh_reject_gpu(pv, fdr) & (ev > 0)
    return mean, std, eff, p, sig


class SurrogateTestMethod(InferenceMethod):
    """Surrogate-based test with effective-TE bias correction."""

    @property
    def name(self) -> str:
        return "surrogate_test"

    def infer(self, te_matrix, variable_names, device, **kwargs) -> SurrogateTestResult:
        """kwargs as the reference (bin_data, n_per_var, b_max, kernel, tau,
        n_surrogates=100, shuffle_method='block', block_length=None,
        p_method='parametric', fdr=0.01, seed=42, fused=None, sources) plus
        rng_device (generator device, default: the compute device),
        targets (restrict the tested pairs to these target names),
        pairs_scope ('all' | 'sources': compute only the tested pairs) and
        learned (None | 'mlp', with learned_kwargs)."""
        bin_data = kwargs.get("bin_data")
        n_per_var = kwargs.get("n_per_var")
        b_max = kwargs.get("b_max")
        kernel_name = kwargs.get("kernel")
        if bin_data is None or n_per_var is None or b_max is None or not kernel_name:
            raise ValueError("surrogate_test requires bin_data, n_per_var, b_max, and kernel. "
                             "Pass a TransferEntropyResult to NetWeaver (these are auto-filled).")
        tau = int(kwargs.get("tau", 1))
        n_surrogates = int(kwargs.get("n_surrogates", 100))
        shuffle_method = kwargs.get("shuffle_method", "block")
        block_length = kwargs.get("block_length")
        p_method = kwargs.get("p_method", "parametric")
        fdr = float(kwargs.get("fdr", 0.01))
        seed = int(kwargs.get("seed", 42))
        if shuffle_method not in ("block", "random"):
            raise ValueError(f"shuffle_method must be 'block' or 'random' (got {shuffle_method!r})")
        if n_surrogates < 2:
            raise ValueError(f"n_surrogates must be >= 2 (got {n_surrogates}); parametric z-test "
                             f"needs a non-degenerate std estimate and MC p-values need at least "
                             f"two samples to be informative.")
        if p_method not in ("parametric", "mc"):
            raise ValueError(f"p_method must be 'parametric' or 'mc' (got {p_method!r})")
        if kwargs.get("learned") is not None and p_method != "parametric":
            raise ValueError(f"learned={kwargs.get('learned')!r} predicts a parametric null: "
                             f"p_method must be 'parametric' (got {p_method!r})")
        bin_data = np.asarray(bin_data)
        if bin_data.ndim == 2:
            bin_data = bin_data[..., None]
        n_vars, T, K = bin_data.shape
        if K != 1:
            raise ValueError(f"only K=1 bin arrays are supported (got K={K})")
        if T - tau < 2:
            raise ValueError(f"T - tau = {T - tau} is too short for TE computation")
        if block_length is None:
            block_length = max(2, int(np.sqrt(T - tau)))
        block_length = int(block_length)
        get_kernel(kernel_name)  # validates the name (KeyError if unknown)
        fused_kw = kwargs.get("fused", None)
        observed_te = np.asarray(te_matrix, dtype=np.float32)
        dev = core.require_gpu(device if device is not None and torch.device(device).type == "cuda"
                               else None)
        rng_device = kwargs.get("rng_device") or dev
        sources = kwargs.get("sources")
        targets = kwargs.get("targets")
        scope = kwargs.get("pairs_scope", "all")
        learned = kwargs.get("learned")
        names_arr = np.asarray(variable_names)
        s_rows = tgt_idx = None
        if sources is not None:
            # rows = the sources, every other variable a target (build_pairs)
            _, s_rows, _ = np.intersect1d(names_arr, sources, return_indices=True)
        if targets is not None:
            # extension: restrict the tested pairs to these target names too
            _, tgt_idx, _ = np.intersect1d(names_arr, targets, return_indices=True)
            tgt_idx = tgt_idx.astype(np.int32)
        src_idx = None
        if scope == "sources" and sources is not None:
            src_idx = s_rows.astype(np.int32)
        vprint(f"[TENEX] surrogate_test: {n_surrogates} shuffles (method={shuffle_method}, "
               f"block_length={block_length}, kernel={kernel_name}, p={p_method}, "
               f"device={dev})")
        bins, npv_d, _, T_ = core.upload_bins(bin_data, dev)
        # the observed matrix travels while the shuffles and TE launches are set up
        obs_up = core.PendingUpload(observed_te, dev)
        npv_in = core.h2d(np.asarray(n_per_var, dtype=np.int32), dev)
        if npv_in.shape[0] == n_vars and bool((npv_in >= npv_d).all()):
            npv_d = npv_in.contiguous()
        prep = core.prepare(bins, npv_d, T, tau)
        rows = np.sort(s_rows) if s_rows is not None else np.arange(n_vars)
        cols = np.sort(tgt_idx) if tgt_idx is not None else np.arange(n_vars)
        tested = (rows, cols) if len(rows) * len(cols) <= (1 << 24) else None

        def tested_mask():
            """(n, n) bool of the tested [source, target] pairs: rows x cols
            off the diagonal."""
            m = torch.zeros((n_vars, n_vars), dtype=torch.bool, device=dev)
            if s_rows is not None:
                m[core.h2d(s_rows.astype(np.int64), dev)] = True
            else:
                m.fill_(True)
            m.fill_diagonal_(False)
            if tgt_idx is not None:
                tm = torch.zeros(n_vars, dtype=torch.bool, device=dev)
                tm[core.h2d(tgt_idx.astype(np.int64), dev)] = True
                m &= tm[None, :]
            return m

        if learned is not None:
            from .learned_null import learned_surrogate_test
            return learned_surrogate_test(
                prep, observed_te, obs_up, variable_names, n_surrogates, shuffle_method,
                block_length, seed, rng_device, p_method, fdr,
                tested_mask if tested is not None else tested_mask(), src_idx, dev,
                tested=tested, **(kwargs.get("learned_kwargs") or {}))
        mask = tested_mask()
        tg_plan = tgt_idx if (scope == "sources" and tgt_idx is not None) else None
        # the parametric decision reads only the sums: count_ge is needed for
        # mc p-values (and the fused / loop paths always count)
        counted = p_method == "mc" or fused_kw is not None
        s, s2, cg = surrogate_accumulate(prep, obs_up.get() if counted else None, n_surrogates,
                                         shuffle_method, block_length, seed, rng_device,
                                         fused=fused_kw, targets=tg_plan, sources=src_idx)
        obs_d = obs_up.get()
        if fused_kw is not False:
            cg.diagonal().add_(n_surrogates)  # the loop path's diagonal (te_k = obs = 0)
        mean, std, eff, p, sig = surrogate_decision(obs_d, s, s2, cg, n_surrogates, p_method,
                                                    fdr, mask, dev, host=False)
        edges = torch.nonzero(sig)
        grn = make_grn(names_arr, edges.cpu().numpy().astype(np.int64),
                       eff[edges[:, 0], edges[:, 1]].cpu().numpy())
        vprint(f"[TENEX] surrogate_test: {len(grn)} significant edges (FDR<{fdr})")
        return SurrogateTestResult(
            observed_te=obs_d, mean_surrogate_te=mean, std_surrogate_te=std,
            effective_te=eff, p_values=p, grn=grn, n_surrogates=n_surrogates,
            shuffle_method=shuffle_method,
            block_length=block_length if shuffle_method == "block" else None,
            p_method=p_method, fdr=fdr)


def surrogate_test_lags(bin_data, n_per_var, variable_names, taus, device, **kwargs):
    """The surrogate test at every lag of taus with all lags in each launch
    (the observed TE of every lag in one launch, then per batch of
    surrogates one TE launch over the copies of every lag).  Per lag it
    returns what SurrogateTestMethod.infer returns for that tau with the
    observed matrix of that lag: the same block orders (block_length
    defaults to max(2, sqrt(T - tau)) per lag), the same statistics and
    decision.  Returns None for options it does not cover (a target
    restriction, the learned null, an explicit fused / loop path) or a plan
    that cannot run as one persistent launch; the caller then runs the lags
    one by one."""
    if kwargs.get("learned") is not None or kwargs.get("fused") is not None or \
            kwargs.get("targets") is not None:
        return None
    n_surrogates = int(kwargs.get("n_surrogates", 100))
    shuffle_method = kwargs.get("shuffle_method", "block")
    block_length = kwargs.get("block_length")
    p_method = kwargs.get("p_method", "parametric")
    fdr = float(kwargs.get("fdr", 0.01))
    seed = int(kwargs.get("seed", 42))
    sources = kwargs.get("sources")
    if shuffle_method not in ("block", "random") or n_surrogates < 2 or \
            p_method not in ("parametric", "mc"):
        return None   # the per-lag path raises the reference's errors
    b = np.asarray(bin_data)
    if b.ndim == 3:
        if b.shape[2] != 1:
            return None
        b = b[:, :, 0]
    n, T = b.shape
    taus = [int(t) for t in taus]
    if min(taus) < 1 or T - max(taus) < 2:
        return None
    dev = core.require_gpu(device if device is not None and torch.device(device).type == "cuda"
                           else None)
    rng_device = kwargs.get("rng_device") or dev
    bls = [int(block_length) if block_length is not None else max(2, int(np.sqrt(T - t)))
           for t in taus]
    bins, npv_d, _, _ = core.upload_bins(b, dev)
    npv_in = core.h2d(np.asarray(n_per_var, dtype=np.int32), dev)
    if npv_in.shape[0] == n and bool((npv_in >= npv_d).all()):
        npv_d = npv_in.contiguous()
    prep = core.prepare(bins, npv_d, T, min(taus))
    names_arr = np.asarray(variable_names)
    s_rows = None
    if sources is not None:
        _, s_rows, _ = np.intersect1d(names_arr, sources, return_indices=True)
    src_idx = None if s_rows is None else np.sort(s_rows).astype(np.int32)
    plan = core.surrogate_plan(prep, None, src_idx)
    obs_int = core.multilag_observed(prep, plan, taus)
    if obs_int is None:
        return None
    obs_pub = [o.t().contiguous() for o in obs_int]
    counted = p_method == "mc"
    stats = core.multilag_surrogate_stats(prep, plan, taus, bls, obs_pub if counted else None,
                                          n_surrogates, shuffle_method, seed, rng_device)
    if stats is None:
        return None
    mask = torch.zeros((n, n), dtype=torch.bool, device=dev)
    if s_rows is not None:
        mask[core.h2d(s_rows.astype(np.int64), dev)] = True
    else:
        mask.fill_(True)
    mask.fill_diagonal_(False)
    out = []
    for l, tau in enumerate(taus):
        s, s2, cg = stats[l]
        cg.diagonal().add_(n_surrogates)   # as the single-lag default path
        mean, std, eff, p, sig = surrogate_decision(obs_pub[l], s, s2, cg, n_surrogates,
                                                    p_method, fdr, mask, dev, host=False)
        edges = torch.nonzero(sig)
        grn = make_grn(names_arr, edges.cpu().numpy().astype(np.int64),
                       eff[edges[:, 0], edges[:, 1]].cpu().numpy())
        out.append(SurrogateTestResult(
            observed_te=obs_pub[l], mean_surrogate_te=mean, std_surrogate_te=std,
            effective_te=eff, p_values=p, grn=grn, n_surrogates=n_surrogates,
            shuffle_method=shuffle_method,
            block_length=bls[l] if shuffle_method == "block" else None,
            p_method=p_method, fdr=fdr))
    return out
