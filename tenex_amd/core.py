"""Device-resident data and launch wrappers over libtenex_hip (torch = memory).

Layout in HBM (one GPU):
  bins    int8   (n_vars, ld)   ld = round_up(T, 64), padding zero
  n_per_var int32 (n_vars,)
  states  uint16 (n_vars, lds)  lds = round_up(T - tau, 64): x[t+tau]*b + x[t]
  S1, S2a float64 (n_vars,)     per-target entropy sums
  xlogx   float32 (2*(N+1),)    c*log2(c) as hi/lo pairs, N = T - tau
  feat    float32 (n_vars, 8)   per-variable features of the learned null
"""
from __future__ import annotations

import ctypes as C
import hashlib
import os
import threading
from collections import OrderedDict
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from ._lib import call, ptr

ROW_ALIGN = 64


def round_up(v: int, m: int) -> int:
    return (v + m - 1) // m * m


def require_gpu(device=None) -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("tenex_amd needs a ROCm GPU (gfx950); none is available")
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    if dev.type != "cuda":
        raise RuntimeError(f"tenex_amd runs on the GPU only (got device {dev})")
    if dev.index is None:
        dev = torch.device("cuda", torch.cuda.current_device())
    _lib.load()
    return dev


def _stream(dev) -> int:
    return torch.cuda.current_stream(dev).cuda_stream


def device_props(dev) -> _lib.DeviceProps:
    p = _lib.DeviceProps()
    call("tnx_device_props", int(dev.index), C.byref(p))
    return p


# ── discretisation ────────────────────────────────────────────────────────

def fsbw_l_gpu(x: torch.Tensor, kp: float = 0.5):
    """FSBW-L with numpy-exact statistics on the GPU.  x float32 (n, T) on
    device.  Returns (raw int32 (n, T), n_bins int32 (n,))."""
    n, T = x.shape
    x = x.contiguous()
    raw = torch.empty((n, T), dtype=torch.int32, device=x.device)
    nb = torch.empty(n, dtype=torch.int32, device=x.device)
    sd = torch.empty(n, dtype=torch.float32, device=x.device)
    sm = torch.empty(n, dtype=torch.float32, device=x.device)
    call("tnx_fsbw_l", ptr(x), n, T, float(kp), ptr(raw), ptr(nb), ptr(sd), ptr(sm),
         _stream(x.device))
    return raw, nb


def remap_gpu(raw: torch.Tensor):
    """Dense per-variable relabelling into padded int8 rows.
    Returns (bins int8 (n, ld), n_per_var int32 (n,) device, b_max)."""
    n, T = raw.shape
    ld = round_up(T, ROW_ALIGN)
    bins = torch.empty((n, ld), dtype=torch.int8, device=raw.device)
    npv = torch.empty(n, dtype=torch.int32, device=raw.device)
    status = torch.zeros(1, dtype=torch.int32, device=raw.device)
    call("tnx_remap_dense", ptr(raw.contiguous()), n, T, ptr(bins), ld, ptr(npv), ptr(status),
         _stream(raw.device))
    st = int(status.item())
    if st & 1:
        raise ValueError("remap: a raw bin index is negative or >= 65536")
    if st & 2:
        raise ValueError("remap: a variable has more than 127 distinct levels; "
                         "int8 bins cannot hold it (use coarsening or a smaller kp)")
    b_max = int(npv.max().item()) if n else 0
    return bins, npv, b_max


def remap_any_gpu(raw: torch.Tensor):
    """Dense relabelling into int8 rows, or, when a variable has more than
    127 levels, into int16 rows as well (the reference's int16 bins).
    Returns (bins int8 (n, ld), bins16 int16 (n, ld) or None, n_per_var,
    b_max)."""
    try:
        bins, npv, b_max = remap_gpu(raw)
        return bins, None, npv, b_max
    except ValueError as e:
        if "more than 127" not in str(e):
            raise
    b16, b8, npv, b_max = remap16_gpu(raw)
    return b8, b16, npv, b_max


def upload_bins_any(bin_arr, device):
    """upload_bins for dense bins of up to 32,767 levels: returns (bins int8,
    bins16 int16 or None (all levels <= 127), npv, b_max, T)."""
    t = torch.as_tensor(np.asarray(bin_arr) if not isinstance(bin_arr, torch.Tensor) else bin_arr)
    if t.dim() == 3:
        if t.shape[2] != 1:
            raise ValueError(f"only K=1 bin arrays are supported (got K={t.shape[2]})")
        t = t[:, :, 0]
    if t.numel() and int(t.max().item()) > MAX_LEVELS_INT8 - 1:
        if t.is_floating_point() or int(t.min().item()) < 0 or \
                int(t.max().item()) > MAX_LEVELS_INT16 - 1:
            raise ValueError(f"bins must satisfy 0 <= bin < n_per_var <= {MAX_LEVELS_INT16}")
        n, T = t.shape
        ld = round_up(T, ROW_ALIGN)
        b16 = torch.zeros((n, ld), dtype=torch.int16, device=device)
        b16[:, :T] = t.to(device=device, dtype=torch.int16)
        b8 = b16.clamp(max=MAX_LEVELS_INT8).to(torch.int8)
        npv = (b16[:, :T].to(torch.int32).amax(dim=1) + 1).to(torch.int32).contiguous()
        return b8, b16, npv, int(npv.max().item()), T
    bins, npv, b_max, T = upload_bins(t, device)
    return bins, None, npv, b_max, T


def remap16_gpu(raw: torch.Tensor):
    """Dense relabelling into padded int16 rows (variables of up to 32,767
    levels).  Returns (bins16 int16 (n, ld), bins int8 (n, ld) with levels
    above 127 clamped, n_per_var, b_max)."""
    n, T = raw.shape
    ld = round_up(T, ROW_ALIGN)
    b16 = torch.empty((n, ld), dtype=torch.int16, device=raw.device)
    npv = torch.empty(n, dtype=torch.int32, device=raw.device)
    status = torch.zeros(1, dtype=torch.int32, device=raw.device)
    call("tnx_remap_dense16", ptr(raw.contiguous()), n, T, ptr(b16), ld, ptr(npv), ptr(status),
         _stream(raw.device))
    st = int(status.item())
    if st & 1:
        raise ValueError("remap: a raw bin index is negative or >= 65536")
    if st & 4:
        raise ValueError(f"remap: a variable has more than {MAX_LEVELS_INT16} distinct levels")
    b8 = b16.clamp(max=MAX_LEVELS_INT8).to(torch.int8)
    return b16, b8, npv, int(npv.max().item()) if n else 0


def upload_bins(bin_arr, device) -> tuple[torch.Tensor, torch.Tensor, int, int]:
    """Dense bins (n, T) or (n, T, 1) from the host or device -> padded int8
    device rows.  Bins must already be dense per variable (0..b-1 with every
    level present) as remap_bins produces.  Returns (bins, npv, b_max, T)."""
    t = torch.as_tensor(np.asarray(bin_arr) if not isinstance(bin_arr, torch.Tensor) else bin_arr)
    if t.dim() == 3:
        if t.shape[2] != 1:
            raise ValueError(f"only K=1 bin arrays are supported (got K={t.shape[2]})")
        t = t[:, :, 0]
    n, T = t.shape
    if t.is_floating_point() or t.dtype == torch.bool or t.is_complex():
        raise ValueError(f"bins must be an integer array (got {t.dtype})")
    ld = round_up(T, ROW_ALIGN)
    bins = torch.zeros((n, ld), dtype=torch.int8, device=device)
    if n and T:
        # the range is checked on the device, after the upload, with one read back
        td = t.to(device=device, non_blocking=t.is_pinned())
        mm = torch.stack([td.min(), td.max()]).to(torch.int64)
        bins[:, :T] = td.to(torch.int8)
        npv = (bins[:, :T].to(torch.int32).amax(dim=1) + 1).to(torch.int32)
        lo, hi, b_max = (int(v) for v in torch.cat([mm, npv.max().view(1).to(torch.int64)]).cpu())
        if lo < 0 or hi > 126:
            raise ValueError(f"bins must satisfy 0 <= bin < n_per_var <= 127 (got values in "
                             f"[{lo}, {hi}]); int8 bins cannot hold them")
    else:
        npv = torch.ones(n, dtype=torch.int32, device=device)
        b_max = 1 if n else 0
    return bins, npv.contiguous(), b_max, T


def coarsen_gpu(bins: torch.Tensor, npv: torch.Tensor, T: int, limit: int):
    call("tnx_coarsen_bins", ptr(bins), bins.shape[0], T, bins.shape[1], ptr(npv), int(limit),
         _stream(bins.device))


# ── per-variable preparation ──────────────────────────────────────────────

_xlogx_cache: dict = {}


def xlogx_table(N: int, device) -> torch.Tensor:
    key = (int(N), str(device))
    t = _xlogx_cache.get(key)
    if t is None:
        t = torch.empty(2 * (N + 1), dtype=torch.float32, device=device)
        call("tnx_xlogx_table", ptr(t), int(N), _stream(device))
        _xlogx_cache[key] = t
    return t


_xlogx_fx_cache: dict = {}


def fx_shift(N: int) -> int:
    """Fraction bits of the fixed-point c*log2(c) table for series of N
    samples: sums of up to N log2 N stay below 2^60."""
    return int(min(48, max(16, 55 - int(np.ceil(np.log2(N + 1))))))


def xlogx_fx_table(N: int, device) -> torch.Tensor:
    """int64 (N + 1,) fixed-point c*log2(c) with fx_shift(N) fraction bits."""
    key = (int(N), str(device))
    t = _xlogx_fx_cache.get(key)
    if t is None:
        t = torch.empty(N + 1, dtype=torch.int64, device=device)
        call("tnx_xlogx_table_fx", ptr(t), int(N), fx_shift(N), _stream(device))
        _xlogx_fx_cache[key] = t
    return t


@dataclass
class Prepared:
    """Device-resident discretised data ready for the TE kernels."""
    bins: torch.Tensor       # int8 (n, ld)
    npv: torch.Tensor        # int32 (n,)
    npv_host: np.ndarray     # int32 (n,)
    T: int
    tau: int
    states: torch.Tensor     # uint16 as int16 storage (n, lds)
    S1: torch.Tensor
    S2a: torch.Tensor
    feat: torch.Tensor       # float32 (n, 8)
    xlogx: torch.Tensor
    plans: dict = field(default_factory=dict)
    hot_zero: bool = False   # key 0 frequent enough to divert (see key0_fraction)
    bins16: torch.Tensor | None = None   # int16 (n, ld) when some variable has > 127 levels

    @property
    def wide(self) -> np.ndarray:
        """Indices of the variables with more than 127 levels."""
        return np.nonzero(self.npv_host > MAX_LEVELS_INT8)[0]

    @property
    def n_vars(self) -> int:
        return self.bins.shape[0]

    @property
    def ld(self) -> int:
        return self.bins.shape[1]

    @property
    def lds(self) -> int:
        return self.states.shape[1]

    @property
    def L(self) -> int:
        return self.T - self.tau

    @property
    def device(self):
        return self.bins.device

    @property
    def hot(self) -> tuple[int, int]:
        """(hot_cw, hot_maxb) of the persistent launch: with a frequent key 0
        (zero-inflated data) and at most HOT_MAX_LEVELS levels everywhere,
        the samples of each target's four hot states (both levels below 2)
        are counted in HOT_CW lane-private copies per cell; else (0, 0)."""
        bmax = int(self.npv_host.max()) if len(self.npv_host) else 0
        if self.hot_zero and HOT_CW > 0 and 1 <= bmax <= HOT_MAX_LEVELS:
            return HOT_CW, bmax
        return 0, 0

    @property
    def hot_words(self) -> int:
        """LDS words a wave slice adds for the hot cells (0 when off)."""
        cw, mb = self.hot
        return 4 * cw * mb


HOT_CW = 4               # wave-slice copies of a hot cell (4, 8 or 16; 0: off)
HOT_CB = 16              # block-role copies of a hot cell (te_args.hpp kHotCB)
HOT_MAX_LEVELS = 32      # the hot cells hold source levels below this


def wave_cap(prep: "Prepared") -> int:
    """Cells of a wave histogram for prep: WAVE_CAP less the hot cells'
    words, so that a slice takes the LDS it takes without them."""
    return WAVE_CAP - prep.hot_words


MAX_LEVELS_INT8 = 127     # levels an int8 bin row holds (the histogram kernels' bound)
MAX_LEVELS_INT16 = 32767  # levels of int16 rows (the occupied-key kernels)


def prepare(bins: torch.Tensor, npv: torch.Tensor, T: int, tau: int,
            bins16: torch.Tensor | None = None) -> Prepared:
    """Per-variable preparation (tnx_gene_prep).  With variables of more than
    127 levels, bins16 (int16 rows, same ld) holds every variable's levels
    (taken from bins when absent): their S1 / S2a come from
    tnx_gene_prep_sort and their pairs run on the occupied-key kernel (the
    int8 rows of those variables are not read)."""
    n, ld = bins.shape
    if not (1 <= tau < T):
        raise ValueError(f"need 1 <= tau < T (tau={tau}, T={T})")
    dev = bins.device
    L = T - tau
    lds = round_up(L, ROW_ALIGN)
    states = torch.empty((n, lds), dtype=torch.int16, device=dev)
    S1 = torch.empty(n, dtype=torch.float64, device=dev)
    S2a = torch.empty(n, dtype=torch.float64, device=dev)
    feat = torch.empty((n, 8), dtype=torch.float32, device=dev)
    npv_host = npv.cpu().numpy().astype(np.int32)
    wide = np.nonzero(npv_host > MAX_LEVELS_INT8)[0]
    if len(wide) and bins16 is None:
        bins16 = bins.to(torch.int16)   # level counts above the values present
    call("tnx_gene_prep", ptr(bins), n, T, ld, int(tau), ptr(npv), ptr(states), lds, ptr(S1),
         ptr(S2a), ptr(feat), int(min(npv_host.max(initial=1), MAX_LEVELS_INT8)), _stream(dev))
    if len(wide):
        if L > TRACE_MAX_L:
            raise NotImplementedError(f"variables of more than {MAX_LEVELS_INT8} levels need "
                                      f"T - tau <= {TRACE_MAX_L} (got {L})")
        wd = h2d(wide.astype(np.int64), dev)
        states.index_fill_(0, wd, 0)
        feat.index_fill_(0, wd, 0.0)
        # device arrays held in locals until the launch is queued: a freed
        # temporary's memory could go to the table allocated in between
        tbl = xlogx_fx_table(L, dev)
        wide_d = h2d(wide.astype(np.int32), dev)
        call("tnx_gene_prep_sort", ptr(bins16), n, T, ld, int(tau), ptr(wide_d), len(wide),
             ptr(tbl), fx_shift(L), ptr(S1), ptr(S2a), _stream(dev))
    return Prepared(bins=bins, npv=npv, npv_host=npv_host, T=T,
                    tau=tau, states=states, S1=S1, S2a=S2a, feat=feat,
                    xlogx=xlogx_table(L, dev),
                    hot_zero=key0_fraction(states, bins, L) >= HOT_ZERO_FRACTION,
                    bins16=bins16)


def require_narrow(prep: Prepared, what: str):
    """Raise for paths that read int8 rows only when prep has wide variables."""
    if len(prep.wide):
        raise NotImplementedError(
            f"{what} takes variables of at most {MAX_LEVELS_INT8} levels (this data has "
            f"{len(prep.wide)} above); use coarsening=True or a smaller kp")


def te_sort_pairs(prep: Prepared, pairs: np.ndarray, out: torch.Tensor, ld_out: int):
    """TE of [target, source] pairs (host int, target != source) by the
    occupied-key kernel into out[target * ld_out + source] (ld_out > 0) or
    out[p] (ld_out == 0)."""
    pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
    if len(pairs) == 0:
        return
    if prep.L > TRACE_MAX_L:
        raise NotImplementedError(f"the occupied-key kernel takes T - tau <= {TRACE_MAX_L}")
    b16 = prep.bins16 if prep.bins16 is not None else prep.bins.to(torch.int16)
    tbl = xlogx_fx_table(prep.L, prep.device)
    pairs_d = h2d(pairs, prep.device)      # held until the launch is queued
    call("tnx_te_sort_pairs", ptr(b16), prep.n_vars, prep.T, prep.ld, prep.tau, ptr(pairs_d),
         len(pairs), ptr(prep.S1), ptr(prep.S2a), ptr(tbl), fx_shift(prep.L), ptr(out),
         int(ld_out), _stream(prep.device))


# A histogram key held by a fraction p of the samples puts about 64 p lanes
# of every wavefront atomic on one LDS address; at p >= 1/16 (four lanes)
# diverting it pays for the two extra VALU operations per sample.
HOT_ZERO_FRACTION = 1.0 / 16.0


def key0_fraction(states: torch.Tensor, bins: torch.Tensor, L: int) -> float:
    """Estimated share of samples with histogram key 0 (target state 0 and
    source level 0) over all ordered pairs: mean P(state 0) * mean P(level 0),
    two device reductions.  Zero-inflated expression data reach 0.2-0.3;
    near-uniform levels stay below 0.03."""
    if states.shape[0] == 0 or L <= 0:
        return 0.0
    ps = (states[:, :L] == 0).float().mean()
    pb = (bins[:, :L] == 0).float().mean()
    return float((ps * pb).item())


# ── plan ─────────────────────────────────────────────────────────────────

@dataclass
class Plan:
    info: _lib.PlanInfo
    tgt_order: torch.Tensor
    src_order: torch.Tensor
    work: torch.Tensor  # int32 (n_items, 4)
    work_host: np.ndarray
    tgt_host: np.ndarray
    src_host: np.ndarray
    cache: dict = field(default_factory=dict)   # derived data (pair slots, copies)


def h2d(a, device) -> torch.Tensor:
    """Host array -> device tensor without a host wait: staged through a new
    pinned buffer (torch's pinned-memory cache keeps it until the copy is
    done) and copied asynchronously on the current stream."""
    t = torch.from_numpy(np.ascontiguousarray(a))
    if torch.device(device).type != "cuda":
        return t.to(device)
    return t.pin_memory().to(device, non_blocking=True)


def d2h_pinned(t: torch.Tensor) -> np.ndarray:
    """Device tensor -> host numpy array backed by pinned memory (the array
    keeps the pinned buffer alive), so a later upload of the same array
    (PendingUpload) needs no staging copy."""
    out = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
    out.copy_(t)
    return out.numpy()


_SIDE_STREAMS: dict = {}


class PendingUpload:
    """A host array's copy to the device that overlaps other work: a
    background thread starts the copy on a side stream (default priority);
    get() waits for the thread and makes the current stream wait for the
    copy, then returns the device tensor.  An array in pinned memory (as
    the engine returns its TE matrix, d2h_pinned) is copied directly;
    others are staged in pinned memory first."""

    def __init__(self, a, device):
        self.device = torch.device(device)
        self._host = torch.from_numpy(np.ascontiguousarray(a))
        self._out = None
        self._err = None
        self._ev = torch.cuda.Event()
        st = _SIDE_STREAMS.get(str(self.device))
        if st is None:
            st = torch.cuda.Stream(self.device)
            _SIDE_STREAMS[str(self.device)] = st
        self._stream = st
        self._thread = threading.Thread(target=self._run, daemon=True)
        self._thread.start()

    CHUNK = 16 << 20   # bytes per copy: small copies of other streams slip in between

    def _run(self):
        try:
            h = self._host
            pinned = h if h.is_pinned() else h.pin_memory()
            with torch.cuda.device(self.device), torch.cuda.stream(self._stream):
                out = torch.empty(h.shape, dtype=h.dtype, device=self.device)
                src, dst = pinned.reshape(-1), out.reshape(-1)
                step = max(1, self.CHUNK // max(1, h.element_size()))
                for o in range(0, src.numel(), step):
                    dst[o:o + step].copy_(src[o:o + step], non_blocking=True)
                self._out = out
                self._ev.record(self._stream)
        except BaseException as e:   # re-raised by get()
            self._err = e

    def get(self) -> torch.Tensor:
        if self._thread is not None:
            self._thread.join()
            self._thread = None
            if self._err is not None:
                raise self._err
            cur = torch.cuda.current_stream(self.device)
            cur.wait_event(self._ev)
            self._out.record_stream(cur)
        return self._out


def _ranges(beg: np.ndarray, end: np.ndarray) -> np.ndarray:
    """Concatenation of arange(beg[i], end[i])."""
    ln = np.maximum(end - beg, 0)
    tot = int(ln.sum())
    if tot == 0:
        return np.zeros(0, np.int64)
    start = np.repeat(beg - np.concatenate([[0], np.cumsum(ln)[:-1]]), ln)
    return start + np.arange(tot, dtype=np.int64)


def plan_pairs(plan: Plan) -> tuple[np.ndarray, np.ndarray]:
    """(target, source) variable indices of every pair of the plan's items
    (target != source), item order."""
    w = plan.work_host.astype(np.int64).reshape(-1, 4)
    xs, ys = [], []
    for j in range(int(w[:, 1].max(initial=0))):
        it = w[w[:, 1] > j]
        src_pos = _ranges(it[:, 2], it[:, 3])
        tgt_pos = np.repeat(it[:, 0] + j, it[:, 3] - it[:, 2])
        xs.append(plan.tgt_host[tgt_pos].astype(np.int64))
        ys.append(plan.src_host[src_pos].astype(np.int64))
    x = np.concatenate(xs) if xs else np.zeros(0, np.int64)
    y = np.concatenate(ys) if ys else np.zeros(0, np.int64)
    keep = x != y
    return x[keep], y[keep]


def replicate_plan(plan: Plan, K: int, m: int) -> Plan:
    """The plan's pairs for K copies of an m-variable problem stacked as
    variables k * m + j: copy k's items read the order arrays at offsets
    k * n_tgt and k * n_src.  Wave items of every copy come first."""
    i = plan.info
    nw = int(i.n_wave_items)
    nt, ns = len(plan.tgt_host), len(plan.src_host)
    w = plan.work_host.astype(np.int64).reshape(-1, 4)
    # the base plan's items were sized to fill the chip alone; K copies
    # fill it with K times fewer, longer items (one-target items, which a
    # wave takes whole, at most WAVE_ITEM_PAIRS pairs)
    pw = lambda q: int((q[:, 1] * (q[:, 3] - q[:, 2])).sum())
    one_target = nw > 0 and bool(np.all(w[:nw, 1] == 1))
    cap = WAVE_ITEM_PAIRS if one_target else 256
    wave = _merge_items(w[:nw], int(np.clip(-(-K * pw(w[:nw]) // FILL_WAVE_ITEMS), 4, cap)))
    # wave items in the planner's L2 order: source chunk dealt round-robin to
    # the XCD slices, then chunk, then target group (te_hist.hip, by_chunk)
    ck = wave[:, 2] // CHUNK_SRC
    wave = wave[np.lexsort((wave[:, 0], ck, ck % N_XCDS))]
    block = _merge_items(w[nw:], int(np.clip(-(-K * pw(w[nw:]) // FILL_BLOCK_ITEMS), 1, 16)))
    nw_m = len(wave)
    ks = np.arange(K, dtype=np.int64)
    reps = []
    for part in (wave, block):
        q = np.repeat(part[None], K, axis=0)          # (K, items, 4)
        q[:, :, 0] += (ks * nt)[:, None]
        q[:, :, 2:] += (ks * ns)[:, None, None]
        reps.append(q.reshape(-1, 4))
    work = np.concatenate(reps).astype(np.int32)
    tgt = (plan.tgt_host[None, :].astype(np.int64) + (ks * m)[:, None]).reshape(-1).astype(np.int32)
    src = (plan.src_host[None, :].astype(np.int64) + (ks * m)[:, None]).reshape(-1).astype(np.int32)
    info = _lib.PlanInfo.from_buffer_copy(i)
    info.n_wave_items, info.n_block_items = nw_m * K, len(block) * K
    info.n_pairs, info.n_tgt, info.n_src = i.n_pairs * K, len(tgt), len(src)
    dev = plan.work.device
    wd = work if len(work) else np.zeros((1, 4), np.int32)
    return Plan(info=info, tgt_order=h2d(tgt, dev), src_order=h2d(src, dev), work=h2d(wd, dev),
                work_host=work, tgt_host=tgt, src_host=src)


def _merge_items(items: np.ndarray, per_pairs: int) -> np.ndarray:
    """Merge work items (tgt_beg, cnt, src_beg, src_end) of the same target
    group over contiguous source ranges into items of about per_pairs pairs
    (items are never split)."""
    if len(items) == 0:
        return items.reshape(0, 4)
    q = items[np.lexsort((items[:, 2], items[:, 1], items[:, 0]))]
    brk = np.ones(len(q), bool)
    brk[1:] = (q[1:, 0] != q[:-1, 0]) | (q[1:, 1] != q[:-1, 1]) | (q[1:, 2] != q[:-1, 3])
    run = np.cumsum(brk) - 1
    run_start = q[brk, 2][run]
    ch = np.maximum(per_pairs // q[:, 1], 1)
    chunk = (q[:, 2] - run_start) // ch
    key = run * (int(chunk.max()) + 1) + chunk
    first = np.ones(len(q), bool)
    first[1:] = key[1:] != key[:-1]
    idx = np.nonzero(first)[0]
    out = q[idx].copy()
    out[:, 3] = np.maximum.reduceat(q[:, 3], idx)
    return out


def make_plan(npv_host: np.ndarray, L: int, device, targets=None, sources=None,
              cap_wave: int = 0, cap_block: int = 0, src_skip=None) -> Plan:
    """Histogram plan (tnx_te_plan).  src_skip (int32 (n_vars,)): target v
    leaves out its first src_skip[v] sources in level-count order (the pairs
    a one-hot plan counts)."""
    npv_host = np.ascontiguousarray(npv_host, dtype=np.int32)
    n = len(npv_host)
    tg = None if targets is None else np.ascontiguousarray(targets, dtype=np.int32)
    sr = None if sources is None else np.ascontiguousarray(sources, dtype=np.int32)
    sk = None if src_skip is None else np.ascontiguousarray(src_skip, dtype=np.int32)
    info = _lib.PlanInfo()
    args = (ptr(npv_host), n, int(L), ptr(tg), 0 if tg is None else len(tg), ptr(sr),
            0 if sr is None else len(sr), int(cap_wave), int(cap_block), ptr(sk))
    call("tnx_te_plan", *args, None, None, None, 0, C.byref(info))
    n_items = info.n_wave_items + info.n_block_items
    tgt = np.empty(info.n_tgt, np.int32)
    src = np.empty(info.n_src, np.int32)
    work = np.empty((max(n_items, 1), 4), np.int32)
    call("tnx_te_plan", *args, ptr(tgt), ptr(src), ptr(work), n_items, C.byref(info))
    return Plan(info=info, tgt_order=h2d(tgt, device), src_order=h2d(src, device),
                work=h2d(work, device), work_host=work[:n_items],
                tgt_host=tgt, src_host=src)


# Plans depend only on the level counts, the pair set and the caps; the last
# few are kept across Prepared objects (a surrogate test prepares its data
# afresh on every call).
_PLAN_CACHE: "OrderedDict[tuple, Plan]" = OrderedDict()
_PLAN_CACHE_SIZE = 16


def _cached_plan(key, build):
    p = _PLAN_CACHE.get(key)
    if p is None:
        p = build()
        _PLAN_CACHE[key] = p
        while len(_PLAN_CACHE) > _PLAN_CACHE_SIZE:
            _PLAN_CACHE.popitem(last=False)
    else:
        _PLAN_CACHE.move_to_end(key)
    return p


def _digest(a) -> bytes | None:
    if a is None:
        return None
    return hashlib.blake2b(np.ascontiguousarray(a).tobytes(), digest_size=16).digest()


def get_plan(prep: Prepared, targets=None, sources=None, cap_wave: int = 0,
             cap_block: int = 0, src_skip=None) -> Plan:
    if cap_wave == 0 and prep.hot_words:
        cap_wave = wave_cap(prep)
    key = ("te", _digest(prep.npv_host), int(prep.L), str(prep.device),
           _digest(None if targets is None else np.asarray(targets, np.int32)),
           _digest(None if sources is None else np.asarray(sources, np.int32)),
           int(cap_wave), int(cap_block),
           _digest(None if src_skip is None else np.asarray(src_skip, np.int32)))
    # variables above 127 levels never enter a histogram plan (their pairs run
    # on the occupied-key kernel); the planner sees them clamped
    npv = np.minimum(prep.npv_host, MAX_LEVELS_INT8) if len(prep.wide) else prep.npv_host
    if len(prep.wide):
        isw = prep.npv_host > MAX_LEVELS_INT8
        for lst in (targets, sources):
            if lst is None or isw[np.asarray(lst, dtype=np.int64)].any():
                raise ValueError("a histogram plan cannot take variables of more than "
                                 f"{MAX_LEVELS_INT8} levels")
    return _cached_plan(key, lambda: make_plan(npv, prep.L, prep.device, targets,
                                               sources, cap_wave, cap_block, src_skip))


# ── TE run ────────────────────────────────────────────────────────────────

# The persistent single launch (te_persist.hip) takes series of up to this
# many samples; TNX_PERSIST=0 in the environment selects the two-launch path.
PERSIST_MAX_L = 12 * 1024
PERSIST_MAX_SLICES_BYTES = 156 * 1024  # four wave slices per workgroup


def _wave_slice(cap: int) -> int:
    return (cap + 64 + 3) & ~3


def persistent_ok(prep: Prepared, plan: Plan) -> bool:
    """Whether the plan runs as one persistent launch: short enough series
    and a block-role histogram (four wave slices) that holds the widest
    block pair's (x_t, y_t) marginal."""
    if os.environ.get("TNX_PERSIST", "1") == "0" or prep.L > PERSIST_MAX_L:
        return False
    i = plan.info
    if 4 * _wave_slice(_wave_area(prep, int(i.cap_wave))) * 4 > PERSIST_MAX_SLICES_BYTES:
        return False
    if i.n_block_items == 0:
        return True
    return block_cap(prep, plan) >= max(int(i.max_block_cols), 4)


def _wave_area(prep: Prepared, cap_wave: int) -> int:
    """Words of a wave slice before its 64 lane words (te_args.hpp wave_area)."""
    cw, mb = prep.hot
    return ((cap_wave + 3) & ~3) + 4 * cw * mb if cw else cap_wave


def block_cap(prep: Prepared, plan: Plan) -> int:
    """Histogram cells of the persistent launch's block role (its four wave
    slices less the widest marginal, two rows of lane words and the hot
    cells): a block pair of more cells runs in several passes."""
    i = plan.info
    mc = max(int(i.max_block_cols), 1)
    return (4 * _wave_slice(_wave_area(prep, int(i.cap_wave))) - mc - 128
            - 4 * HOT_CB * prep.hot[1]) & ~3


def plan_queue(plan: Plan, npv_host: np.ndarray, L: int, item_per_wave: bool):
    """Device work queues of the persistent launch for plan (tnx_te_queue),
    cached on the plan: (items, offsets, counters)."""
    key = ("queue", bool(item_per_wave))
    got = plan.cache.get(key)
    if got is None:
        i = plan.info
        nw, nb = int(i.n_wave_items), int(i.n_block_items)
        work = np.ascontiguousarray(plan.work_host.reshape(-1, 4)[: nw + nb], dtype=np.int32)
        npv = np.ascontiguousarray(npv_host, dtype=np.int32)
        n_we = C.c_int64(0)
        args = (ptr(work), nw, nb, ptr(npv), len(npv), ptr(plan.tgt_host), ptr(plan.src_host),
                int(L), int(i.cap_wave), int(bool(item_per_wave)))
        call("tnx_te_queue", *args, None, 0, None, C.byref(n_we))
        n_ent = int(n_we.value) + nb
        items = np.empty(max(n_ent, 1), np.int32)
        off = np.zeros(2 * _lib.TNX_QUEUES + 1, np.int32)
        call("tnx_te_queue", *args, ptr(items), n_ent, ptr(off), C.byref(n_we))
        dev = plan.work.device
        ctr = torch.zeros(_lib.TNX_QUEUE_CTR_INTS, dtype=torch.int32, device=dev)
        got = (h2d(items, dev), h2d(off, dev), ctr)
        plan.cache[key] = got
    return got


def te_args(prep: Prepared, plan: Plan, out: torch.Tensor | None, out_mode: int, ld_out: int = 0,
            pair_offset: int = 0, n_pairs_local: int = 0, stats=None, stage_vars: int = 0,
            stage_stride: int = 0, item_per_wave: bool | None = None,
            persistent: bool | None = None, counts=None, var_L=None,
            cap_sort: int = 0) -> _lib.TeArgs:
    """stats: (sum, sumsq, count_ge, observed) public-layout tensors for
    TNX_OUT_STATS; stage_vars / stage_stride for TNX_OUT_STAGE.  item_per_wave None: chosen from the
    plan (item_per_wave_default).  persistent None: one persistent launch
    whenever persistent_ok.  counts = (map int32 (n, n), out int32 (rows,
    ld)) turns on the kernels' joint-count dump (persistent launch).  var_L
    (device int32 (n_vars,)): per-target series lengths of stacked copies
    prepared with different lags (persistent launch)."""
    i = plan.info
    if item_per_wave is None:
        item_per_wave = item_per_wave_default(plan)
    if persistent is None:
        persistent = persistent_ok(prep, plan)
    if counts is not None and not persistent:
        raise ValueError("the joint-count dump needs the persistent launch")
    if var_L is not None and not persistent:
        raise ValueError("per-variable series lengths need the persistent launch")
    q = plan_queue(plan, prep.npv_host, prep.L, item_per_wave) if persistent else (None,) * 3
    cm, cd = (counts[0], counts[1]) if counts is not None else (None, None)
    st = stats if stats is not None else (None, None, None, None)
    return _lib.TeArgs(
        bins=ptr(prep.bins), states=ptr(prep.states), n_per_var=ptr(prep.npv), S1=ptr(prep.S1),
        S2a=ptr(prep.S2a), xlogx=ptr(prep.xlogx), tgt_order=ptr(plan.tgt_order),
        src_order=ptr(plan.src_order), work=ptr(plan.work), n_vars=prep.n_vars, T=prep.T,
        ld=prep.ld, lds=prep.lds, tau=prep.tau, out_mode=out_mode,
        n_wave_items=i.n_wave_items, n_block_items=i.n_block_items, cap_wave=i.cap_wave,
        cap_block=i.cap_block, max_block_cols=max(i.max_block_cols, 1), out=ptr(out),
        ld_out=ld_out, pair_offset=pair_offset, n_pairs_local=n_pairs_local,
        sum_te=ptr(st[0]), sumsq_te=ptr(st[1]), count_ge=ptr(st[2]), observed=ptr(st[3]),
        hot_zero=int(prep.hot_zero), stage_vars=int(stage_vars), stage_stride=int(stage_stride),
        item_per_wave=int(bool(item_per_wave)), queue_items=ptr(q[0]), queue_off=ptr(q[1]),
        queue_ctr=ptr(q[2]), counts_map=ptr(cm), counts_dbg=ptr(cd),
        counts_ld=int(cd.shape[1]) if cd is not None else 0,
        xlogx_fx=ptr(xlogx_fx_table(prep.L, prep.device)), fx_shift=fx_shift(prep.L),
        var_L=ptr(var_L), cap_sort=int(cap_sort), hot_cw=prep.hot[0], hot_maxb=prep.hot[1])


def te_out_args(prep: Prepared, out: torch.Tensor, out_mode: int, ld_out: int) -> _lib.TeArgs:
    """TE args with prep's per-variable data and an output only (no
    histogram plan): what the one-hot kernel reads."""
    return _lib.TeArgs(
        bins=ptr(prep.bins), states=ptr(prep.states), n_per_var=ptr(prep.npv), S1=ptr(prep.S1),
        S2a=ptr(prep.S2a), xlogx=ptr(prep.xlogx), n_vars=prep.n_vars, T=prep.T, ld=prep.ld,
        lds=prep.lds, tau=prep.tau, out_mode=out_mode, out=ptr(out), ld_out=ld_out,
        xlogx_fx=ptr(xlogx_fx_table(prep.L, prep.device)), fx_shift=fx_shift(prep.L))


WAVE_ITEM_PAIRS = 32   # pairs of a one-target item that a single wave takes whole


def item_per_wave_default(plan: Plan) -> bool:
    """A wave per work item when every wave item holds one target and at
    most WAVE_ITEM_PAIRS pairs (pair-list plans, where the four waves would
    otherwise share one target's few pairs) and there are enough items to
    fill the chip that way."""
    got = plan.cache.get("item_per_wave")
    if got is None:
        nw = int(plan.info.n_wave_items)
        w = np.asarray(plan.work_host).reshape(-1, 4)[:nw]
        got = bool(nw >= 4 * FILL_WAVE_ITEMS and np.all(w[:, 1] == 1)
                   and int((w[:, 3] - w[:, 2]).max(initial=0)) <= WAVE_ITEM_PAIRS)
        plan.cache["item_per_wave"] = got
    return got


def te_run(prep: Prepared, plan: Plan, out: torch.Tensor, out_mode: int, ld_out: int = 0,
           pair_offset: int = 0, n_pairs_local: int = 0, args: _lib.TeArgs | None = None):
    if args is None:
        args = te_args(prep, plan, out, out_mode, ld_out, pair_offset, n_pairs_local)
    call("tnx_te_run", C.byref(args), _stream(prep.device))


# (cap_wave, cap_block) candidates of the autotuner: the wave kernel's LDS
# histogram per wave (which pairs go to the block kernel, and how many
# workgroups fit a CU) and the block kernel's histogram per pass.
TUNE_CAPS = ((0, 0), (2400, 0), (4096, 0), (0, 8192))
_TUNE_CACHE: dict = {}


def autotune(prep: Prepared, targets=None, sources=None, sample_targets: int = 256,
             reps: int = 3, min_gain: float = 0.02) -> dict:
    """Launch parameters for the TE kernels on prep, chosen by timing: each
    (cap_wave, cap_block) of TUNE_CAPS, with and without the hot-state (key-0)
    diversion, runs the plan of an evenly spaced sample of at most
    sample_targets targets against all sources; a candidate replaces the
    defaults only when it is min_gain faster.  Cached per (levels, L,
    targets, sources, device).  Returns dict(cap_wave, cap_block, hot_zero,
    ms) and sets prep.hot_zero."""
    tg_all = np.arange(prep.n_vars, dtype=np.int32) if targets is None else \
        np.asarray(targets, dtype=np.int32)
    key = (_digest(prep.npv_host), prep.L, str(prep.device), _digest(tg_all),
           _digest(None if sources is None else np.asarray(sources, np.int32)))
    best = _TUNE_CACHE.get(key)
    if best is None:
        # sample targets spread over the level counts (sorted by levels)
        order = tg_all[np.argsort(prep.npv_host[tg_all], kind="stable")]
        step = max(1, len(order) // sample_targets)
        tg = np.sort(order[::step][:sample_targets])
        out = torch.empty((prep.n_vars, prep.n_vars), dtype=torch.float32, device=prep.device)
        hz0 = prep.hot_zero
        timings = []
        for hz in (hz0, not hz0):
            prep.hot_zero = hz
            for cw, cb in TUNE_CAPS:
                # cw 0: the default cap, less the hot cells' words when they are on
                plan = make_plan(prep.npv_host, prep.L, prep.device, tg, sources,
                                 cw or wave_cap(prep), cb)
                args = te_args(prep, plan, out, _lib.TNX_OUT_MATRIX, ld_out=prep.n_vars)
                te_run(prep, plan, out, _lib.TNX_OUT_MATRIX, args=args)
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
                for _ in range(reps):
                    te_run(prep, plan, out, _lib.TNX_OUT_MATRIX, args=args)
                e1.record()
                e1.synchronize()
                timings.append((e0.elapsed_time(e1) / reps, cw, cb, hz))
        prep.hot_zero = hz0
        default = timings[0]
        fastest = min(timings)
        pick = fastest if fastest[0] < (1.0 - min_gain) * default[0] else default
        best = dict(cap_wave=pick[1], cap_block=pick[2], hot_zero=pick[3], ms=pick[0],
                    default_ms=default[0], sample_targets=len(tg))
        _TUNE_CACHE[key] = best
    prep.hot_zero = best["hot_zero"]
    return best


def capture_te_step(prep: Prepared, plan: Plan, out: torch.Tensor, out_mode: int,
                    ld_out: int = 0, oh=None) -> torch.cuda.CUDAGraph:
    """The per-variable preparation and the plan's TE launches (with oh, a
    one-hot plan, its pack and MFMA launches too) captured once in a HIP
    graph (torch.cuda.CUDAGraph): replay() re-runs them on prep's current
    bins with no host launch work.  The launches are issued once eagerly
    first (kernel attributes set outside the capture); nothing in them is a
    memset, and the buffers are fixed, so every replay is equivalent to the
    eager sequence."""
    from . import onehot
    args = te_args(prep, plan, out, out_mode, ld_out=ld_out)

    def step():
        gene_prep_into(prep)
        if oh is not None:
            onehot.run(prep, oh, args, _stream(prep.device))
        te_run(prep, plan, out, out_mode, args=args)

    step()
    torch.cuda.synchronize(prep.device)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        step()
    g._tnx_keep = (args, plan, out, oh)  # the graph's kernels read these
    return g


def onehot_enabled() -> bool:
    """Whether TE matrices count the pairs of few levels on INT8 MFMA
    (onehot.py); TNX_ONEHOT=0 in the environment counts every pair in the
    histogram kernels."""
    return os.environ.get("TNX_ONEHOT", "1") != "0"


def get_te_plan(prep: Prepared, targets=None, sources=None, cap_wave: int = 0, cap_block: int = 0,
                thr: int | None = None) -> Plan:
    """Histogram plan for targets x sources that leaves the pairs of few
    levels to a one-hot plan (onehot.py), kept as plan.cache["onehot"]
    (absent when onehot_enabled() is off or no pair qualifies).  thr: the
    one-hot threshold (None: onehot.ONEHOT_THR)."""
    oh = get_onehot_plan(prep, targets, sources, thr) if onehot_enabled() else None
    plan = get_plan(prep, targets, sources, cap_wave, cap_block,
                    src_skip=None if oh is None else oh.src_skip)
    if oh is not None:
        plan.cache["onehot"] = oh
    return plan


def plan_all_pairs(plan: Plan, npv_host: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """(target, source) of every pair a plan counts: its histogram items and
    its one-hot plan's pairs (target != source)."""
    x, y = plan_pairs(plan)
    oh = plan.cache.get("onehot")
    if oh is not None:
        from . import onehot
        cov = onehot.covered_pairs(oh, npv_host)
        x, y = np.concatenate([x, cov[:, 0]]), np.concatenate([y, cov[:, 1]])
    return x, y


def get_onehot_plan(prep: Prepared, targets=None, sources=None, thr: int | None = None,
                    src_ie: bool = True):
    """Cached one-hot plan of prep's narrow variables (None when no pair
    qualifies or when the data hold variables of more than 127 levels)."""
    from . import onehot
    thr = onehot.default_threshold(prep.L) if thr is None else int(thr)
    if len(prep.wide) or thr <= 0:
        return None
    key = ("onehot", _digest(prep.npv_host), int(prep.L), str(prep.device),
           _digest(None if targets is None else np.asarray(targets, np.int32)),
           _digest(None if sources is None else np.asarray(sources, np.int32)), thr,
           bool(src_ie))
    return _cached_plan(key, lambda: onehot.plan(prep.npv_host, prep.L, prep.device, targets,
                                                 sources, thr, src_ie))


# Per-pair cost of the occupied-key kernel in microseconds of one CU by sort
# size (measured on MI355X: 39,800 pairs per launch), and of a block-role
# pair (block_pair_cost's units, te_persist.hip, at 1.04e-4 us per unit).
SORT_US = {1024: 4.6, 2048: 9.5, 4096: 20.0, 8192: 55.0, 16384: 134.0}
DENSE_US_PER_UNIT = 1.04e-4


_SORT_THRESHOLDS: dict = {}


def sort_threshold(L: int, cap_wave: int = 0) -> int:
    """Smallest b_t^2 b_s above which the occupied-key kernel is expected to
    beat the block role's dense passes by 1.5x for series of L samples (0:
    never, including every series the LDS sort cannot hold).  Cached."""
    cap_wave = int(cap_wave) or WAVE_CAP
    key = (int(L), cap_wave)
    if key not in _SORT_THRESHOLDS:
        _SORT_THRESHOLDS[key] = _sort_threshold(int(L), cap_wave)
    return _SORT_THRESHOLDS[key]


def _sort_threshold(L: int, cap_wave: int) -> int:
    if L > TRACE_MAX_L:
        return 0
    P = max(1024, 1 << (max(1, int(L)) - 1).bit_length())
    sort_us = SORT_US[P]
    cap_b = 4 * _wave_slice(cap_wave) - 2 * 64 - 128
    best = 0
    for bx in range(2, MAX_LEVELS_INT8 + 1):
        for by in range(2, MAX_LEVELS_INT8 + 1):
            cells = bx * bx * by
            if cells <= cap_wave:
                continue
            ncol = bx * by
            per = max(1, cap_b // ncol)
            passes = -(-bx // per)
            dense = (L * (1.8 + 0.35 * passes) + 2.0 * cells + 6000.0) * DENSE_US_PER_UNIT
            if dense > 1.5 * sort_us and (best == 0 or cells < best):
                best = cells
    return max(best - 1, cap_wave + 1) if best else 0


def _sort_split(prep: Prepared, tg: np.ndarray, sr: np.ndarray, cap: int) -> np.ndarray:
    """[target, source] pairs of tg x sr (target != source) with more than
    cap histogram cells."""
    npv = prep.npv_host.astype(np.int64)
    order = np.argsort(npv[sr], kind="stable")
    s_sorted = sr[order]
    b_sorted = npv[s_sorted]
    out = []
    for t in tg:
        lim = cap // (npv[t] * npv[t])              # sources with by > lim exceed cap
        k = np.searchsorted(b_sorted, lim, side="right")
        if k < len(s_sorted):
            s = s_sorted[k:]
            s = s[s != t]
            if len(s):
                out.append(np.stack([np.full(len(s), t), s], 1))
    return np.concatenate(out).astype(np.int64) if out else np.zeros((0, 2), np.int64)


def te_matrix_internal(prep: Prepared, targets=None, sources=None, cap_wave: int = 0,
                       cap_block: int = 0) -> torch.Tensor:
    """Internal-layout TE matrix M[target, source] = TE(source -> target) on the
    device (zeros where not computed)."""
    n = prep.n_vars
    out = torch.zeros((n, n), dtype=torch.float32, device=prep.device)
    if len(prep.wide):
        # pairs touching a variable of more than 127 levels: occupied keys;
        # the rest: the histogram kernels
        isw = np.zeros(n, dtype=bool)
        isw[prep.wide] = True
        tg = np.arange(n) if targets is None else np.asarray(targets, dtype=np.int64)
        sr = np.arange(n) if sources is None else np.asarray(sources, dtype=np.int64)
        tn, sn = tg[~isw[tg]], sr[~isw[sr]]
        if len(tn) and len(sn):
            plan = get_plan(prep, tn.astype(np.int32), sn.astype(np.int32), cap_wave, cap_block)
            te_run(prep, plan, out, _lib.TNX_OUT_MATRIX, ld_out=n)
        tw, sw = tg[isw[tg]], sr[isw[sr]]
        parts = [np.stack(np.meshgrid(tw, sr, indexing="ij"), -1).reshape(-1, 2),
                 np.stack(np.meshgrid(tn, sw, indexing="ij"), -1).reshape(-1, 2)]
        wp = np.concatenate(parts)
        te_sort_pairs(prep, wp[wp[:, 0] != wp[:, 1]], out, n)
        return out
    plan = get_te_plan(prep, targets, sources, cap_wave, cap_block)
    oh = plan.cache.get("onehot")
    if oh is not None:
        from . import onehot
        onehot.run(prep, oh, te_args(prep, plan, out, _lib.TNX_OUT_MATRIX, ld_out=n),
                   _stream(prep.device))
    cap = sort_threshold(prep.L, int(plan.info.cap_wave)) if plan.info.n_block_items else 0
    if cap:
        tg = np.arange(n) if targets is None else np.asarray(targets, dtype=np.int64)
        sr = np.arange(n) if sources is None else np.asarray(sources, dtype=np.int64)
        npv = prep.npv_host.astype(np.int64)
        if npv[tg].max(initial=0) ** 2 * npv[sr].max(initial=0) <= cap:
            cap = 0                         # no pair of the grid is that large
        else:
            sp = _sort_split(prep, tg, sr, cap)
            if len(sp) == 0:
                cap = 0
    if not cap:
        te_run(prep, plan, out, _lib.TNX_OUT_MATRIX, ld_out=n)
        return out
    # pairs whose dense histograms dwarf the series: occupied keys
    args = te_args(prep, plan, out, _lib.TNX_OUT_MATRIX, ld_out=n, cap_sort=cap)
    te_run(prep, plan, out, _lib.TNX_OUT_MATRIX, args=args)
    te_sort_pairs(prep, sp, out, n)
    return out


def te_pair_list(prep: Prepared, pairs: np.ndarray, cap_wave: int = 0) -> torch.Tensor:
    """TE of an explicit [target, source] pair list as a (P,) float32 tensor
    (0 where target == source), costing the listed pairs only: a pair-list
    plan over the variables the pairs touch, run into a matrix over those
    variables (a private re-prepared copy of their rows when they are fewer
    than all)."""
    p = np.ascontiguousarray(pairs, dtype=np.int64).reshape(-1, 2)
    P = len(p)
    out = torch.zeros(P, dtype=torch.float32, device=prep.device)
    if P and (p.min() < 0 or p.max() >= prep.n_vars):
        raise ValueError(f"pair indices outside [0, {prep.n_vars})")
    valid = p[:, 0] != p[:, 1]
    if len(prep.wide):
        isw = np.zeros(prep.n_vars, dtype=bool)
        isw[prep.wide] = True
        wmask = valid & (isw[p[:, 0]] | isw[p[:, 1]])
        if wmask.any():
            tmp = torch.empty(int(wmask.sum()), dtype=torch.float32, device=prep.device)
            te_sort_pairs(prep, p[wmask], tmp, 0)
            out[h2d(np.nonzero(wmask)[0], prep.device)] = tmp
        valid &= ~wmask
    if not valid.any():
        return out
    pv = p[valid]
    involved = np.unique(pv)
    m = len(involved)
    if m == prep.n_vars:
        sub, loc = prep, pv
    else:
        sub = var_subset(prep, involved, cache=False)
        gene_prep_into(sub, features=False)
        loc = np.searchsorted(involved, pv)
    plan = te_pairlist_plan(sub.npv_host, np.unique(loc, axis=0), prep.device,
                            cap_wave or wave_cap(sub))
    M = torch.zeros((m, m), dtype=torch.float32, device=prep.device)
    te_run(sub, plan, M, _lib.TNX_OUT_MATRIX, ld_out=m)
    lt = h2d(loc, prep.device)
    out[h2d(np.nonzero(valid)[0], prep.device)] = M[lt[:, 0], lt[:, 1]]
    return out


TRACE_MAX_L = 16384     # trace.hip sorts one entropy's keys in LDS


def trace_te(bins: torch.Tensor, T: int, tau: int, mode: int, var_ids, conn_lists,
             return_sums: bool = False):
    """TRACE OutTE (mode 0, conn = each variable's targets) or InTE (mode 1,
    conn = its sources) of the variables var_ids on device int8 bins (n, ld)
    through tnx_trace_entropy; returns float32 (len(var_ids),) on the device
    (and the int64 (jobs, 4) fixed-point entropy sums with return_sums)."""
    dev = bins.device
    var_ids = np.ascontiguousarray(var_ids, dtype=np.int32)
    J = len(var_ids)
    te = torch.zeros(J, dtype=torch.float32, device=dev)
    sums = torch.zeros((J, 4), dtype=torch.int64, device=dev)
    if J == 0:
        return (te, sums) if return_sums else te
    L = int(T) - int(tau)
    if L > TRACE_MAX_L:
        raise NotImplementedError(f"TRACE on the GPU takes T - tau <= {TRACE_MAX_L} (got {L})")
    n = bins.shape[0]
    lens = np.array([len(c) for c in conn_lists], dtype=np.int64)
    off = np.zeros(J + 1, dtype=np.int64)
    np.cumsum(lens, out=off[1:])
    flat = (np.concatenate([np.asarray(c, dtype=np.int64) for c in conn_lists])
            if off[-1] else np.zeros(1, np.int64))
    if (var_ids.min() < 0 or var_ids.max() >= n or flat.min() < 0 or flat.max() >= n):
        raise ValueError(f"TRACE variable indices outside [0, {n})")
    tbl = xlogx_fx_table(L, dev)
    v_d = h2d(var_ids, dev)
    c_d = h2d(flat.astype(np.int32), dev)
    o_d = h2d(off, dev)
    call("tnx_trace_entropy", ptr(bins), n, int(T), bins.shape[1], int(tau), int(mode), ptr(v_d),
         ptr(c_d), ptr(o_d), J, ptr(tbl), fx_shift(L), ptr(sums), ptr(te), _stream(dev))
    return (te, sums) if return_sums else te


def te_pairfree(prep: Prepared, pair_offset: int = 0, n_pairs_local: int | None = None,
                cap_wave: int = 0, cap_block: int = 0) -> torch.Tensor:
    """TE in pair-free linear order for ids [pair_offset, pair_offset + n_local)."""
    require_narrow(prep, "the pair-free launch")
    n = prep.n_vars
    total = n * (n - 1)
    if n_pairs_local is None:
        n_pairs_local = total - pair_offset
    if pair_offset < 0 or n_pairs_local < 0 or pair_offset + n_pairs_local > total:
        raise ValueError("pair range outside [0, n*(n-1))")
    out = torch.empty(n_pairs_local, dtype=torch.float32, device=prep.device)
    if n_pairs_local == 0:
        return out
    nm1 = n - 1
    t0 = pair_offset // nm1
    t1 = (pair_offset + n_pairs_local - 1) // nm1 + 1
    plan = get_plan(prep, np.arange(t0, t1, dtype=np.int32), None, cap_wave, cap_block)
    te_run(prep, plan, out, _lib.TNX_OUT_PAIRFREE, pair_offset=pair_offset,
           n_pairs_local=n_pairs_local)
    return out


def counts_plan(prep: Prepared, plan: Plan, pairs: np.ndarray, max_cells: int, oh=None):
    """Joint counts of the given [target, source] pairs as the production
    launches count them (the persistent kernel's count dump; with oh, a
    one-hot plan, its MFMA launch's too): int32 (P, max_cells) in layout
    (a * b_t + b) * b_s + c, and the launches' TE matrix (internal
    layout)."""
    require_narrow(prep, "the joint-count dump")
    pairs = np.ascontiguousarray(pairs, dtype=np.int64).reshape(-1, 2)
    n, dev = prep.n_vars, prep.device
    cells = prep.npv_host[pairs[:, 0]].astype(np.int64) ** 2 * prep.npv_host[pairs[:, 1]]
    if len(cells) and cells.max() > max_cells:
        raise ValueError(f"a pair has {int(cells.max())} cells > max_cells {max_cells}")
    key, inv = np.unique(pairs[:, 0] * n + pairs[:, 1], return_inverse=True)
    cmap = torch.full((n * n,), -1, dtype=torch.int32, device=dev)
    cmap[h2d(key, dev)] = torch.arange(len(key), dtype=torch.int32, device=dev)
    counts = torch.zeros((max(len(key), 1), max_cells), dtype=torch.int32, device=dev)
    out = torch.zeros((n, n), dtype=torch.float32, device=dev)
    args = te_args(prep, plan, out, _lib.TNX_OUT_MATRIX, ld_out=n, persistent=True,
                   counts=(cmap, counts))
    if oh is not None:
        from . import onehot
        onehot.run(prep, oh, args, _stream(dev))
    te_run(prep, plan, out, _lib.TNX_OUT_MATRIX, args=args)
    return counts[h2d(inv.reshape(-1), dev)], out


def counts_debug(prep: Prepared, pairs: np.ndarray, max_cells: int = 8192):
    """Raw joint counts of explicit [target, source] pairs through the
    production launch (a pair-list plan in the persistent kernel, wave and
    block roles); returns (counts (P, max_cells), TE (P,)).  Series longer
    than the persistent launch takes use tnx_te_counts_debug (wave path)."""
    pairs_np = np.ascontiguousarray(pairs, dtype=np.int64).reshape(-1, 2)
    if prep.L <= PERSIST_MAX_L and os.environ.get("TNX_PERSIST", "1") != "0":
        plan = te_pairlist_plan(prep.npv_host, pairs_np, prep.device, wave_cap(prep))
        if persistent_ok(prep, plan):
            counts, m = counts_plan(prep, plan, pairs_np, max_cells)
            te = m[h2d(pairs_np[:, 0], prep.device), h2d(pairs_np[:, 1], prep.device)]
            return counts, te
    pairs = np.ascontiguousarray(pairs, dtype=np.int32)
    p_d = torch.from_numpy(pairs).to(prep.device)
    P = len(pairs)
    counts = torch.zeros((P, max_cells), dtype=torch.int32, device=prep.device)
    te = torch.empty(P, dtype=torch.float32, device=prep.device)
    call("tnx_te_counts_debug", ptr(prep.bins), ptr(prep.states), ptr(prep.npv), ptr(prep.S1),
         ptr(prep.S2a), ptr(prep.xlogx), prep.n_vars, prep.T, prep.ld, prep.lds, prep.tau,
         ptr(p_d), P, int(max_cells), ptr(counts), ptr(te), _stream(prep.device))
    return counts, te


def gene_prep_into(prep: Prepared, features: bool = True):
    """Re-run the per-variable preparation into prep's existing buffers
    (features=False: states and entropies only)."""
    call("tnx_gene_prep", ptr(prep.bins), prep.n_vars, prep.T, prep.ld, int(prep.tau),
         ptr(prep.npv), ptr(prep.states), prep.lds, ptr(prep.S1), ptr(prep.S2a),
         ptr(prep.feat) if features else None, int(prep.npv_host.max(initial=1)),
         _stream(prep.device))


def lds_atomic_rate(device, mode: int = 0, n_blocks: int | None = None, iters: int = 4096,
                    reps: int = 5) -> float:
    """Measured chip-wide LDS atomic increments per second (median of reps)."""
    props = device_props(device)
    if n_blocks is None:
        n_blocks = int(props.n_cu) * 8
    out = torch.empty(n_blocks, dtype=torch.int32, device=device)
    s = _stream(device)
    call("tnx_bench_lds_atomics", int(mode), int(n_blocks), int(iters), ptr(out), s)
    rates = []
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        call("tnx_bench_lds_atomics", int(mode), int(n_blocks), int(iters), ptr(out), s)
        e1.record()
        e1.synchronize()
        rates.append(n_blocks * 256.0 * iters * 8 / (e0.elapsed_time(e1) / 1e3))
    return float(np.median(rates))


MFMA_OPS = {0: 32 * 32 * 64 * 2, 1: 32 * 32 * 64 * 2, 2: 32 * 32 * 32 * 2}


def mfma_rate(device, mode: int = 0, n_blocks: int | None = None, iters: int = 2048,
              reps: int = 3) -> float:
    """Measured chip-wide matrix-core operations per second from registers
    (tnx_bench_mfma: mode 0 FP4, 1 FP8, 2 INT8; median of reps)."""
    props = device_props(device)
    if n_blocks is None:
        n_blocks = int(props.n_cu) * 4
    out = torch.empty(n_blocks * 256, dtype=torch.float32, device=device)
    s = _stream(device)
    call("tnx_bench_mfma", int(mode), int(n_blocks), int(iters), ptr(out), s)
    rates = []
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        call("tnx_bench_mfma", int(mode), int(n_blocks), int(iters), ptr(out), s)
        e1.record()
        e1.synchronize()
        n_mfma = n_blocks * 4.0 * iters * 4
        rates.append(n_mfma * MFMA_OPS[int(mode)] / (e0.elapsed_time(e1) / 1e3))
    return float(np.median(rates))


# ── surrogates ────────────────────────────────────────────────────────────

SURR_MAX_CELLS = 32768


def surrogate_cap(npv_host: np.ndarray, targets=None, sources=None) -> int:
    """Histogram cells needed by the surrogate kernel: max b_t^2 * b_s over
    the pairs, rounded up to an even count >= 256."""
    bt = npv_host if targets is None else npv_host[np.asarray(targets)]
    bs = npv_host if sources is None else npv_host[np.asarray(sources)]
    need = int(bt.max(initial=1)) ** 2 * int(bs.max(initial=1))
    return max(256, need + (need & 1))


def block_perm_device(n_vars: int, T: int, shuffle_method: str, block_length: int, k_start: int,
                      k_end: int, seed: int, rng_device, out_device, rows=None):
    """Block orders for surrogates [k_start, k_end): perm[k] =
    argsort(rand(n_vars, n_blocks)) after manual_seed(seed + k) on a torch
    Generator on rng_device (tenex/inference/surrogate_test.py:165-207).
    rows: only these variables' orders (the same draws, the other rows
    discarded).  The sort is stable, so tied draws order by block index
    whatever the number of rows sorted.  Returns (perm int32
    (k, n_vars or len(rows), n_blocks) on out_device, eff_bl, n_blocks)."""
    rnd_all = shuffle_method == "random" or block_length <= 1 or block_length >= T
    eff, nb = (1, T) if rnd_all else (int(block_length), T // int(block_length))
    if rows is None or (isinstance(rows, torch.Tensor) and rows.device == torch.device(rng_device)):
        rows_t = rows
    else:
        rows_t = h2d(np.asarray(rows.cpu() if isinstance(rows, torch.Tensor) else rows,
                                dtype=np.int64), rng_device)
    m = n_vars if rows is None else int(rows_t.numel())
    nk = k_end - k_start
    gen = torch.Generator(device=rng_device)
    if nk * n_vars * nb <= PERM_DRAW_FLOATS:
        # the batch's draws side by side, one sort; the orders depend only on
        # (seed, k, n_vars, n_blocks), so the last few batches are kept
        key = (n_vars, nb, seed, k_start, k_end, str(torch.device(rng_device)),
               str(torch.device(out_device)))
        full = _PERM_CACHE.get(key)
        if full is None:
            r = torch.empty((nk, n_vars, nb), dtype=torch.float32, device=rng_device)
            for ki, k in enumerate(range(k_start, k_end)):
                gen.manual_seed(seed + k)
                torch.rand(n_vars, nb, device=rng_device, generator=gen, out=r[ki])
            full = torch.argsort(r, dim=2, stable=True).to(device=out_device, dtype=torch.int32)
            _PERM_CACHE[key] = full
            while sum(v.numel() for v in _PERM_CACHE.values()) > PERM_CACHE_INTS:
                _PERM_CACHE.popitem(last=False)
        else:
            _PERM_CACHE.move_to_end(key)
        if rows_t is None:
            return full, eff, nb
        return full.index_select(1, rows_t.to(full.device)), eff, nb
    out = torch.empty((nk, m, nb), dtype=torch.int32, device=out_device)
    for ki, k in enumerate(range(k_start, k_end)):
        gen.manual_seed(seed + k)
        r = torch.rand(n_vars, nb, device=rng_device, generator=gen)
        if rows_t is not None:
            r = r.index_select(0, rows_t)
        out[ki].copy_(torch.argsort(r, dim=1, stable=True).to(torch.int32))
    return out, eff, nb


def surrogate_run(prep: Prepared, plan: Plan, perm: torch.Tensor, block_length: int,
                  observed: torch.Tensor, sum_te: torch.Tensor, sumsq_te: torch.Tensor,
                  count_ge: torch.Tensor, te_k_dbg: torch.Tensor | None = None):
    """Accumulate surrogate statistics for every pair of the plan into the
    public-layout (n, n) sum / sumsq (float64) and count_ge (int32)."""
    n_surr, n, nb = perm.shape
    if n != prep.n_vars:
        raise ValueError("perm does not match n_vars")
    if plan.info.n_block_items:
        raise ValueError("surrogate plans must put every pair in the wave kernel")
    args = _lib.SurrArgs(
        bins=ptr(prep.bins), n_per_var=ptr(prep.npv), xlogx=ptr(prep.xlogx), perm=ptr(perm),
        tgt_order=ptr(plan.tgt_order), src_order=ptr(plan.src_order), work=ptr(plan.work),
        observed=ptr(observed), n_vars=n, T=prep.T, ld=prep.ld, tau=prep.tau, n_surr=n_surr,
        n_blocks=nb, block_length=int(block_length), n_wave_items=plan.info.n_wave_items,
        n_block_items=0, cap_wave=plan.info.cap_wave, cap_block=plan.info.cap_block,
        max_block_cols=max(plan.info.max_block_cols, 1), sum_te=ptr(sum_te),
        sumsq_te=ptr(sumsq_te), count_ge=ptr(count_ge), te_k_dbg=ptr(te_k_dbg),
        max_levels=int(prep.npv_host[plan.tgt_host].max(initial=0)),
        max_tgt_cnt=int(plan.work_host[:, 1].max(initial=0)))
    call("tnx_surrogate_run", C.byref(args), _stream(prep.device))


def block_shuffle(bins: torch.Tensor, T: int, perm_k: torch.Tensor, block_length: int,
                  out: torch.Tensor, copies: int = 1):
    """out = block-shuffled copies of bins (n_vars rows): row k * n_vars + v
    of out is row v shuffled with perm_k row k * n_vars + v (perm_k int32
    (copies * n_vars, n_blocks))."""
    n, ld = bins.shape
    perm_k = perm_k.contiguous()
    call("tnx_block_shuffle_copies", ptr(bins), n, int(copies), T, ld, ptr(perm_k),
         perm_k.shape[1], int(block_length), ptr(out), _stream(bins.device))


STAGE_FLOATS = 1 << 29    # staging for batched surrogate copies: 2 GiB of float32
COPY_BYTES = 6 << 30      # stacked surrogate copies (bins and states) held at once
PERM_DRAW_FLOATS = 32 << 20   # uniform draws of one batch of block orders held at once
PERM_CACHE_INTS = 128 << 20   # block orders kept across calls (512 MiB of int32)
_PERM_CACHE: "OrderedDict[tuple, torch.Tensor]" = OrderedDict()


def surrogate_stats_materialized(prep: Prepared, plan: Plan, perm: torch.Tensor,
                                 block_length: int, observed: torch.Tensor | None,
                                 sum_te: torch.Tensor, sumsq_te: torch.Tensor,
                                 count_ge: torch.Tensor | None, work: dict | None = None):
    """Surrogate statistics through the TE kernels: for each surrogate k of
    perm (k, n_vars, n_blocks), a block-shuffled copy of the bins, its
    per-variable states and entropies (tnx_gene_prep), then the plan's TE,
    added in surrogate order into the public-layout sum / sumsq (FP64) and
    count_ge (int32; not counted when count_ge or observed is None).  Up to
    STAGE_FLOATS / n_vars^2 surrogates (and COPY_BYTES of copies) go
    through one launch of each kernel: their copies are stacked as
    variables k * n_vars + j, the TE kernels write each copy's TE to a
    staging slab (TNX_OUT_STAGE) and tnx_stats_accumulate folds the slab
    into the statistics copy by copy; otherwise one TNX_OUT_STATS run per
    surrogate.  Both give the same additions in the same order.  `work`
    caches the buffers across calls."""
    n = prep.n_vars
    if work is None:
        work = {}
    if count_ge is None or observed is None:
        count_ge = observed = None
    n_k = perm.shape[0]
    kb = copies_per_launch(prep, n_k, n * n)
    if kb >= 2:
        slots = _plan_slots(plan, n, prep.device, prep.npv_host)
        for k0 in range(0, n_k, kb):
            kk = min(kb, n_k - k0)
            _staged_batch(prep, plan, perm[k0:k0 + kk], block_length, work, slots, observed,
                          sum_te, sumsq_te, count_ge)
        return
    sp = work.get("prep")
    if sp is None:
        sb = torch.empty_like(prep.bins)
        sp = Prepared(bins=sb, npv=prep.npv, npv_host=prep.npv_host, T=prep.T, tau=prep.tau,
                      states=torch.empty_like(prep.states), S1=torch.empty_like(prep.S1),
                      S2a=torch.empty_like(prep.S2a), feat=torch.empty_like(prep.feat),
                      xlogx=prep.xlogx, hot_zero=prep.hot_zero)  # block shuffles keep levels
        work["prep"] = sp
    if count_ge is None:   # the kernels' statistics epilogue always counts: a discarded count
        count_ge = work.get("discard_count")
        if count_ge is None:
            count_ge = torch.zeros_like(sum_te, dtype=torch.int32)
            work["discard_count"] = count_ge
        observed = torch.zeros_like(sum_te, dtype=torch.float32)
    args = te_args(sp, plan, None, _lib.TNX_OUT_STATS, ld_out=n,
                   stats=(sum_te, sumsq_te, count_ge, observed))
    oh = plan.cache.get("onehot")
    for kk in range(n_k):
        block_shuffle(prep.bins, prep.T, perm[kk], block_length, sp.bins)
        gene_prep_into(sp, features=False)
        if oh is not None:
            from . import onehot
            onehot.run(sp, oh, args, _stream(prep.device))
        te_run(sp, plan, None, _lib.TNX_OUT_STATS, args=args)


def copies_per_launch(prep: Prepared, n_k: int, stage_per_copy: int) -> int:
    """Surrogate copies of prep's variables that go through one launch: up
    to n_k, within STAGE_FLOATS of staging and COPY_BYTES of stacked bins
    and states."""
    per_copy = prep.n_vars * (prep.ld + 2 * prep.lds + 64)
    return int(min(n_k, STAGE_FLOATS // max(1, stage_per_copy), COPY_BYTES // max(1, per_copy)))


def _staged_batch(prep: Prepared, plan: Plan, perm: torch.Tensor, block_length: int, work: dict,
                  slots: torch.Tensor, observed, sum_te, sumsq_te, count_ge):
    """One launch of each kernel for the K = perm.shape[0] surrogate copies
    of prep's variables: shuffle, prepare, TE to the (K, n, n) staging slab,
    fold the slab into the public-layout statistics at `slots` in copy
    order (count_ge None: no counting)."""
    kk, n = perm.shape[0], prep.n_vars
    bw = work.get(("batch", kk))
    if bw is None:
        bw = _stacked_copies(prep, plan, kk)
        work[("batch", kk)] = bw
    sp, args, stage, ohK = bw
    block_shuffle(prep.bins, prep.T, perm.reshape(kk * n, -1), block_length, sp.bins, kk)
    gene_prep_into(sp, features=False)
    if ohK is not None:
        from . import onehot
        onehot.run(sp, ohK, args, _stream(prep.device))
    te_run(sp, None, None, _lib.TNX_OUT_STAGE, args=args)
    call("tnx_stats_accumulate", ptr(stage), kk, n * n, ptr(slots), slots.numel(),
         ptr(observed), ptr(sum_te), ptr(sumsq_te), ptr(count_ge), _stream(prep.device))


def _stacked_copies(prep: Prepared, plan: Plan, K: int):
    """Buffers for K stacked surrogate copies of prep's variables: the
    copies' Prepared, TE args for the K-fold plan writing to a (K, n, n)
    staging slab (only the plan's slots are written and read), and the
    slab."""
    n = prep.n_vars
    sp = Prepared(bins=torch.empty((K * n, prep.ld), dtype=prep.bins.dtype, device=prep.device),
                  npv=prep.npv.repeat(K).contiguous(),
                  npv_host=np.tile(prep.npv_host, K), T=prep.T, tau=prep.tau,
                  states=torch.empty((K * n, prep.lds), dtype=prep.states.dtype,
                                     device=prep.device),
                  S1=torch.empty(K * n, dtype=torch.float64, device=prep.device),
                  S2a=torch.empty(K * n, dtype=torch.float64, device=prep.device),
                  feat=torch.empty((K * n, prep.feat.shape[1]), dtype=torch.float32,
                                   device=prep.device),
                  xlogx=prep.xlogx, hot_zero=prep.hot_zero)
    planK = plan.cache.get(("copies", K))
    if planK is None:
        planK = replicate_plan(plan, K, n)
        plan.cache[("copies", K)] = planK
    stage = torch.empty(K * n * n, dtype=torch.float32, device=prep.device)
    args = te_args(sp, planK, stage, _lib.TNX_OUT_STAGE, ld_out=n, stage_vars=n,
                   stage_stride=n * n)
    ohK = None
    oh = plan.cache.get("onehot")
    if oh is not None:   # the one-hot pairs of each copy, all copies in one launch
        ohK = plan.cache.get(("onehot_copies", K))
        if ohK is None:
            from . import onehot
            ohK = onehot.replicate(oh, K, n, prep.device)
            plan.cache[("onehot_copies", K)] = ohK
    return sp, args, stage, ohK


# ── several lags in one launch ───────────────────────────────────────────

def _lag_stack(prep: Prepared, plan: Plan, taus, K: int):
    """Buffers for K copies of prep's variables at each lag of taus, stacked
    lag-major (variable (l K + k) n + j is copy k of variable j at lag
    taus[l]; prep's tau is the smallest lag): the stacked Prepared, the
    per-variable series lengths, TE args for the K len(taus)-fold plan
    writing a (K len(taus), n, n) staging slab, and the slab.  None when the
    plan cannot run as one persistent launch."""
    n, NL = prep.n_vars, len(taus)
    KT = K * NL
    sp = Prepared(bins=torch.empty((KT * n, prep.ld), dtype=prep.bins.dtype, device=prep.device),
                  npv=prep.npv.repeat(KT).contiguous(), npv_host=np.tile(prep.npv_host, KT),
                  T=prep.T, tau=prep.tau,
                  states=torch.empty((KT * n, prep.lds), dtype=prep.states.dtype,
                                     device=prep.device),
                  S1=torch.empty(KT * n, dtype=torch.float64, device=prep.device),
                  S2a=torch.empty(KT * n, dtype=torch.float64, device=prep.device),
                  feat=torch.empty((1, prep.feat.shape[1]), dtype=torch.float32,
                                   device=prep.device),
                  xlogx=prep.xlogx, hot_zero=prep.hot_zero)
    planK = plan.cache.get(("copies", KT))
    if planK is None:
        planK = replicate_plan(plan, KT, n)
        plan.cache[("copies", KT)] = planK
    if not persistent_ok(sp, planK):
        return None
    var_L = h2d(np.repeat(np.array([prep.T - t for t in taus], np.int32), K * n), prep.device)
    stage = torch.empty(KT * n * n, dtype=torch.float32, device=prep.device)
    args = te_args(sp, planK, stage, _lib.TNX_OUT_STAGE, ld_out=n, stage_vars=n,
                   stage_stride=n * n, persistent=True, var_L=var_L)
    return sp, args, stage, var_L


def _lag_prep(sp: Prepared, taus, K: int, n: int):
    """tnx_gene_prep of each lag's block of stacked copies with that lag."""
    bmax = int(sp.npv_host.max(initial=1))
    for l, tau in enumerate(taus):
        r0, r1 = l * K * n, (l + 1) * K * n
        call("tnx_gene_prep", ptr(sp.bins[r0:r1]), r1 - r0, sp.T, sp.ld, int(tau),
             ptr(sp.npv[r0:r1]), ptr(sp.states[r0:r1]), sp.lds, ptr(sp.S1[r0:r1]),
             ptr(sp.S2a[r0:r1]), None, bmax, _stream(sp.device))


def multilag_observed(prep: Prepared, plan: Plan, taus) -> list | None:
    """Internal-layout TE matrices of plan's pairs at every lag of taus in
    one launch (prep prepared with min(taus)); None when the stacked plan
    cannot run persistently."""
    n, NL = prep.n_vars, len(taus)
    got = _lag_stack(prep, plan, taus, 1)
    if got is None:
        return None
    sp, args, stage, _ = got
    for l in range(NL):
        sp.bins[l * n:(l + 1) * n].copy_(prep.bins)
    _lag_prep(sp, taus, 1, n)
    stage.zero_()
    te_run(sp, None, None, _lib.TNX_OUT_STAGE, args=args)
    slab = stage.view(NL, n, n)       # [lag, source, target]
    return [slab[l].t().contiguous() for l in range(NL)]


def multilag_surrogate_stats(prep: Prepared, plan: Plan, taus, block_lengths, observed,
                             n_surr: int, shuffle_method: str, seed: int, rng_device):
    """Surrogate statistics of plan's pairs at every lag of taus with each
    batch of surrogates in one TE launch for all lags: lag l's copies are
    block-shuffled with block_lengths[l] (the same orders as a run at that
    lag alone), prepared with taus[l] and counted with T - taus[l] samples.
    observed: per lag, public-layout float32 (n, n) or None (no counting).
    Returns per lag (sum, sumsq, count_ge) public layout, or None when the
    stacked plan cannot run persistently."""
    n, NL = prep.n_vars, len(taus)
    dev = prep.device
    # copies of every lag share the launch's staging and copy budget; the
    # surrogates are split into even batches (no small tail launch)
    kmax = max(1, copies_per_launch(prep, n_surr * NL, n * n) // NL)
    n_batches = -(-n_surr // kmax)
    kb = -(-n_surr // n_batches)
    slots = _plan_slots(plan, n, dev, prep.npv_host)
    out = []
    for l in range(NL):
        s = torch.zeros((n, n), dtype=torch.float64, device=dev)
        out.append((s, torch.zeros_like(s), torch.zeros((n, n), dtype=torch.int32, device=dev)))
    bufs = {}
    for k0 in range(0, n_surr, kb):
        k1 = min(k0 + kb, n_surr)
        kk = k1 - k0
        b = bufs.get(kk)
        if b is None:
            b = _lag_stack(prep, plan, taus, kk)
            if b is None:
                return None
            bufs[kk] = b
        sp, args, stage, _ = b
        for l in range(NL):
            perm, eff = copy_perms(n, prep.T, shuffle_method, int(block_lengths[l]), k0, k1, 0,
                                   n_surr, seed, rng_device, dev)
            r0 = l * kk * n
            block_shuffle(prep.bins, prep.T, perm.reshape(kk * n, -1), eff,
                          sp.bins[r0:r0 + kk * n], kk)
        _lag_prep(sp, taus, kk, n)
        te_run(sp, None, None, _lib.TNX_OUT_STAGE, args=args)
        for l in range(NL):
            s, s2, cg = out[l]
            obs = observed[l] if observed is not None else None
            call("tnx_stats_accumulate", ptr(stage[l * kk * n * n:]), kk, n * n, ptr(slots),
                 slots.numel(), ptr(obs), ptr(s), ptr(s2), ptr(cg if obs is not None else None),
                 _stream(dev))
    return out


def copy_perms(n_vars: int, T: int, shuffle_method: str, block_length: int, k0: int, k1: int,
               k_start: int, k_end: int, seed: int, rng_device, device, rows=None,
               perm_batch: int = 32):
    """Block orders of surrogates [k0, k1) inside a run over [k_start,
    k_end): taken from the perm_batch-aligned draws block_perm_device makes
    (and caches) for the run, so a launch may span several draws.  Returns
    (perm int32 (k1 - k0, rows, n_blocks), eff_bl)."""
    parts = []
    q0 = k_start + (k0 - k_start) // perm_batch * perm_batch
    eff = 1
    while q0 < k1:
        q1 = min(q0 + perm_batch, k_end)
        p_, eff, _ = block_perm_device(n_vars, T, shuffle_method, block_length, q0, q1, seed,
                                       rng_device, device, rows=rows)
        parts.append(p_[max(k0, q0) - q0:min(k1, q1) - q0])
        q0 = q1
    return (parts[0] if len(parts) == 1 else torch.cat(parts)), eff


def var_subset(prep: Prepared, vars_host: np.ndarray, cache: bool = True) -> Prepared:
    """The rows `vars_host` (sorted variable indices) of prep's bins as a
    source for the batched surrogate statistics: shuffling and re-preparing
    only the variables a plan touches.  Cached on prep unless cache=False
    (then a private copy whose states are left to the caller)."""
    vars_host = np.ascontiguousarray(vars_host, dtype=np.int64)
    key = ("subset", hash(vars_host.tobytes()), len(vars_host))
    sub = prep.plans.get(key) if cache else None
    if sub is None:
        vd = h2d(vars_host, prep.device)
        m = len(vars_host)
        sub = Prepared(bins=prep.bins.index_select(0, vd).contiguous(),
                       npv=prep.npv.index_select(0, vd).contiguous(),
                       npv_host=np.ascontiguousarray(prep.npv_host[vars_host]), T=prep.T,
                       tau=prep.tau, states=torch.empty((m, prep.lds), dtype=prep.states.dtype,
                                                        device=prep.device),
                       S1=torch.empty(m, dtype=torch.float64, device=prep.device),
                       S2a=torch.empty(m, dtype=torch.float64, device=prep.device),
                       feat=torch.empty((m, prep.feat.shape[1]), dtype=torch.float32,
                                        device=prep.device),
                       xlogx=prep.xlogx, hot_zero=prep.hot_zero)
        sub.plans["rows"] = vd    # the subset's variable indices on the device
        if cache:
            prep.plans[key] = sub
    return sub


def _plan_slots(plan: Plan, m: int, device, npv_host: np.ndarray) -> torch.Tensor:
    """Sorted public slots source * m + target of the plan's pairs (its
    one-hot plan's too) over m variables (device int64).  Cached on the
    plan."""
    got = plan.cache.get("slots")
    if got is None:
        x, y = plan_all_pairs(plan, npv_host)
        got = h2d(np.unique(y * m + x), device)
        plan.cache["slots"] = got
    return got


def _stats_run(sub: Prepared, plan: Plan, rows, n_vars: int, observed, s, s2, cg, n_surr: int,
               shuffle_method: str, block_length: int, seed: int, rng_device, k_start: int,
               perm_batch: int):
    """Surrogates [k_start, k_start + n_surr) of sub's variables (rows of an
    n_vars problem) into sub's public-layout statistics, in launches of
    copies_per_launch copies."""
    m = sub.n_vars
    k_end = k_start + n_surr
    kb = max(1, copies_per_launch(sub, n_surr, m * m))
    slots = _plan_slots(plan, m, sub.device, sub.npv_host)
    work = {}
    for k0 in range(k_start, k_end, kb):
        k1 = min(k0 + kb, k_end)
        perm, eff = copy_perms(n_vars, sub.T, shuffle_method, block_length, k0, k1, k_start,
                               k_end, seed, rng_device, sub.device, rows=rows,
                               perm_batch=perm_batch)
        _staged_batch(sub, plan, perm, eff, work, slots, observed, s, s2, cg)


def surrogate_stats_subset(prep: Prepared, vars_host: np.ndarray, plan: Plan,
                           observed_pub: torch.Tensor | None, n_surr: int, shuffle_method: str,
                           block_length: int, seed: int, rng_device, k_start: int = 0,
                           perm_batch: int = 32):
    """Surrogate statistics of the pairs of `plan`, a plan over local indices
    of the variable subset vars_host (sorted): the subset's bins are
    block-shuffled with the same orders as in the full data (rows of
    block_perm_device), so every pair's statistics equal those of the full
    computation.  Returns public-layout (m, m) sum, sumsq (float64) and
    count_ge (int32; zero when observed_pub is None)."""
    sub = var_subset(prep, vars_host)
    m = len(vars_host)
    vd = sub.plans["rows"]
    dev = prep.device
    s = torch.zeros((m, m), dtype=torch.float64, device=dev)
    s2 = torch.zeros_like(s)
    cg = torch.zeros((m, m), dtype=torch.int32, device=dev)
    obs = None if observed_pub is None else (
        observed_pub.index_select(0, vd).index_select(1, vd).contiguous())
    _stats_run(sub, plan, vd, prep.n_vars, obs, s, s2, None if obs is None else cg, n_surr,
               shuffle_method, block_length, seed, rng_device, k_start, perm_batch)
    return s, s2, cg


def surrogate_stats_pairs(prep: Prepared, pairs: np.ndarray, observed: torch.Tensor | None,
                          n_surr: int, shuffle_method: str, block_length: int, seed: int,
                          rng_device, perm_batch: int = 32):
    """Surrogate statistics of an explicit list of distinct [target, source]
    pairs (host int (P, 2), target != source) as (P,) tensors: sum, sumsq
    (float64) and count_ge (int32, against observed, float32 (P,); zero
    when observed is None).  Only the variables the pairs touch are
    shuffled, with the rows of the full draw's block orders, so each pair's
    statistics equal those of the full computation."""
    pairs = np.ascontiguousarray(pairs, dtype=np.int64).reshape(-1, 2)
    P = len(pairs)
    dev = prep.device
    if P == 0 or n_surr <= 0:
        z = torch.zeros(P, dtype=torch.float64, device=dev)
        return z, z.clone(), torch.zeros(P, dtype=torch.int32, device=dev)
    if np.any(pairs[:, 0] == pairs[:, 1]):
        raise ValueError("surrogate_stats_pairs: target == source pairs are not tested")
    involved = np.unique(pairs)
    loc = np.searchsorted(involved, pairs)
    m = len(involved)
    pub = loc[:, 1] * m + loc[:, 0]
    if len(np.unique(pub)) != P:
        raise ValueError("surrogate_stats_pairs: pairs must be distinct")
    sub = var_subset(prep, involved)
    vd = sub.plans["rows"]
    plan = te_pairlist_plan(sub.npv_host, loc, dev, wave_cap(sub))
    s = torch.zeros((m, m), dtype=torch.float64, device=dev)
    s2 = torch.zeros_like(s)
    cg = None
    obs = None
    pd = h2d(pub, dev)
    if observed is not None:
        cg = torch.zeros((m, m), dtype=torch.int32, device=dev)
        obs = torch.zeros((m, m), dtype=torch.float32, device=dev)
        obs.view(-1)[pd] = observed.to(device=dev, dtype=torch.float32)
    _stats_run(sub, plan, vd, prep.n_vars, obs, s, s2, cg, n_surr, shuffle_method,
               block_length, seed, rng_device, 0, perm_batch)
    out_cg = (torch.zeros(P, dtype=torch.int32, device=dev) if cg is None
              else cg.view(-1)[pd])
    return s.view(-1)[pd], s2.view(-1)[pd], out_cg


def surrogate_plan(prep: Prepared, targets=None, sources=None) -> Plan:
    require_narrow(prep, "the surrogate test")
    cap = surrogate_cap(prep.npv_host, targets, sources)
    if cap > SURR_MAX_CELLS:
        raise NotImplementedError(
            f"surrogate test needs b_t^2*b_s <= {SURR_MAX_CELLS} histogram cells (got {cap}); "
            f"use coarsening=True or a smaller kp")
    return get_plan(prep, targets, sources, cap_wave=cap, cap_block=0)


WAVE_CAP = 3136          # cells per wave histogram (te_hist.hip's default)
FILL_WAVE_ITEMS = 256 * 3 * 4    # items that fill the chip: 4 per resident workgroup
CHUNK_SRC, N_XCDS = 64, 8        # the planner's source chunk and gfx950's XCD count
FILL_BLOCK_ITEMS = 256 * 2 * 4
BLOCK_LDS = 80 * 1024 - 512 - 8 * 256   # as te_hist.hip's kBlockLdsTarget


def te_pairlist_plan(npv_host: np.ndarray, pairs: np.ndarray, device,
                     cap_wave: int = WAVE_CAP) -> Plan:
    """Cached _te_pairlist_plan."""
    pairs = np.ascontiguousarray(pairs, dtype=np.int64).reshape(-1, 2)
    key = ("pairs", _digest(np.asarray(npv_host, np.int32)), _digest(pairs), str(device),
           int(cap_wave))
    return _cached_plan(key, lambda: _te_pairlist_plan(npv_host, pairs, device, cap_wave))


def _te_pairlist_plan(npv_host: np.ndarray, pairs: np.ndarray, device,
                      cap_wave: int = WAVE_CAP) -> Plan:
    """TE plan for an explicit [target, source] pair list: pairs grouped by
    target, each target's sources sorted by level count; the wave-kernel
    pairs (b_t^2 b_s <= cap_wave) in items of up to 256 sources, the rest in
    block-kernel items of up to 16, with the block histogram sized like
    tnx_te_plan's (two workgroups per CU)."""
    pairs = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    info = _lib.PlanInfo()
    bt = npv_host[pairs[:, 0]].astype(np.int64)
    bs = npv_host[pairs[:, 1]].astype(np.int64)
    is_wave = bt * bt * bs <= cap_wave
    order = np.lexsort((bs, ~is_wave, pairs[:, 0]))   # target, wave first, levels ascending
    p = pairs[order]
    w = is_wave[order]
    tg, start, cnt = np.unique(p[:, 0], return_index=True, return_counts=True)
    # pairs per item: up to 256 (wave) / 16 (block), fewer for small lists so
    # that the items fill the chip (as tnx_te_plan does)
    nwp = int(w.sum())
    per_w = int(np.clip(-(-nwp // FILL_WAVE_ITEMS), 1, 256))
    per_b = int(np.clip(-(-(len(p) - nwp) // FILL_BLOCK_ITEMS), 1, 16))
    nw = np.add.reduceat(w.astype(np.int64), start) if len(start) else np.zeros(0, np.int64)
    gi = np.arange(len(tg), dtype=np.int64)
    wave = _chunk_items(gi, start, start + nw, per_w)
    block = _chunk_items(gi, start + nw, start + cnt, per_b)
    blk = ~w
    maxcols = int((npv_host[p[blk, 0]].astype(np.int64) * npv_host[p[blk, 1]]).max(initial=0))
    cap_block = min(16384, BLOCK_LDS // 4 - maxcols) & ~3
    cap_block = max(cap_block, (maxcols + 3) & ~3, 4)
    work = np.concatenate([wave, block]).astype(np.int32).reshape(-1, 4)
    info.n_wave_items, info.n_block_items = len(wave), len(block)
    info.n_pairs = len(p)
    info.cap_wave, info.cap_block, info.max_block_cols = int(cap_wave), int(cap_block), maxcols
    info.n_tgt, info.n_src = len(tg), len(p)
    wd = work if len(work) else np.zeros((1, 4), np.int32)
    return Plan(info=info, tgt_order=h2d(tg.astype(np.int32), device),
                src_order=h2d(p[:, 1].astype(np.int32), device), work=h2d(wd, device),
                work_host=work, tgt_host=tg.astype(np.int32), src_host=p[:, 1].astype(np.int32))


def _chunk_items(group: np.ndarray, beg: np.ndarray, end: np.ndarray, per: int) -> np.ndarray:
    """Work items (group, 1, b, e) splitting each [beg, end) into runs of at
    most `per`, groups in order."""
    n = np.maximum(-(-(end - beg) // per), 0)
    b = np.repeat(beg, n) + _ranges(np.zeros_like(n), n) * per
    e = np.minimum(b + per, np.repeat(end, n))
    return np.stack([np.repeat(group, n), np.ones_like(b), b, e], 1)


def pairlist_plan(npv_host: np.ndarray, pairs: np.ndarray, device, cap_wave: int,
                  per_item: int = 256) -> Plan:
    """Plan for an explicit [target, source] pair list with every pair in the
    wave kernel: pairs grouped by target, one item per target chunk of at most
    per_item sources (src_order is the concatenation of per-target source
    lists).  Every pair must satisfy b_t^2 * b_s <= cap_wave."""
    pairs = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    info = _lib.PlanInfo()
    if len(pairs) == 0:
        info.cap_wave, info.cap_block = int(cap_wave), 16128
        z = torch.zeros(1, dtype=torch.int32, device=device)
        return Plan(info=info, tgt_order=z, src_order=z, work=torch.zeros((1, 4), dtype=torch.int32,
                    device=device), work_host=np.zeros((0, 4), np.int32),
                    tgt_host=np.zeros(0, np.int32), src_host=np.zeros(0, np.int32))
    cells = npv_host[pairs[:, 0]].astype(np.int64) ** 2 * npv_host[pairs[:, 1]]
    if cells.max() > cap_wave:
        raise ValueError("a pair exceeds cap_wave histogram cells")
    order = np.lexsort((pairs[:, 1], pairs[:, 0]))
    p = pairs[order]
    tg, start, cnt = np.unique(p[:, 0], return_index=True, return_counts=True)
    items = []
    for ti, (s0, c) in enumerate(zip(start, cnt)):
        for a in range(0, c, per_item):
            items.append((ti, 1, s0 + a, s0 + min(c, a + per_item)))
    work = np.asarray(items, np.int32)
    info.n_wave_items, info.n_block_items = len(work), 0
    info.n_pairs = int((p[:, 0] != p[:, 1]).sum())
    info.cap_wave, info.cap_block, info.max_block_cols = int(cap_wave), 16128, 1
    info.n_tgt, info.n_src = len(tg), len(p)
    return Plan(info=info, tgt_order=torch.from_numpy(tg.astype(np.int32)).to(device),
                src_order=torch.from_numpy(p[:, 1].astype(np.int32)).to(device),
                work=torch.from_numpy(work).to(device), work_host=work,
                tgt_host=tg.astype(np.int32), src_host=p[:, 1].astype(np.int32))
