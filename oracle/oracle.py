"""ctypes bindings to the C oracle plus numpy restatements (test infrastructure).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module.  Reference citations are paths inside the cxinsys/tenex tree.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "libtnx_oracle.so")
_lib = None


def build(force: bool = False) -> str:
    """Compile tnx_oracle.c with oracle/Makefile (gcc)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []),
                       check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        p = C.c_void_p
        i64 = C.c_int64
        L.orc_np_rowsum_f32.restype = C.c_float
        L.orc_np_rowsum_f32.argtypes = [p, i64]
        L.orc_fsbw_l.argtypes = [p, i64, i64, C.c_float, p, p, p, p]
        L.orc_remap.argtypes = [p, i64, i64, p, p]
        L.orc_joint_counts.argtypes = [p, i64, i64, C.c_int, p, i64, i64, p]
        L.orc_gene_entropies.argtypes = [p, i64, i64, C.c_int, p, p, p]
        L.orc_te_pairs.argtypes = [p, i64, i64, C.c_int, p, p, i64, p, p]
        L.orc_te_b2.argtypes = [p, i64, i64, C.c_int, p, p, p, p]
        L.orc_surrogate.argtypes = [p, i64, i64, C.c_int, p, p, i64, i64, i64,
                                    p, i64, p, p, p, p, p]
        L.orc_mlp_n_params.restype = i64
        L.orc_mlp_n_params.argtypes = [i64, i64, i64]
        L.orc_mlp_forward.argtypes = [p, i64, i64, i64, p, i64, p]
        L.orc_mlp_loss_grad.restype = C.c_double
        L.orc_mlp_loss_grad.argtypes = [p, i64, i64, i64, p, p, p, i64, p]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _chk(rc, what):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed with status {rc}")


# ── discretisation ─────────────────────────────────────────────────────────

def np_rowsum_f32(a: np.ndarray) -> np.float32:
    a = np.ascontiguousarray(a, dtype=np.float32)
    return np.float32(lib().orc_np_rowsum_f32(_ptr(a), a.size))


def fsbw_l(arr: np.ndarray, kp: float = 0.5):
    """C restatement of bin_fsbw_l_numpy (preprocess.py:61-88) + clamp(min=0)
    (preprocess.py:622).  Returns (raw int32 (n,T), n_bins, std, shifted_min).
    Pinned bit-exact against fsbw_l_numpy (numpy itself) in tests."""
    arr = np.ascontiguousarray(arr, dtype=np.float32)
    n, T = arr.shape
    raw = np.empty((n, T), np.int32)
    nb = np.empty(n, np.int32)
    sd = np.empty(n, np.float32)
    sm = np.empty(n, np.float32)
    _chk(lib().orc_fsbw_l(_ptr(arr), n, T, float(kp), _ptr(raw), _ptr(nb),
                          _ptr(sd), _ptr(sm)), "fsbw_l")
    return raw, nb, sd, sm


def fsbw_l_numpy(arr: np.ndarray, kp: float = 0.5):
    """numpy restatement of bin_fsbw_l_numpy (preprocess.py:61-88) followed by
    clamp(min=0) (preprocess.py:622); uses the same numpy calls, so it is the
    library's own arithmetic."""
    arr = np.asarray(arr, dtype=np.float32)
    stds = np.std(arr, axis=1, ddof=1)
    stds = np.where((stds == 0) | np.isnan(stds), 1.0, stds)
    mins = np.min(arr, axis=1)
    maxs = np.max(arr, axis=1)
    n_bins = np.ceil((maxs - mins) / stds).astype(np.int32)
    shifted_min = mins - kp * stds
    bin_arr = np.floor((arr.T - shifted_min) / stds).T.astype(np.int32)
    return np.maximum(bin_arr, 0), n_bins, stds.astype(np.float32), shifted_min


def remap(raw: np.ndarray):
    """C restatement of remap_bins (preprocess.py:401-459): per-gene dense rank.
    Returns (dense int32, n_per_var int32, b_max)."""
    raw = np.ascontiguousarray(raw, dtype=np.int32)
    n = raw.shape[0]
    TK = raw.size // max(n, 1)
    dense = np.empty_like(raw)
    npv = np.empty(n, np.int32)
    _chk(lib().orc_remap(_ptr(raw), n, TK, _ptr(dense), _ptr(npv)), "remap")
    return dense, npv, int(npv.max()) if n else 0


def remap_numpy(raw: np.ndarray):
    """numpy.unique(return_inverse) per gene: the semantics of torch.unique
    used by remap_bins (preprocess.py:453-456)."""
    raw = np.asarray(raw)
    out = np.empty(raw.shape, np.int32)
    npv = np.empty(raw.shape[0], np.int32)
    for g in range(raw.shape[0]):
        u, inv = np.unique(raw[g], return_inverse=True)
        out[g] = inv.reshape(raw[g].shape)
        npv[g] = len(u)
    return out, npv, int(npv.max()) if len(npv) else 0


def discretize(arr: np.ndarray, kp: float = 0.5):
    """FSBW-L (numpy-exact) + clamp + dense remap: (bins int32, n_bins, npv, b_max)."""
    raw, nb, _, _ = fsbw_l(arr, kp)
    dense, npv, bmax = remap(raw)
    return dense, nb, npv, bmax


def coarsen_bins(bins: np.ndarray, npv: np.ndarray, limit: int):
    """coarsen_bins (tenex/preprocess.py:462-551, CPU branch :531-534):
    variables with more than `limit` levels map level v to v * limit // ng
    and report exactly `limit` levels; the rest are unchanged.  Returns
    (bins int32, npv int32, b_max)."""
    bins = np.array(bins, dtype=np.int64, copy=True)
    npv = np.array(npv, dtype=np.int64, copy=True)
    big = npv > limit
    bins[big] = bins[big] * int(limit) // npv[big][:, None]
    npv[big] = int(limit)
    return bins.astype(np.int32), npv.astype(np.int32), int(npv.max())


def smem_bin_limit(smem_optin: int, warps: int = 32) -> int:
    """compute_smem_bin_limit (tenex/kernels/adaptive_smem.py:701-726): the
    largest B >= 2 with (B^3 + 2 B^2 + B + warps) * 4 <= smem_optin, by
    downward search from the cube-root estimate."""
    B = max(2, int((smem_optin / 4) ** (1.0 / 3.0)))
    while B > 2 and (B ** 3 + 2 * B ** 2 + B + warps) * 4 > smem_optin:
        B -= 1
    return B


# ── pairs ──────────────────────────────────────────────────────────────────

def pairfree_pairs(n: int) -> np.ndarray:
    """Pair-free linear order (full_smem.cu:190-195): pair id = tgt*(n-1)+s,
    src = s + (s >= tgt).  Returns int64 (n(n-1), 2) [target, source]."""
    ids = np.arange(n * (n - 1), dtype=np.int64)
    tgt = ids // (n - 1)
    src = ids % (n - 1)
    src = src + (src >= tgt)
    return np.stack([tgt, src], axis=1)


def build_pairs_engine(n: int, source_idx=None) -> np.ndarray:
    """TransferEntropyEngine._build_pairs (transferentropy.py:456-498):
    target-major [target, source] with a stable sort by target."""
    if source_idx is not None:
        srcs = np.asarray(source_idx, dtype=np.int64)
        blocks = []
        for t in range(n):
            s = srcs[srcs != t]
            if len(s):
                b = np.empty((len(s), 2), np.int64)
                b[:, 0] = t
                b[:, 1] = s
                blocks.append(b)
        pairs = np.concatenate(blocks) if blocks else np.empty((0, 2), np.int64)
    else:
        pairs = pairfree_pairs(n)
    return pairs[np.argsort(pairs[:, 0], kind="stable")]


def pairfree_to_matrix(ent: np.ndarray, n: int) -> np.ndarray:
    """_assemble_result_pairfree (transferentropy.py:1394-1414): returns
    M[i, j] = TE(i -> j)."""
    m = np.zeros((n, n), dtype=ent.dtype)
    e2 = ent.reshape(n, n - 1)
    for i in range(n):
        m[i, :i] = e2[i, :i]
        m[i, i + 1:] = e2[i, i:]
    return m.T


# ── TE ─────────────────────────────────────────────────────────────────────

def joint_counts(bins, npv, tgt, src, tau=1):
    bins = np.ascontiguousarray(bins, dtype=np.int32)
    npv = np.ascontiguousarray(npv, dtype=np.int32)
    n, T = bins.shape[:2]
    bt, bs = int(npv[tgt]), int(npv[src])
    out = np.empty(bt * bt * bs, np.int32)
    _chk(lib().orc_joint_counts(_ptr(bins), n, T, tau, _ptr(npv), tgt, src,
                                _ptr(out)), "joint_counts")
    return out.reshape(bt, bt, bs)


def gene_entropies(bins, npv, tau=1):
    bins = np.ascontiguousarray(bins, dtype=np.int32)
    npv = np.ascontiguousarray(npv, dtype=np.int32)
    n, T = bins.shape[:2]
    s1 = np.empty(n)
    s2 = np.empty(n)
    _chk(lib().orc_gene_entropies(_ptr(bins), n, T, tau, _ptr(npv), _ptr(s1),
                                  _ptr(s2)), "gene_entropies")
    return s1, s2


def te_pairs(bins, npv, pairs, tau=1, direct=False):
    """FP64 TE(src -> tgt) for pairs [target, source] (entropy S-form); with
    direct=True also returns the reference's log-ratio form."""
    bins = np.ascontiguousarray(bins, dtype=np.int32)
    if bins.ndim == 3:
        bins = np.ascontiguousarray(bins[:, :, 0])
    npv = np.ascontiguousarray(npv, dtype=np.int32)
    pairs = np.ascontiguousarray(pairs, dtype=np.int64)
    n, T = bins.shape
    te = np.empty(len(pairs))
    td = np.empty(len(pairs)) if direct else None
    _chk(lib().orc_te_pairs(_ptr(bins), n, T, tau, _ptr(npv), _ptr(pairs),
                            len(pairs), _ptr(te), _ptr(td)), "te_pairs")
    return (te, td) if direct else te


def te_pairs_sparse(bins, pairs, tau=1):
    """FP64 TE of [target, source] pairs from the occupied joint states only
    (np.unique counts of (x_{t+tau}, x_t, y_t), (x_t, y_t), (x_{t+tau}, x_t)
    and x_t; S-form), for level counts whose dense b_t^2 b_s tables would not
    fit.  Pinned against te_pairs (dense counts) in the tests."""
    bins = np.asarray(bins)
    if bins.ndim == 3:
        bins = bins[:, :, 0]
    T = bins.shape[1]
    L = T - tau

    def S(*cols):
        c = np.unique(np.stack(cols), axis=1, return_counts=True)[1].astype(np.float64)
        c = c[c > 1]
        return float(np.sum(c * np.log2(c)))
    out = np.empty(len(pairs))
    for i, (t, s_) in enumerate(np.asarray(pairs)):
        xf, xp, yp = bins[t, tau:], bins[t, :L], bins[s_, :L]
        out[i] = (S(xf, xp, yp) + S(xp) - S(xf, xp) - S(xp, yp)) / L
    return out


def te_matrix(bins, npv, tau=1):
    """FP64 public-layout matrix M[i, j] = TE(i -> j), diagonal 0."""
    n = bins.shape[0]
    pairs = pairfree_pairs(n)
    return pairfree_to_matrix(te_pairs(bins, npv, pairs, tau), n)


def te_b2(bins, tau=1, counts=False):
    """GEMM-B2 closed form (gemm_b2.py:210-285) in FP64: internal layout
    te[i, j] = TE(j -> i).  counts=True also returns C, D, E (int64)."""
    bins = np.ascontiguousarray(bins, dtype=np.int32)
    if bins.ndim == 3:
        bins = np.ascontiguousarray(bins[:, :, 0])
    n, T = bins.shape
    te = np.empty((n, n))
    if counts:
        Cm, Dm, Em = (np.empty((n, n), np.int64) for _ in range(3))
    else:
        Cm = Dm = Em = None
    _chk(lib().orc_te_b2(_ptr(bins), n, T, tau, _ptr(te), _ptr(Cm), _ptr(Dm),
                         _ptr(Em)), "te_b2")
    return (te, Cm, Dm, Em) if counts else te


# ── surrogates ─────────────────────────────────────────────────────────────

def surrogate_stats(bins, npv, perm, block_length, pairs, obs, tau=1,
                    return_te_k=False):
    """FP64 per-pair (sum, sumsq, count_ge[, te_k]) over the surrogates in
    perm (n_surr, n_vars, n_blocks) int32 (full_smem_surrogate_test.cu:94-246)."""
    bins = np.ascontiguousarray(bins, dtype=np.int32)
    if bins.ndim == 3:
        bins = np.ascontiguousarray(bins[:, :, 0])
    npv = np.ascontiguousarray(npv, dtype=np.int32)
    perm = np.ascontiguousarray(perm, dtype=np.int32)
    pairs = np.ascontiguousarray(pairs, dtype=np.int64)
    obs = np.ascontiguousarray(obs, dtype=np.float32)
    n, T = bins.shape
    ns, nv, nb = perm.shape
    assert nv == n
    P = len(pairs)
    s = np.empty(P)
    s2 = np.empty(P)
    cg = np.empty(P, np.int32)
    tk = np.empty((P, ns)) if return_te_k else None
    _chk(lib().orc_surrogate(_ptr(bins), n, T, tau, _ptr(npv), _ptr(perm), ns, nb,
                             int(block_length), _ptr(pairs), P, _ptr(obs), _ptr(s),
                             _ptr(s2), _ptr(cg), _ptr(tk)), "surrogate")
    return (s, s2, cg, tk) if return_te_k else (s, s2, cg)


def block_perm(n_vars, T, shuffle_method, block_length, k_start, k_end, seed,
               device="cpu"):
    """_build_block_perm (surrogate_test.py:165-207) with torch's generator on
    `device`: perm[k] = argsort(rand(n_vars, n_blocks)) after
    manual_seed(seed + k).  Returns (perm int32 numpy, eff_block_length, n_blocks)."""
    import torch
    rnd_all = (shuffle_method == "random" or block_length <= 1 or block_length >= T)
    if rnd_all:
        eff, nb = 1, T
    else:
        eff, nb = int(block_length), T // int(block_length)
    out = np.empty((k_end - k_start, n_vars, nb), np.int32)
    gen = torch.Generator(device=device)
    for ki, k in enumerate(range(k_start, k_end)):
        gen.manual_seed(seed + k)
        r = torch.rand(n_vars, nb, device=device, generator=gen)
        out[ki] = torch.argsort(r, dim=1).to(torch.int32).cpu().numpy()
    return out, eff, nb


def surrogate_decision(observed, sum_s, sumsq_s, count_ge, n_surr, p_method="parametric",
                       fdr=0.01, mask=None):
    """numpy restatement of SurrogateTestMethod.infer's statistics
    (surrogate_test.py:890-929) on (n, n) public-layout matrices.
    Returns dict(mean, std, eff, p, adj, sig)."""
    from scipy.special import erfc
    observed = np.asarray(observed, np.float32)
    n = observed.shape[0]
    mean = (sum_s / n_surr).astype(np.float32)
    var = sumsq_s / n_surr - (sum_s / n_surr) ** 2
    std = np.sqrt(np.maximum(var, 0.0)).astype(np.float32)
    eff = (observed - mean).astype(np.float32)
    if p_method == "parametric":
        eps = np.float32(1e-12)
        z = (observed - mean) / np.maximum(std, eps)
        p = (0.5 * erfc(z / np.sqrt(2.0))).astype(np.float32)
    else:
        p = ((1.0 + count_ge) / (1.0 + n_surr)).astype(np.float32)
    np.fill_diagonal(p, 1.0)
    if mask is None:
        mask = ~np.eye(n, dtype=bool)
    adj_off = bh_fdr(p[mask])
    sig = np.zeros((n, n), bool)
    sig[mask] = adj_off < fdr
    sig &= eff > 0
    adj = np.ones((n, n), np.float32)
    adj[mask] = adj_off
    return dict(mean=mean, std=std, eff=eff, p=p, adj=adj, sig=sig)


def bh_fdr(p: np.ndarray) -> np.ndarray:
    """bh_fdr_gpu (inference/__init__.py:115-126) restated in numpy float32:
    sorted*m/rank, clamp(max=1), reverse cumulative min, scatter back.
    Ties among equal p-values do not change the output."""
    p = np.asarray(p, np.float32)
    m = p.shape[0]
    if m == 0:
        return p
    idx = np.argsort(p, kind="stable")
    sp = p[idx]
    rank = np.arange(1, m + 1, dtype=np.float32)
    adj = np.minimum((sp * np.float32(m)) / rank, np.float32(1.0)).astype(np.float32)
    adj = np.minimum.accumulate(adj[::-1])[::-1]
    out = np.empty_like(adj)
    out[idx] = adj
    return out


def fdr_threshold(te: np.ndarray, pairs: np.ndarray, fdr: float):
    """fdr_threshold_gpu (inference/__init__.py:129-138) restated in numpy:
    z = (te - mean) / std with the unbiased (n - 1) standard deviation,
    p = erfc(z / sqrt 2) / 2, BH-adjusted (bh_fdr above), kept where the
    adjusted p < fdr; a constant vector keeps nothing.  Mean and std are
    taken in float64 and rounded to float32 (torch accumulates float32
    reductions in wider precision), p in float64 rounded to float32."""
    from scipy.special import erfc
    te = np.asarray(te, np.float32)
    std = np.float32(te.astype(np.float64).std(ddof=1))
    if std == 0:
        return te[:0], pairs[:0]
    mean = np.float32(te.astype(np.float64).mean())
    z = ((te - mean) / std).astype(np.float32)
    p = (0.5 * erfc(z.astype(np.float64) / np.sqrt(2.0))).astype(np.float32)
    keep = bh_fdr(p) < np.float32(fdr)
    return te[keep], pairs[keep]


# ── DPI trimming ───────────────────────────────────────────────────────────

def trim_dpi(A: np.ndarray, batch_size: int = 512, trim_threshold: float = 0.0):
    """_trim (tenex/inference/fdr.py:61-133) restated in numpy float32 on the
    thresholded network A (n x n): mediators are rows with positive sum
    (:69-71); no mediator returns the input edges unchanged (:74-75); per
    batch of rows, the active mediators are the batch's nonzero columns
    that are mediators (:92-95); a batch without one keeps every nonzero
    entry (:97-103); otherwise indirect = max(0, max over active k of
    min(A[i,k], A[k,j])) (:106-114), plus trim_threshold when nonzero
    (:116-117), and entries >= indirect survive (:119).  Returns
    (pairs (m, 2) int64 row-major, values float32), or None for the
    unchanged-input case."""
    A = np.asarray(A, np.float32)
    n = A.shape[0]
    mediator = A.sum(axis=1, dtype=np.float32) > 0
    if not mediator.any():
        return None
    thr = np.float32(trim_threshold)
    pairs, vals = [], []
    for r0 in range(0, n, batch_size):
        batch = A[r0:r0 + batch_size]
        if not (batch != 0).any():
            continue
        cols = np.unique(np.nonzero(batch)[1])
        active = cols[mediator[cols]]
        if len(active) == 0:
            kept = batch
        else:
            indirect = np.zeros_like(batch)
            for k in active:
                indirect = np.maximum(indirect, np.minimum(batch[:, k:k + 1], A[k][None, :]))
            if trim_threshold != 0.0:
                indirect = (indirect + thr).astype(np.float32)
            kept = np.where(batch >= indirect, batch, np.float32(0))
        r, c = np.nonzero(kept)
        pairs.append(np.stack([r + r0, c], 1).astype(np.int64))
        vals.append(kept[r, c])
    if not pairs:
        return np.zeros((0, 2), np.int64), np.zeros(0, np.float32)
    return np.concatenate(pairs), np.concatenate(vals).astype(np.float32)


# ── TRACE ──────────────────────────────────────────────────────────────────

TRACE_P = np.uint64(1000000007)


def trace_hash(rows) -> np.ndarray:
    """Horner polynomial hash of the rows (d, L) with wrapping 64-bit
    arithmetic: h = h * P + v per row in order (trace.py:497-502,
    trace_sort_entropy.cu:46-50); uint64 bits equal the reference's int64."""
    rows = np.atleast_2d(np.asarray(rows)).astype(np.uint64)
    h = np.zeros(rows.shape[1], np.uint64)
    with np.errstate(over="ignore"):
        for r in rows:
            h = h * TRACE_P + r
    return h


def _pow_p(d: int) -> np.uint64:
    v = np.uint64(1)
    with np.errstate(over="ignore"):
        for _ in range(d):
            v = v * TRACE_P
    return v


def trace_keys(bins, var: int, conn, mode: int, tau: int = 1):
    """The four key vectors of one TRACE job as the reference's 4-way hash
    fusion forms them (compute_outte_cpu / compute_inte_cpu, trace.py:475-547):
    mode 0 (OutTE, conn = X's targets): h(Yf,Yp), h(Yp), h(Yf,Xp,Yp), h(Xp,Yp);
    mode 1 (InTE, conn = X's sources): h(Xf,Xp), h(Xp), h(Xf,Xp,Yp), h(Xp,Yp)."""
    bins = np.asarray(bins)
    if bins.ndim == 3:
        bins = bins[:, :, 0]
    T = bins.shape[1]
    L = T - tau
    conn = [int(c) for c in conn]
    Pd = _pow_p(len(conn))
    xp = bins[var, :L].astype(np.uint64)
    with np.errstate(over="ignore"):
        if mode == 0:
            h_yf = trace_hash(bins[conn, tau:tau + L])
            h_yp = trace_hash(bins[conn, :L])
            return (h_yf * Pd + h_yp, h_yp, (h_yf * TRACE_P + xp) * Pd + h_yp, xp * Pd + h_yp)
        xf = bins[var, tau:tau + L].astype(np.uint64)
        h_yp = trace_hash(bins[conn, :L])
        h_xfxp = xf * TRACE_P + xp
        return (h_xfxp, xp, h_xfxp * Pd + h_yp, xp * Pd + h_yp)


def _entropy_np(h) -> float:
    _, counts = np.unique(h, return_counts=True)
    p = counts / len(h)
    return -float(np.sum(p * np.log2(p)))


def trace_te(bins, var: int, conn, mode: int, tau: int = 1) -> float:
    """OutTE (mode 0) or InTE (mode 1) of one variable as the reference's CPU
    path computes it: FP64 plug-in entropies of the hashed states,
    max(e0 - e1 - e2 + e3, 0) (trace.py:505-510, :541-547); 0 without
    connections."""
    if len(conn) == 0:
        return 0.0
    e = [_entropy_np(k) for k in trace_keys(bins, var, conn, mode, tau)]
    return max(e[0] - e[1] - e[2] + e[3], 0.0)


def trace_te_tuples(bins, var: int, conn, mode: int, tau: int = 1) -> float:
    """The same quantity from the states themselves (np.unique over rows of
    the (L, k) state matrix, no hashing), the structured-dtype entropy the
    reference validates its hash against (trace.py:417-440)."""
    bins = np.asarray(bins)
    if bins.ndim == 3:
        bins = bins[:, :, 0]
    if len(conn) == 0:
        return 0.0
    L = bins.shape[1] - tau
    conn = [int(c) for c in conn]
    f = lambda rows: bins[rows, tau:tau + L]  # noqa: E731
    pst = lambda rows: bins[rows, :L]  # noqa: E731

    def H(*blocks):
        m = np.concatenate([np.atleast_2d(b) for b in blocks], axis=0).T
        _, c = np.unique(m, axis=0, return_counts=True)
        p = c / L
        return -float(np.sum(p * np.log2(p)))
    if mode == 0:
        e = (H(f(conn), pst(conn)), H(pst(conn)), H(f(conn), pst([var]), pst(conn)),
             H(pst([var]), pst(conn)))
    else:
        e = (H(f([var]), pst([var])), H(pst([var])), H(f([var]), pst([var]), pst(conn)),
             H(pst([var]), pst(conn)))
    return max(e[0] - e[1] - e[2] + e[3], 0.0)


def trace_network(te_matrix, n_surrogates: int = 100, significance: float = 2.0,
                  seed: int = 42, gpu: bool = True):
    """build_causal_network (trace.py:550-612): the threshold mean +
    significance * std (numpy float32, ddof 0) of n_surrogates off-diagonal
    entries drawn with default_rng(seed); edges te[i, j] > threshold with the
    diagonal zeroed first (the GPU branch, gpu=True) or excluded after the
    comparison (the CPU branch).  Returns (network dict target -> sorted
    sources, threshold)."""
    te = np.asarray(te_matrix, np.float32)
    n = te.shape[0]
    rng = np.random.default_rng(seed)
    rows = rng.integers(0, n, size=n_surrogates)
    cols = rng.integers(0, n, size=n_surrogates)
    m = rows == cols
    cols[m] = (cols[m] + 1) % n
    sv = te[rows, cols]
    thr = float(sv.mean() + significance * sv.std())
    if gpu:
        t2 = te.copy()
        np.fill_diagonal(t2, 0.0)
        edges = t2 > thr
    else:
        edges = te > thr
        np.fill_diagonal(edges, False)
    src, tgt = np.nonzero(edges)
    network = {j: [] for j in range(n)}
    order = np.argsort(tgt, kind="stable")
    for s_, t_ in zip(src[order], tgt[order]):
        network[int(t_)].append(int(s_))
    return network, thr


def trace_infer(te_matrix, bins, tau: int = 1, n_surrogates: int = 100,
                significance: float = 2.0, seed: int = 42, gpu: bool = True):
    """TRACEMethod.infer (trace.py:621-760) on the CPU: network, then OutTE
    over each variable's targets and InTE over its sources (float32).
    Returns (outte, inte, network, edge pairs [src, tgt] target-major)."""
    network, _ = trace_network(te_matrix, n_surrogates, significance, seed, gpu)
    n = len(network)
    outgoing = {i: [] for i in range(n)}
    pairs = []
    for j in range(n):
        for i in network[j]:
            outgoing[i].append(j)
            pairs.append((i, j))
    outte = np.zeros(n, np.float32)
    inte = np.zeros(n, np.float32)
    for i in range(n):
        outte[i] = trace_te(bins, i, sorted(outgoing[i]), 0, tau)
        inte[i] = trace_te(bins, i, network[i], 1, tau)
    return outte, inte, network, np.asarray(pairs, np.int64).reshape(-1, 2)


# ── MLP ────────────────────────────────────────────────────────────────────

def mlp_n_params(D, H, O):
    return int(lib().orc_mlp_n_params(D, H, O))


def mlp_forward(params, D, H, O, X):
    params = np.ascontiguousarray(params, np.float64)
    X = np.ascontiguousarray(X, np.float64)
    Y = np.empty((X.shape[0], O))
    _chk(lib().orc_mlp_forward(_ptr(params), D, H, O, _ptr(X), X.shape[0], _ptr(Y)),
         "mlp_forward")
    return Y


def mlp_loss_grad(params, D, H, O, X, Tgt, wout=None):
    params = np.ascontiguousarray(params, np.float64)
    X = np.ascontiguousarray(X, np.float64)
    Tgt = np.ascontiguousarray(Tgt, np.float64)
    w = None if wout is None else np.ascontiguousarray(wout, np.float64)
    g = np.empty_like(params)
    loss = lib().orc_mlp_loss_grad(_ptr(params), D, H, O, _ptr(X), _ptr(Tgt), _ptr(w),
                                   X.shape[0], _ptr(g))
    return float(loss), g
