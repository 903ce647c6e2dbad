"""Learned surrogate null (no reference counterpart).

A 3-layer MLP (width 64, bf16 weights, fp32 accumulation; mlp.hip) predicts
each pair's surrogate-null mean and standard deviation of TE from per-gene
marginal features (level count, H(X_t), H(X_t+1 | X_t), zero and modal
fractions, ...; tnx_gene_prep).  It is trained on the GPU by the project's
own forward / backward / Adam kernels on plain surrogate statistics of a
random subset of pairs.  Pairs whose predicted z lies within `band` of the
BH threshold are re-run with the plain surrogate test, which stays both the
fallback and the judge; every other pair keeps its predicted null.  The band
is max(band, band_sigmas * RMS z error on held-out plain pairs): 1.0 when the
model is accurate, wider when it is not.

Because surrogate k of a variable depends only on (seed, k, variable), the
plain statistics of a re-run pair equal those of the full plain test.
"""
from __future__ import annotations

import ctypes as C
import math
import time

import numpy as np
import torch

from .. import _lib, core
from .._log import vprint
from ..core import call, ptr
from . import make_grn

H = 64


def _bf16(t: torch.Tensor) -> torch.Tensor:
    """Round-to-nearest-even fp32 -> bf16 bit patterns (int16 storage)."""
    u = t.contiguous().view(torch.int32).to(torch.int64) & 0xFFFFFFFF
    u = u + 0x7FFF + ((u >> 16) & 1)
    return ((u >> 16) & 0xFFFF).to(torch.int32).to(torch.int16)


class NullMLP:
    """Device-resident MLP: fp32 master parameters (W1, b1, W2, b2, W3, b3)
    plus the bf16 weight copies the kernels read."""

    def __init__(self, D: int, O: int = 2, device=None, seed: int = 0):
        self.D, self.O = D, O
        self.dev = core.require_gpu(device)
        g = torch.Generator().manual_seed(seed)
        shapes = [(H, D), (H,), (H, H), (H,), (O, H), (O,)]
        parts = []
        for s in shapes:
            if len(s) == 2:
                parts.append((torch.randn(s, generator=g) * math.sqrt(2.0 / s[1])).ravel())
            else:
                parts.append(torch.zeros(s))
        self.master = torch.cat(parts).to(torch.float32).to(self.dev)
        self.n_params = self.master.numel()
        self.m1 = torch.zeros_like(self.master)
        self.m2 = torch.zeros_like(self.master)
        self.t = 0
        self.W1 = torch.empty(H * D, dtype=torch.int16, device=self.dev)
        self.W2 = torch.empty(H * H, dtype=torch.int16, device=self.dev)
        self.W3 = torch.empty(O * H, dtype=torch.int16, device=self.dev)
        self.b1 = torch.empty(H, dtype=torch.float32, device=self.dev)
        self.b2 = torch.empty(H, dtype=torch.float32, device=self.dev)
        self.b3 = torch.empty(O, dtype=torch.float32, device=self.dev)
        self.sync_copies()

    def views(self):
        D, O = self.D, self.O
        o = np.cumsum([0, H * D, H, H * H, H, O * H, O])
        return [self.master[o[i]:o[i + 1]] for i in range(6)]

    def sync_copies(self):
        w1, b1, w2, b2, w3, b3 = self.views()
        self.W1.copy_(_bf16(w1))
        self.W2.copy_(_bf16(w2))
        self.W3.copy_(_bf16(w3))
        self.b1.copy_(b1)
        self.b2.copy_(b2)
        self.b3.copy_(b3)

    def struct(self) -> _lib.Mlp:
        return _lib.Mlp(W1=ptr(self.W1), W2=ptr(self.W2), W3=ptr(self.W3), b1=ptr(self.b1),
                        b2=ptr(self.b2), b3=ptr(self.b3), D=self.D, H=H, O=self.O)

    def _s(self):
        return torch.cuda.current_stream(self.dev).cuda_stream

    def forward(self, X: torch.Tensor) -> torch.Tensor:
        X = X.contiguous().to(torch.float32)
        Y = torch.empty((X.shape[0], self.O), dtype=torch.float32, device=self.dev)
        m = self.struct()
        call("tnx_mlp_forward", C.byref(m), ptr(X), X.shape[0], ptr(Y), self._s())
        return Y

    def loss_grad(self, X, T, w=None):
        X = X.contiguous().to(torch.float32)
        T = T.contiguous().to(torch.float32)
        grad = torch.zeros(self.n_params, dtype=torch.float32, device=self.dev)
        loss = torch.zeros(1, dtype=torch.float32, device=self.dev)
        m = self.struct()
        call("tnx_mlp_backward", C.byref(m), ptr(X), ptr(T), ptr(w), X.shape[0], ptr(grad),
             ptr(loss), self._s())
        return loss, grad

    def adam_step(self, grad, lr=3e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.t += 1
        m = self.struct()
        call("tnx_mlp_adam", C.byref(m), ptr(self.master), ptr(grad), ptr(self.m1), ptr(self.m2),
             self.n_params, float(lr), float(b1), float(b2), float(eps), self.t, self._s())

    def fit(self, X, T, steps=400, lr=3e-3, batch=8192, seed=0):
        n = X.shape[0]
        g = torch.Generator(device="cpu").manual_seed(seed)
        losses = []
        for s in range(steps):
            if n > batch:
                idx = torch.randint(0, n, (batch,), generator=g).to(self.dev)
                xb, tb = X[idx], T[idx]
            else:
                xb, tb = X, T
            loss, grad = self.loss_grad(xb, tb)
            self.adam_step(grad, lr=lr * (0.5 * (1 + math.cos(math.pi * s / steps))) + 1e-5)
            if s % 50 == 0 or s == steps - 1:
                losses.append(float(loss.item()))
        return losses


def pair_features(prep: core.Prepared, tgt: np.ndarray, src: np.ndarray) -> torch.Tensor:
    """Features of pairs: per-gene marginals of source and target plus the
    log of the plug-in bias scale b_t (b_t - 1)(b_s - 1) / (2 L)."""
    dev = prep.device
    t = torch.from_numpy(np.asarray(tgt, np.int64)).to(dev)
    s = torch.from_numpy(np.asarray(src, np.int64)).to(dev)
    f = prep.feat
    bt = f[t, 0]
    bs = f[s, 0]
    bias = torch.log((bt * (bt - 1).clamp(min=1) * (bs - 1).clamp(min=1)) / (2.0 * prep.L))
    return torch.cat([f[s], f[t], bias[:, None]], dim=1).contiguous()


def _norm_stats(X: torch.Tensor):
    mu = X.mean(0)
    sd = X.std(0).clamp(min=1e-6)
    return mu, sd


def plain_stats_pairs(prep, pairs, obs_pub, n_surr, shuffle_method, block_length, seed,
                      rng_device, perm_batch=32):
    """Plain surrogate (sum, sumsq, count) of an explicit [target, source]
    pair list, as (P,) arrays (float64, float64, int32)."""
    n, T = prep.n_vars, prep.T
    dev = prep.device
    cap = core.surrogate_cap(prep.npv_host, pairs[:, 0], pairs[:, 1])
    plan = core.pairlist_plan(prep.npv_host, pairs, dev, cap)
    s = torch.zeros((n, n), dtype=torch.float64, device=dev)
    s2 = torch.zeros_like(s)
    cg = torch.zeros((n, n), dtype=torch.int32, device=dev)
    for k0 in range(0, n_surr, perm_batch):
        k1 = min(k0 + perm_batch, n_surr)
        perm, eff, nb = core.block_perm_device(n, T, shuffle_method, block_length, k0, k1, seed,
                                               rng_device, dev)
        core.surrogate_run(prep, plan, perm, eff, obs_pub, s, s2, cg)
    slot_s = torch.from_numpy(pairs[:, 1]).to(dev)
    slot_t = torch.from_numpy(pairs[:, 0]).to(dev)
    return (s[slot_s, slot_t].cpu().numpy(), s2[slot_s, slot_t].cpu().numpy(),
            cg[slot_s, slot_t].cpu().numpy())


def learned_surrogate_test(prep, observed_te, obs_d, variable_names, n_surrogates,
                           shuffle_method, block_length, seed, rng_device, p_method, fdr, mask,
                           src_idx, dev, train_frac: float = 0.05, band: float = 1.0,
                           steps: int = 400, lr: float = 3e-3, min_train: int = 2000,
                           model_seed: int = 0, band_sigmas: float = 3.0):
    """Learned-null surrogate test; returns a SurrogateTestResult with an
    extra `learned_info` dict (training size, re-run size, timings, losses)."""
    from scipy.special import erfc

    from . import bh_fdr_gpu
    from .surrogate_test import SurrogateTestResult
    t0 = time.perf_counter()
    n = prep.n_vars
    pub = np.argwhere(mask)                  # [source, target] of tested pairs
    if src_idx is not None:
        pub = pub[np.isin(pub[:, 0], src_idx)]
    pairs = np.stack([pub[:, 1], pub[:, 0]], 1).astype(np.int64)   # [target, source]
    P = len(pairs)
    rng = np.random.default_rng(seed)
    n_train = min(P, max(min_train, int(train_frac * P)))
    tr = rng.choice(P, size=n_train, replace=False)
    # 1. plain surrogates on the training subset
    s, s2, cg = plain_stats_pairs(prep, pairs[tr], obs_d, n_surrogates, shuffle_method,
                                  block_length, seed, rng_device)
    mean_tr = s / n_surrogates
    std_tr = np.sqrt(np.maximum(s2 / n_surrogates - mean_tr ** 2, 0.0))
    t_train_plain = time.perf_counter() - t0
    # 2. train the MLP on (mean, log std)
    X_all = pair_features(prep, pairs[:, 0], pairs[:, 1])
    mu, sd = _norm_stats(X_all)
    Xn = (X_all - mu) / sd
    Y = np.stack([mean_tr, np.log(np.maximum(std_tr, 1e-9))], 1)
    ymu, ysd = Y.mean(0), np.maximum(Y.std(0), 1e-9)
    Yn = torch.from_numpy(((Y - ymu) / ysd).astype(np.float32)).to(dev)
    # hold out a fifth of the plain pairs to measure the model's z error
    n_fit = max(1, int(0.8 * n_train))
    fit_rows = torch.from_numpy(np.arange(n_fit)).to(dev)
    trd = torch.from_numpy(tr).to(dev)
    model = NullMLP(Xn.shape[1], 2, dev, seed=model_seed)
    losses = model.fit(Xn[trd][fit_rows], Yn[fit_rows], steps=steps, lr=lr, seed=seed)
    t_fit = time.perf_counter() - t0 - t_train_plain
    # 3. predict for every pair
    Yp = model.forward(Xn).cpu().numpy().astype(np.float64) * ysd + ymu
    mean_p = Yp[:, 0]
    std_p = np.exp(Yp[:, 1])
    obs = observed_te[pub[:, 0], pub[:, 1]].astype(np.float64)
    z = (obs - mean_p) / np.maximum(std_p, 1e-12)
    ho = tr[n_fit:]
    if len(ho):
        z_true = (obs[ho] - mean_tr[n_fit:]) / np.maximum(std_tr[n_fit:], 1e-12)
        z_err = float(np.sqrt(np.mean((z[ho] - z_true) ** 2)))
    else:
        z_err = 0.0
    band = max(band, band_sigmas * z_err)
    mean_p[tr] = mean_tr
    std_p[tr] = std_tr
    z = (obs - mean_p) / np.maximum(std_p, 1e-12)
    p_pred = (0.5 * erfc(z / math.sqrt(2.0))).astype(np.float32)
    adj = bh_fdr_gpu(torch.from_numpy(p_pred).to(dev), alpha=fdr).cpu().numpy()
    sig = (adj < fdr) & (obs - mean_p > 0)
    if sig.any():
        z_thr = float(z[sig].min())
    else:
        z_thr = float(-np.log(fdr / max(P, 1)))  # no predicted edge: a high bar
    border = np.abs(z - z_thr) < band
    border[tr] = False
    rr = np.nonzero(border)[0]
    # 4. plain surrogates for the borderline pairs
    if len(rr):
        s_b, s2_b, cg_b = plain_stats_pairs(prep, pairs[rr], obs_d, n_surrogates,
                                            shuffle_method, block_length, seed, rng_device)
        mb = s_b / n_surrogates
        mean_p[rr] = mb
        std_p[rr] = np.sqrt(np.maximum(s2_b / n_surrogates - mb ** 2, 0.0))
    t_total = time.perf_counter() - t0
    # 5. decision with the reference arithmetic on the merged null
    mean_m = np.zeros((n, n), np.float32)
    std_m = np.zeros((n, n), np.float32)
    mean_m[pub[:, 0], pub[:, 1]] = mean_p
    std_m[pub[:, 0], pub[:, 1]] = std_p
    eff = (observed_te - mean_m).astype(np.float32)
    eps = np.float32(1e-12)
    zz = (observed_te - mean_m) / np.maximum(std_m, eps)
    pv = (0.5 * erfc(zz / np.sqrt(2.0))).astype(np.float32)
    np.fill_diagonal(pv, 1.0)
    adj_off = bh_fdr_gpu(torch.from_numpy(pv[mask]).to(dev), alpha=fdr).cpu().numpy()
    sigm = np.zeros((n, n), bool)
    sigm[mask] = adj_off < fdr
    sigm &= eff > 0
    edges = np.argwhere(sigm).astype(np.int64)
    grn = make_grn(np.asarray(variable_names), edges, eff[edges[:, 0], edges[:, 1]])
    vprint(f"[TENEX] learned null: trained on {n_train} pairs, re-ran {len(rr)} of {P}; "
           f"{len(grn)} edges")
    res = SurrogateTestResult(observed_te=observed_te, mean_surrogate_te=mean_m,
                              std_surrogate_te=std_m, effective_te=eff, p_values=pv, grn=grn,
                              n_surrogates=n_surrogates, shuffle_method=shuffle_method,
                              block_length=block_length if shuffle_method == "block" else None,
                              p_method="parametric", fdr=fdr)
    res.learned_info = dict(n_pairs=P, n_train=n_train, n_rerun=int(len(rr)), band=band,
                            z_err_holdout=z_err,
                            plain_fraction=(n_train + len(rr)) / max(P, 1), z_threshold=z_thr,
                            t_train_plain=t_train_plain, t_fit=t_fit, t_total=t_total,
                            losses=losses)
    return res
