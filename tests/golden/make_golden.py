"""Write small golden fixtures from the reference implementation (CPU only).

Run once in a container that holds the reference source tree at
/root/reference; the outputs are committed under tests/golden/ and the tests
read only those .npz files (the reference itself never travels).

Fixtures
--------
te_c0_100x500.npz      engine.compute(accelerator="cpu") on the 100-gene x
                       500-cell zero-inflated log-normal workload (seed 0):
                       input data, bins, n_per_var, the float32 TE matrix of
                       the reference's scatter_add path.
surrogate_40x400.npz   NetWeaver(result).infer("surrogate_test") on the CPU loop
                       path (torch CPU generator), 40 genes x 400 cells, 20
                       block-shuffle surrogates, block length 25, seed 42,
                       FDR 0.05: observed, mean, std, effective TE, p-values,
                       edge list.
surrogate_120x1000.npz the same at 120 genes x 1,000 cells with 8 planted
                       lagged dependencies, 50 surrogates, block length 40,
                       seed 42, FDR 0.05 (more edges pinned than the 40-gene
                       fixture).
fdr_300x600.npz        the FDR method's arithmetic on the 300-gene x 600-cell
                       TE matrix of the reference's CPU engine (seed 5, 24
                       planted lagged dependencies of graded strength and 8
                       chains a -> b -> c with a weaker direct a -> c): the
                       reference's
                       fdr_threshold_gpu (z score, erfc p-values, BH at 0.05,
                       cut at FDR 0.01) evaluated on CPU tensors in the
                       public pair order of build_pairs, then its DPI _trim
                       on the CPU device: thresholded and trimmed edge lists
                       with their weights.
trace_80x700.npz       NetWeaver(result).infer("trace", device="cpu") on 80
                       genes x 700 cells (seed 21, 10 hubs driving two
                       targets each): the TE matrix, bins, OutTE, InTE, the
                       thresholded network and its GRN (100 surrogate
                       entries, significance 2.0, the reference's CPU
                       hash-entropy path).
multilag_50x500.npz    Pipeline(engine, fdr=0.05).infer("surrogate_test",
                       taus=[1, 2, 3]) on the CPU (50 genes x 500 cells, seed
                       33, dependencies planted at lags 1, 2 and 3; 20
                       surrogates, seed 42, default block length per lag):
                       the combined statistics and edges of lag_combine
                       'max', and the edges of 'adaptive_select'.
binarize_60x400.npz    engine.compute(accelerator="cpu", binarization=...) for
                       'median', 'quantile' (q 0.8), 'mean' and 'nonzero' on
                       60 genes x 400 cells (seed 44, a constant variable and
                       a tied one): bins, n_per_var and the TE matrix.
scrna_small.npz        load_scrna on the CSV / pseudotime / branch / sources
                       files of scrna_small/ (7 genes x 40 cells, tied
                       pseudotimes): the aligned branch-1 data with sources,
                       the binarised branch-2 data, and the .npy cache names.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"


def zero_inflated_lognormal(n, T, seed):
    rng = np.random.default_rng(seed)
    x = rng.lognormal(1.0, 1.0, size=(n, T)).astype(np.float32)
    x[rng.random((n, T)) < 0.3] = 0.0
    return x


def main(which):
    sys.path.insert(0, REF)
    import tenex as tnx  # the reference package

    if "te_c0" in which:
        write_te_c0(tnx)
    if "surrogate_40x400" in which:
        write_surrogate_40(tnx)
    if "surrogate_120x1000" in which:
        write_surrogate_120(tnx)
    if "fdr_300x600" in which:
        write_fdr_300(tnx)
    if "trace_80x700" in which:
        write_trace_80(tnx)
    if "multilag_50x500" in which:
        write_multilag_50(tnx)
    if "binarize_60x400" in which:
        write_binarize_60(tnx)
    if "scrna_small" in which:
        write_scrna_small(tnx)
    print("golden fixtures written:", ", ".join(which))


def write_te_c0(tnx):
    x = zero_inflated_lognormal(100, 500, 0)
    names = np.array([f"g{i}" for i in range(100)])
    res = tnx.TransferEntropyEngine(x, names).compute(accelerator="cpu")
    np.savez_compressed(os.path.join(HERE, "te_c0_100x500.npz"), data=x,
                        bins=np.asarray(res.bin_arrs)[:, :, 0].astype(np.int8),
                        n_per_var=np.asarray(res.n_per_var, np.int32),
                        matrix=np.asarray(res.matrix, np.float32), b_max=res.b_max,
                        kernel=str(res.kernel))


def write_surrogate_40(tnx):
    x2 = zero_inflated_lognormal(40, 400, 8)
    x2[7, 1:] = x2[3, :-1] + 0.1 * np.random.default_rng(1).random(399).astype(np.float32)
    names2 = np.array([f"g{i}" for i in range(40)])
    r2 = tnx.TransferEntropyEngine(x2, names2).compute(accelerator="cpu")
    st = tnx.NetWeaver(r2, fdr=0.05).infer(method="surrogate_test", device="cpu",
                                            n_surrogates=20, block_length=25, seed=42)
    np.savez_compressed(os.path.join(HERE, "surrogate_40x400.npz"), data=x2,
                        bins=np.asarray(r2.bin_arrs)[:, :, 0].astype(np.int8),
                        n_per_var=np.asarray(r2.n_per_var, np.int32),
                        observed=st.observed_te, mean=st.mean_surrogate_te,
                        std=st.std_surrogate_te, effective=st.effective_te, p=st.p_values,
                        edges=np.asarray(st.grn.pairs, np.int64), n_surrogates=20,
                        block_length=25, seed=42, fdr=0.05)


def write_surrogate_120(tnx):
    x3 = zero_inflated_lognormal(120, 1000, 13)
    rng = np.random.default_rng(3)
    for s_, t_ in rng.choice(120, (8, 2), replace=False):
        x3[t_, 1:] = x3[s_, :-1] * (0.5 + rng.random(999).astype(np.float32))
    names3 = np.array([f"g{i}" for i in range(120)])
    r3 = tnx.TransferEntropyEngine(x3, names3).compute(accelerator="cpu")
    st3 = tnx.NetWeaver(r3, fdr=0.05).infer(method="surrogate_test", device="cpu",
                                             n_surrogates=50, block_length=40, seed=42)
    np.savez_compressed(os.path.join(HERE, "surrogate_120x1000.npz"), data=x3,
                        bins=np.asarray(r3.bin_arrs)[:, :, 0].astype(np.int8),
                        n_per_var=np.asarray(r3.n_per_var, np.int32),
                        observed=st3.observed_te, mean=st3.mean_surrogate_te,
                        std=st3.std_surrogate_te, effective=st3.effective_te, p=st3.p_values,
                        edges=np.asarray(st3.grn.pairs, np.int64), n_surrogates=50,
                        block_length=40, seed=42, fdr=0.05)


def write_fdr_300(tnx):
    import torch
    from tenex.inference import build_pairs, fdr_threshold_gpu
    from tenex.inference.fdr import _trim
    x4 = zero_inflated_lognormal(300, 600, 5)
    rng = np.random.default_rng(5)
    g = rng.choice(300, 24 * 2 + 8 * 3, replace=False)
    # 24 single lagged dependencies of graded strength (some near the cut)
    for i, w in enumerate(np.linspace(0.15, 1.0, 24)):
        s_, t_ = g[2 * i], g[2 * i + 1]
        x4[t_, 1:] = (1 - w) * x4[t_, 1:] + w * x4[s_, :-1] * (0.5 + rng.random(599))
    # 8 chains a -> b -> c with a weaker direct a -> c: DPI removes a -> c
    for i in range(8):
        a, b, c = g[48 + 3 * i: 51 + 3 * i]
        x4[b, 1:] = x4[a, :-1] * (0.5 + rng.random(599))
        x4[c, 1:] = x4[b, :-1] * (0.5 + rng.random(599)) + 0.4 * x4[a, :-1]
    x4 = x4.astype(np.float32)
    names4 = np.array([f"g{i}" for i in range(300)])
    r4 = tnx.TransferEntropyEngine(x4, names4).compute(accelerator="cpu")
    M = np.asarray(r4.matrix, np.float32)
    pairs = build_pairs(names4)
    pairs_t = torch.from_numpy(pairs).to(torch.long)
    te_t = torch.from_numpy(M)[pairs_t[:, 0], pairs_t[:, 1]]
    te_cut, pairs_cut = fdr_threshold_gpu(te_t, pairs_t, 0.01)
    te_cut, pairs_cut = te_cut.numpy(), pairs_cut.numpy()
    trimmed = _trim(M, names4, pairs_cut, te_cut, device=torch.device("cpu"), batch_size=512,
                    trim_threshold=0.0)
    np.savez_compressed(os.path.join(HERE, "fdr_300x600.npz"), matrix=M,
                        pairs_cut=np.asarray(pairs_cut, np.int64),
                        te_cut=np.asarray(te_cut, np.float32),
                        trimmed_pairs=np.asarray(trimmed.pairs, np.int64),
                        trimmed_te=np.asarray(trimmed.te, np.float32), fdr=0.01)


def write_trace_80(tnx):
    x5 = zero_inflated_lognormal(80, 700, 21)
    rng = np.random.default_rng(21)
    g = rng.choice(80, 30, replace=False)
    for i in range(10):                       # hubs: one source drives two targets
        s_, t1, t2 = g[3 * i: 3 * i + 3]
        x5[t1, 1:] = x5[s_, :-1] * (0.5 + rng.random(699))
        x5[t2, 1:] = 0.5 * x5[t2, 1:] + x5[s_, :-1] * (0.5 + rng.random(699))
    x5 = x5.astype(np.float32)
    names5 = np.array([f"g{i}" for i in range(80)])
    r5 = tnx.TransferEntropyEngine(x5, names5).compute(accelerator="cpu")
    tr = tnx.NetWeaver(r5).infer(method="trace", device="cpu", n_surrogates=100,
                                 significance=2.0)
    net = np.array([[int(s_), j] for j in range(80) for s_ in tr.network[j]],
                   np.int64).reshape(-1, 2)
    np.savez_compressed(os.path.join(HERE, "trace_80x700.npz"), data=x5,
                        bins=np.asarray(r5.bin_arrs)[:, :, 0].astype(np.int8),
                        matrix=np.asarray(r5.matrix, np.float32),
                        outte=np.asarray(tr.outte, np.float32),
                        inte=np.asarray(tr.inte, np.float32), network=net,
                        grn_pairs=np.asarray(tr.grn.pairs, np.int64),
                        grn_te=np.asarray(tr.grn.te, np.float32), n_surrogates=100,
                        significance=2.0)


def write_multilag_50(tnx):
    x6 = zero_inflated_lognormal(50, 500, 33)
    rng = np.random.default_rng(33)
    g = rng.choice(50, 12, replace=False)
    for i, lag in enumerate((1, 2, 3, 1, 2, 3)):
        s_, t_ = g[2 * i], g[2 * i + 1]
        x6[t_, lag:] = x6[s_, :-lag] * (0.5 + rng.random(500 - lag))
    x6 = x6.astype(np.float32)
    names6 = np.array([f"g{i}" for i in range(50)])
    out = {}
    for combine in ("max", "adaptive_select"):
        eng = tnx.TransferEntropyEngine(x6, names6)
        pipe = tnx.Pipeline(eng, fdr=0.05).configure(accelerator="cpu")
        r = pipe.infer("surrogate_test", taus=[1, 2, 3], lag_combine=combine, device="cpu",
                       n_surrogates=20, seed=42)
        out[combine] = r
    r = out["max"]
    ra = out["adaptive_select"]
    np.savez_compressed(os.path.join(HERE, "multilag_50x500.npz"), data=x6,
                        observed=r.observed_te, mean=r.mean_surrogate_te,
                        std=r.std_surrogate_te, effective=r.effective_te, p=r.p_values,
                        edges=np.asarray(r.grn.pairs, np.int64),
                        edges_adaptive=np.asarray(ra.grn.pairs, np.int64),
                        effective_adaptive=ra.effective_te, taus=np.array([1, 2, 3]),
                        n_surrogates=20, seed=42, fdr=0.05)


def write_binarize_60(tnx):
    x7 = zero_inflated_lognormal(60, 400, 44)
    x7[5] = 1.0                                  # a constant variable
    x7[9, ::2] = x7[9, 1::2]                     # ties around the median
    names7 = np.array([f"g{i}" for i in range(60)])
    out = {"data": x7}
    for tag, method, kw in (("median", "median", None), ("q80", "quantile", {"q": 0.8}),
                            ("mean", "mean", None), ("nonzero", "nonzero", None)):
        r = tnx.TransferEntropyEngine(x7, names7).compute(
            accelerator="cpu", binarization=method, binarization_kwargs=kw)
        out[f"bins_{tag}"] = np.asarray(r.bin_arrs)[:, :, 0].astype(np.int8)
        out[f"npv_{tag}"] = np.asarray(r.n_per_var, np.int32)
        out[f"matrix_{tag}"] = np.asarray(r.matrix, np.float32)
    np.savez_compressed(os.path.join(HERE, "binarize_60x400.npz"), **out)


def write_scrna_small(tnx):
    """A tiny CSV / pseudotime / branch set under tests/golden/scrna_small
    and the reference's load_scrna output for it (run on a scratch copy, so
    the reference's .npy caches do not land in the fixture folder)."""
    import shutil
    import tempfile
    d = os.path.join(HERE, "scrna_small")
    os.makedirs(d, exist_ok=True)
    rng = np.random.default_rng(55)
    n_genes, n_cells = 7, 40
    expr = np.round(rng.lognormal(0, 1, (n_cells, n_genes)) * (rng.random((n_cells, n_genes)) < 0.6),
                    3)
    genes = [f"G{i}" for i in range(n_genes)]
    with open(os.path.join(d, "expression.csv"), "w") as f:
        f.write("cell," + ",".join(genes) + "\n")
        for c in range(n_cells):
            f.write(f"c{c}," + ",".join(f"{v:.3f}" for v in expr[c]) + "\n")
    pt = np.round(rng.random(n_cells), 2)          # ties on purpose
    br = rng.integers(1, 3, n_cells)
    np.savetxt(os.path.join(d, "pseudotime.txt"), pt, fmt="%.2f")
    np.savetxt(os.path.join(d, "branch.txt"), br, fmt="%d")
    with open(os.path.join(d, "sources.txt"), "w") as f:
        f.write("G1\nG4\n")
    from tenex.io import load_scrna
    tmp = tempfile.mkdtemp()
    try:
        for fn in ("expression.csv", "pseudotime.txt", "branch.txt", "sources.txt"):
            shutil.copy(os.path.join(d, fn), tmp)
        r = load_scrna(os.path.join(tmp, "expression.csv"), os.path.join(tmp, "pseudotime.txt"),
                       os.path.join(tmp, "branch.txt"), branch_id=1,
                       sources=os.path.join(tmp, "sources.txt"))
        rb = load_scrna(os.path.join(tmp, "expression.csv"), os.path.join(tmp, "pseudotime.txt"),
                        os.path.join(tmp, "branch.txt"), branch_id=2, make_binary=True)
        caches = sorted(f for f in os.listdir(tmp) if f.endswith(".npy"))
    finally:
        shutil.rmtree(tmp)
    np.savez_compressed(os.path.join(HERE, "scrna_small.npz"), data=r.data,
                        gene_names=np.asarray(r.gene_names).astype("U8"),
                        sources=np.asarray(r.sources).astype("U8"), data_b2=rb.data,
                        caches=np.array(caches, dtype="U64"))


ALL = ("te_c0", "surrogate_40x400", "surrogate_120x1000", "fdr_300x600", "trace_80x700",
       "multilag_50x500", "binarize_60x400", "scrna_small")

if __name__ == "__main__":
    # python make_golden.py [fixture ...]   (default: all of ALL)
    main(sys.argv[1:] or ALL)
