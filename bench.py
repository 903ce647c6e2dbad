#!/usr/bin/env python
"""Benchmark: pairwise transfer entropy on one MI355X.

Default workload (the headline metric, --workload c2): 4,000 genes x 10,000
cells, mixed-cardinality dense bins (per-gene levels: 80% in [2, 8], 18% in
[9, 32], 2% in [33, 64]), tau = 1, all 15,996,000 ordered pairs.  One step =
the per-gene preparation (target states, S1, S2a) plus every pair's joint
histogram and TE, from bins already resident in HBM, into a device TE matrix.

Other workloads (their own JSON line, same fields):
  c0  100 x 500 zero-inflated log-normal, engine.compute end to end
  c1  2,000 x 5,000 binary (200 planted lagged pairs): GEMM-B2 on INT8 MFMA
  c3  1,000 x 3,000 plain surrogate test, 100 block shuffles (block 50)
  c4  6,000 x 3,000, 200 TF x 2,000 target pairs: learned vs plain null
  dpi 4,000-gene FDR network (1% of ordered pairs): DPI trimming

Prints one JSON line with the throughput, the roofline of the dominant
kernel and the CPU baseline (the torch-CPU restatement of the reference's
scatter_add path, timed on a bounded sample of the same workload).

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--workload c2]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
MFMA_INT8_PEAK_TOPS = 5000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def time_launches(fn, reps):
    """Average duration (s) of fn() over reps launches, HIP events on torch's
    current stream (the stream the library launches on)."""
    import torch
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    fn()
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    e1.synchronize()
    return e0.elapsed_time(e1) / 1e3 / reps


def timed_steps(step, steps, warmup):
    import torch
    for i in range(warmup):
        step()
        torch.cuda.synchronize()
        log(f"warmup {i + 1}/{warmup} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    log(f"timed {steps} steps: {dt * 1e3:.2f} ms/step")
    return dt


def host_cores():
    import torch
    return len(os.sched_getaffinity(0)), torch.get_num_threads()


def sample_pairs(npv, n_sample, seed, stratify=True):
    rng = np.random.default_rng(seed)
    n = len(npv)
    tg = rng.integers(0, n, size=8 * n_sample)
    sr = rng.integers(0, n, size=8 * n_sample)
    keep = tg != sr
    tg, sr = tg[keep], sr[keep]
    if stratify:
        cost = npv[tg].astype(np.int64) ** 2 * npv[sr]
        cls = np.digitize(cost, [2048, 16384])
        pick = np.concatenate([np.nonzero(cls == c)[0][: int(n_sample * f)]
                               for c, f in ((0, 0.6), (1, 0.3), (2, 0.1))])
    else:
        pick = np.arange(n_sample)
    pairs = np.stack([tg[pick], sr[pick]], axis=1).astype(np.int64)
    return pairs[np.argsort(pairs[:, 0], kind="stable")]


def cpu_baseline_pairs(bins_np, npv, pairs, gpu_te_internal=None):
    """scatter_add restatement (oracle/scatter_add_ref.py) timed on a pair
    sample; returns (dict, max |GPU - FP64 oracle| on the sample)."""
    from oracle import oracle as O
    from oracle import scatter_add_ref as S
    b_max = int(npv.max())
    t_cpu, te_cpu = S.time_pairs(bins_np, b_max, pairs, tau=1, repeats=1, warmup=0)
    te64 = O.te_pairs(bins_np, npv, pairs, tau=1)
    err = None
    if gpu_te_internal is not None:
        gpu = gpu_te_internal[pairs[:, 0], pairs[:, 1]]
        err = float(np.abs(gpu.astype(np.float64) - te64).max())
    aff, threads = host_cores()
    return {
        "value": len(pairs) / t_cpu, "unit": "pairs/s", "cores": int(threads), "kind": "port",
        "sample": (f"{len(pairs)} pairs, scatter_add batches of 2048 at global B={b_max}; "
                   f"{t_cpu:.2f} s on {threads} threads ({aff} cores in affinity); "
                   f"extrapolated pairs/s"),
        "max_abs_err_vs_fp64_oracle": float(np.abs(te_cpu.astype(np.float64) - te64).max()),
    }, err


TRAFFIC_FILE = "profiles/te_wave_c2_4000x10000_traffic.md"


def measured_traffic():
    """HBM bytes per launch of the C2 wave kernel, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (copy-calibrated); None if
    absent."""
    import re
    path = os.path.join(HERE, TRAFFIC_FILE)
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*(\d+)", open(path).read())
    return int(m.group(1)) if m else None


def run_traffic_probe(args):
    """One C2 TE matrix, then a 256 MiB device copy: the program the
    FETCH_SIZE / WRITE_SIZE counter passes run (the copy calibrates the
    counters)."""
    import torch
    from tenex_amd import _lib, core, synthetic
    dev = core.require_gpu("cuda:0")
    n, T = 4000, 10000
    bins, npv, b_max, T = core.upload_bins(synthetic.mixed_cardinality_bins(n, T, seed=args.seed),
                                           dev)
    prep = core.prepare(bins, npv, T, 1)
    plan = core.get_plan(prep)
    out = torch.zeros((n, n), dtype=torch.float32, device=dev)
    core.te_run(prep, plan, out, _lib.TNX_OUT_MATRIX, ld_out=n)
    src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    log("traffic probe done: one C2 matrix, one 256 MiB copy")


def run_c2(args):
    import torch
    from tenex_amd import _lib, core, synthetic
    dev = core.require_gpu("cuda:0")
    n, T, tau = 4000, 10000, 1
    log(f"workload C2: {n} genes x {T} cells mixed-cardinality bins, seed {args.seed}")
    bins_np = synthetic.mixed_cardinality_bins(n, T, seed=args.seed)
    bins, npv, b_max, T = core.upload_bins(bins_np, dev)
    prep = core.prepare(bins, npv, T, tau)
    tuned = None
    cw = cb = 0
    if args.autotune:
        tuned = core.autotune(prep)
        cw, cb = tuned["cap_wave"], tuned["cap_block"]
        log(f"autotune: {tuned}")
    t0 = time.perf_counter()
    plan = core.get_plan(prep, cap_wave=cw, cap_block=cb)
    plan_s = time.perf_counter() - t0
    info = plan.info
    n_pairs = n * (n - 1)
    assert info.n_pairs == n_pairs
    out = torch.zeros((n, n), dtype=torch.float32, device=dev)
    targs = core.te_args(prep, plan, out, _lib.TNX_OUT_MATRIX, ld_out=n)

    def step():
        core.gene_prep_into(prep)
        core.te_run(prep, plan, out, _lib.TNX_OUT_MATRIX, args=targs)

    if args.graph:  # the same launches captured once in a HIP graph, replayed per step
        graph = core.capture_te_step(prep, plan, out, _lib.TNX_OUT_MATRIX, ld_out=n)
        step = graph.replay  # noqa: F811

    log(f"plan: {info.n_wave_items} wave items, {info.n_block_items} block items, "
        f"{plan_s * 1e3:.1f} ms host")
    dt = timed_steps(step, args.steps, args.warmup)
    # per-kernel durations: the wave items alone, the block items alone
    wave_args = core.te_args(prep, plan, out, _lib.TNX_OUT_MATRIX, ld_out=n)
    wave_args.n_block_items = 0
    block_args = core.te_args(prep, plan, out, _lib.TNX_OUT_MATRIX, ld_out=n)
    block_args.n_wave_items = 0
    block_args.work = plan.work.data_ptr() + 16 * info.n_wave_items  # block items follow
    reps = max(2, args.steps)
    t_wave = time_launches(lambda: core.te_run(prep, plan, out, 1, args=wave_args), reps)
    t_block = time_launches(lambda: core.te_run(prep, plan, out, 1, args=block_args), reps)
    t_prep = time_launches(lambda: core.gene_prep_into(prep), reps)
    log(f"te_wave kernel {t_wave * 1e3:.2f} ms, te_block kernel {t_block * 1e3:.2f} ms, "
        f"gene_prep {t_prep * 1e3:.3f} ms")
    wh = plan.work_host[: info.n_wave_items]
    # pair slots of the wave items (at most n of them are target == source)
    wave_pairs = int(np.sum(wh[:, 1].astype(np.int64) * (wh[:, 3] - wh[:, 2])))
    L = T - tau
    # algorithmic HBM bytes of one wave-kernel launch: each bin row read once
    # (target states 2 B / sample, source 1 B / sample) + 4 B per TE written
    alg_bytes = n * L * 3 + 4 * wave_pairs
    hbm_achieved = alg_bytes / t_wave / 1e9
    lds_peak = core.lds_atomic_rate(dev, mode=0)
    lds_rand = core.lds_atomic_rate(dev, mode=1)
    upd_rate = wave_pairs * L / t_wave
    te_host = out.cpu().numpy()
    result = {
        "metric": "TE gene-pairs/sec at 4k genes x 10k cells (mixed-cardinality, all ordered pairs)",
        "value": n_pairs / dt, "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic (seeded mixed-cardinality bins)",
        "config": {"workload": "C2 mixed-cardinality TE matrix, 4000 genes x 10000 cells",
                   "n_genes": n, "n_cells": T, "tau": tau, "n_pairs": n_pairs,
                   "b_max": int(b_max), "levels": "80% [2,8], 18% [9,32], 2% [33,64]",
                   "seed": args.seed,
                   "arith": "int32 counts, fp32 compensated entropy terms, fp64 final combine"},
        "roofline": {"bound": "hbm", "kernel": "te_wave_reg_kernel", "achieved": hbm_achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": hbm_achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(),
                     "traffic_source": TRAFFIC_FILE,
                     "note": "histogram path is LDS-atomic bound; see lds_roofline"},
        "lds_roofline": {"bound": "lds_atomic", "kernel": "te_wave_reg_kernel",
                         "achieved": upd_rate / 1e9, "peak": lds_peak / 1e9,
                         "peak_random_keys": lds_rand / 1e9, "unit": "Gupdates/s",
                         "frac": upd_rate / lds_peak, "updates_per_launch": wave_pairs * L,
                         "step_frac": n_pairs * L / dt / lds_peak},
        "kernels_ms": {"te_wave_kernel": t_wave * 1e3, "te_block_kernel": t_block * 1e3,
                       "gene_prep_kernel": t_prep * 1e3},
        "wave_pairs": wave_pairs, "block_pairs": n_pairs - wave_pairs,
        "plan_host_ms": plan_s * 1e3,
    }
    if tuned is not None:
        result["config"]["autotune"] = tuned
    result["config"]["hip_graph"] = bool(args.graph)
    if not args.no_cpu_baseline:
        log("cpu baseline (scatter_add restatement) on a bounded sample ...")
        pairs = sample_pairs(prep.npv_host, args.cpu_sample, args.seed)
        cb, err = cpu_baseline_pairs(bins_np.astype(np.int32), prep.npv_host, pairs, te_host)
        cb["sample"] = "stratified by b_t^2*b_s (60% <=2048, 30% <=16384, 10% larger); " + cb["sample"]
        result["cpu_baseline"] = cb
        result["max_abs_err_vs_fp64_oracle_sample"] = err
    return result


def run_c0(args):
    import torch
    import tenex_amd as tnx
    from tenex_amd import core, synthetic
    core.require_gpu("cuda:0")
    x = synthetic.zero_inflated_lognormal(100, 500, seed=0)
    names = np.array([f"g{i}" for i in range(100)])

    def step():
        eng = tnx.TransferEntropyEngine(x, names)
        return eng.compute(accelerator="gpu")

    dt = timed_steps(step, args.steps, args.warmup)
    res = step()
    result = {"metric": "TE gene-pairs/sec, 100 genes x 500 cells end to end (engine.compute)",
              "value": 9900 / dt, "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
              "warmup": args.warmup, "ms_per_step": dt * 1e3, "higher_is_better": True,
              "scaling": "weak", "vs_baseline": None, "dtype": "int32",
              "data": "synthetic zero-inflated log-normal (seed 0)",
              "config": {"workload": "C0 plain TE matrix 100x500", "kernel": res.kernel,
                         "b_max": int(res.b_max)}}
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        d = res.bin_arrs[:, :, 0].astype(np.int32)
        pairs = O.build_pairs_engine(100)
        cb, _ = cpu_baseline_pairs(d, res.n_per_var, pairs)
        cb["sample"] = "all 9,900 pairs; " + cb["sample"].replace("extrapolated pairs/s", "measured")
        result["cpu_baseline"] = cb
        M = O.te_matrix(d, res.n_per_var)
        result["max_abs_err_vs_fp64_oracle"] = float(np.abs(res.matrix - M).max())
    return result


def run_c1(args):
    import torch
    from tenex_amd import core, synthetic
    from tenex_amd.kernels.gemm_b2 import te_gemm_b2_device
    dev = core.require_gpu("cuda:0")
    n, T = 2000, 5000
    x, planted = synthetic.bernoulli_planted(n, T, n_planted=200, seed=7)
    raw, nb = core.fsbw_l_gpu(torch.from_numpy(x).to(dev))
    bins, npv, b_max = core.remap_gpu(raw)
    prep = core.prepare(bins, npv, T, 1)
    assert b_max == 2
    from tenex_amd import _lib
    ws_bytes = int(_lib.load().tnx_gemm_b2_workspace(n, T, 1))
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    out = torch.empty((n, n), dtype=torch.float32, device=dev)
    s = torch.cuda.current_stream().cuda_stream

    def launch():
        _lib.call("tnx_te_gemm_b2", prep.bins.data_ptr(), n, T, prep.ld, 1, prep.S1.data_ptr(),
                  prep.S2a.data_ptr(), prep.xlogx.data_ptr(), ws.data_ptr(), ws_bytes,
                  out.data_ptr(), n, None, s)

    dt = timed_steps(launch, args.steps, args.warmup)
    t_k = time_launches(launch, max(3, args.steps))
    L = T - 1
    ops = 2.0 * 3 * n * n * L
    tops = ops / t_k / 1e12
    result = {"metric": "TE gene-pairs/sec, binary 2000 genes x 5000 cells (GEMM-B2)",
              "value": n * (n - 1) / dt, "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
              "warmup": args.warmup, "ms_per_step": dt * 1e3, "higher_is_better": True,
              "scaling": "weak", "vs_baseline": None, "dtype": "int8 x int8 -> int32",
              "data": "synthetic Bernoulli with 200 planted lagged pairs (seed 7)",
              "config": {"workload": "C1 GEMM-B2 2000x5000", "n_pairs": n * (n - 1)},
              "roofline": {"bound": "mfma", "kernel": "b2_mfma_kernel (+ b2_pack_kernel)",
                           "achieved": tops, "peak": MFMA_INT8_PEAK_TOPS, "unit": "TOP/s",
                           "frac": tops / MFMA_INT8_PEAK_TOPS, "traffic": None}}
    if not args.no_cpu_baseline:
        d = prep.bins[:, :T].cpu().numpy().astype(np.int32)
        pairs = sample_pairs(prep.npv_host, args.cpu_sample, args.seed, stratify=False)
        cb, err = cpu_baseline_pairs(d, prep.npv_host, pairs, out.cpu().numpy())
        result["cpu_baseline"] = cb
        result["max_abs_err_vs_fp64_oracle_sample"] = err
    return result


def run_c3(args):
    import torch
    import tenex_amd as tnx
    from tenex_amd import synthetic
    n, T = 1000, 3000
    x = synthetic.zero_inflated_lognormal(n, T, seed=3)
    names = np.array([f"g{i}" for i in range(n)])
    res = tnx.TransferEntropyEngine(x, names).compute(accelerator="gpu")
    nw = tnx.NetWeaver(res, fdr=0.05)

    def step():
        return nw.infer("surrogate_test", n_surrogates=100, block_length=50)

    dt = timed_steps(step, args.steps, args.warmup)
    st = step()
    L = T - 1
    upd = n * (n - 1) * 100 * L
    result = {"metric": "surrogate-test seconds, 1000 genes x 3000 cells, 100 block shuffles",
              "value": dt, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
              "ms_per_step": dt * 1e3, "higher_is_better": False, "scaling": "weak",
              "vs_baseline": None, "dtype": "int32",
              "data": "synthetic zero-inflated log-normal (seed 3)",
              "config": {"workload": "C3 plain surrogate test", "b_max": int(res.b_max),
                         "n_surrogates": 100, "block_length": 50, "edges": len(st.grn)},
              "lds_roofline": {"bound": "lds_atomic", "achieved": upd / dt / 1e9,
                               "unit": "Gupdates/s"}}
    return result


def run_c4(args):
    import torch
    import tenex_amd as tnx
    from tenex_amd import synthetic
    n, T = 6000, 3000
    x = synthetic.zero_inflated_lognormal(n, T, seed=4)
    rng = np.random.default_rng(4)
    names = np.array([f"g{i}" for i in range(n)])
    tf = rng.choice(n, 200, replace=False)
    rest = np.setdiff1d(np.arange(n), tf)
    tg = rng.choice(rest, 2000, replace=False)
    for s, t in zip(tf[:100], tg[:100]):       # planted regulation
        x[t, 1:] = x[s, :-1] * (0.5 + rng.random(T - 1).astype(np.float32))
    res = tnx.TransferEntropyEngine(x, names, sources=names[tf]).compute(accelerator="gpu")
    nw = tnx.NetWeaver(res, fdr=0.05)
    kw = dict(n_surrogates=100, block_length=50, targets=names[tg], pairs_scope="sources")

    def timed(fn):
        for _ in range(max(1, args.warmup)):
            out = fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            out = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / args.steps, out

    t_plain, plain = timed(lambda: nw.infer("surrogate_test", **kw))
    t_learn, learned = timed(lambda: nw.infer("surrogate_test", learned="mlp", **kw))
    a = set(map(tuple, plain.grn.pairs.tolist()))
    b = set(map(tuple, learned.grn.pairs.tolist()))
    jac = len(a & b) / max(1, len(a | b))
    info = {k: v for k, v in learned.learned_info.items() if k != "losses"}
    return {"metric": "surrogate-test seconds, learned vs plain (6000x3000, 200 TF x 2000 targets)",
            "value": t_learn, "unit": "s", "n_gpus": 1, "steps": args.steps,
            "warmup": max(1, args.warmup),
            "ms_per_step": t_learn * 1e3, "higher_is_better": False, "scaling": "weak",
            "vs_baseline": None, "dtype": "int32 counts; bf16 weights / fp32 accumulate MLP",
            "data": "synthetic zero-inflated log-normal with 100 planted TF->target edges",
            "config": {"workload": "C4 learned-null surrogate test", "n_surrogates": 100,
                       "block_length": 50},
            "plain_s": t_plain, "learned_s": t_learn, "speedup": t_plain / t_learn,
            "edge_jaccard": jac, "edges_plain": len(a), "edges_learned": len(b),
            "learned_info": info}


# fp32 VALU issue: 256 CUs x 64 lanes x 2.4 GHz = 39.3e12 lane-instructions/s
# (157.3 TFLOP/s counts a packed FMA as four flops).  One max-min update
# (acc = max(acc, min(a, b))) takes 1.5 instructions: v_min plus half a v_max3.
VALU_LANE_OPS = 256 * 64 * 2.4e9
MAXMIN_PEAK = VALU_LANE_OPS / 1.5


def run_dpi(args):
    """DPI trimming of a 4,000-gene network holding 1% of the ordered pairs
    (uniform weights): every row is a mediator, so the max-min product is
    4,000 x 4,000 x 4,000."""
    import torch
    from tenex_amd import core
    from tenex_amd.inference.fdr import _maxmin_keep, trim_dpi
    dev = core.require_gpu("cuda:0")
    n = 4000
    rng = np.random.default_rng(args.seed)
    A = np.where(rng.random((n, n)) < 0.01, rng.random((n, n)), 0).astype(np.float32)
    np.fill_diagonal(A, 0)
    p = np.argwhere(A != 0)
    names = np.array([f"g{i}" for i in range(n)])
    log(f"workload dpi: {n} genes, {len(p)} edges")

    def step():
        return trim_dpi(A, names, p, A[p[:, 0], p[:, 1]], dev)

    dt = timed_steps(step, args.steps, args.warmup)
    kept = step()
    Ad = torch.from_numpy(A).to(dev)
    out_probe = _maxmin_keep(Ad, Ad, Ad, 0.0, dev)  # noqa: F841 (warm)
    t_k = time_launches(lambda: _maxmin_keep(Ad, Ad, Ad, 0.0, dev), max(3, args.steps))
    upd = float(n) ** 3    # max-min updates per launch (every row a mediator)
    result = {
        "metric": "DPI trim seconds, 4000-gene FDR network (1% of pairs)",
        "value": dt, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt * 1e3, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic random network (seeded)",
        "config": {"workload": "DPI trim of the FDR network", "n_genes": n,
                   "n_edges": int(len(p)), "kept": int(len(kept)), "seed": args.seed},
        "roofline": {"bound": "valu", "kernel": "dpi_maxmin_kernel",
                     "achieved": upd / t_k / 1e12, "peak": MAXMIN_PEAK / 1e12,
                     "unit": "T max-min updates/s", "frac": upd / t_k / MAXMIN_PEAK,
                     "traffic": None, "updates_per_launch": upd},
        "kernels_ms": {"dpi_maxmin_kernel": t_k * 1e3},
    }
    if not args.no_cpu_baseline:
        rows = 200
        Ac = torch.from_numpy(A)
        t0 = time.perf_counter()
        _maxmin_keep(Ac[:rows].contiguous(), Ac, Ac[:rows].contiguous(), 0.0,
                     torch.device("cpu"))
        tc = (time.perf_counter() - t0) * n / rows
        result["cpu_baseline"] = {"value": tc, "unit": "s", "cores": torch.get_num_threads(),
                                  "kind": "port",
                                  "sample": f"max-min product of the first {rows} rows on the "
                                            f"torch CPU path, extrapolated x{n // rows}"}
    return result


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c2", choices=["c0", "c1", "c2", "c3", "c4", "dpi"])
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--cpu-sample", type=int, default=20480)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--graph", action="store_true",
                    help="c2: time replays of the step captured in a HIP graph")
    ap.add_argument("--autotune", action="store_true",
                    help="c2: time the launch candidates on a sample first (core.autotune)")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="run one C2 matrix and a 256 MiB copy (for rocprofv3 --pmc), no JSON")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build runs one MI355X per process and has no "
              "multi-GPU path (--gpus must be 1)", file=sys.stderr)
        sys.exit(2)
    if args.traffic_probe:
        run_traffic_probe(args)
        return
    fn = {"c0": run_c0, "c1": run_c1, "c2": run_c2, "c3": run_c3, "c4": run_c4,
          "dpi": run_dpi}[args.workload]
    res = fn(args)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
