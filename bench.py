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
  trace     TRACE OutTE / InTE, 1,000 x 3,000 with planted hubs
  multilag  the c3 surrogate test at lags 1-3, every lag in each launch
  sortpairs 400 x 1,000 with 2-120 levels: large pairs on occupied keys

Prints one JSON line with the throughput, the roofline of the dominant
kernel and the CPU baseline (the torch-CPU restatement of the reference's
scatter_add path, timed on a bounded sample of the same workload).

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--workload c2]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
MFMA_INT8_PEAK_TOPS = 5000.0
MFMA_FP4_PEAK_TOPS = 10100.0   # dense FP4 MFMA (published MI355X figure)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def onehot_kp(L):
    return -(-int(L) // 128) * 128


def onehot_roofline(dev, oh, n_pairs, thr, L, t_mfma):
    """The one-hot MFMA launch against the FP4 matrix-core peaks (published,
    and measured from registers in this run)."""
    if oh is None:
        return {"pairs": 0, "threshold": thr}
    from tenex_amd import core
    ops = 2.0 * oh.cells * onehot_kp(L)
    measured = core.mfma_rate(dev, 0)
    return {"pairs": n_pairs, "tiles": oh.n_tiles, "threshold": thr, "kernel": "oh_mfma_kernel",
            "bound": "mfma", "fp4_mfma_ops": ops, "ms": t_mfma * 1e3,
            "achieved": ops / t_mfma / 1e12, "unit": "TOP/s", "peak": MFMA_FP4_PEAK_TOPS,
            "frac": ops / t_mfma / 1e12 / MFMA_FP4_PEAK_TOPS,
            "peak_measured_registers": measured / 1e12,
            "frac_of_measured": ops / t_mfma / measured,
            "traffic": measured_traffic("onehot_traffic_bytes_per_launch"),
            "traffic_source": TRAFFIC_FILE}


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


TRAFFIC_FILE = "profiles/te_persist_c2_4000x10000_traffic.md"


def measured_traffic(key: str = "traffic_bytes_per_launch"):
    """HBM bytes per launch of a C2 kernel (the persistent launch, or with
    key onehot_traffic_bytes_per_launch the one-hot MFMA launch), from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (copy-calibrated);
    None if absent."""
    import re
    path = os.path.join(HERE, TRAFFIC_FILE)
    if not os.path.exists(path):
        return None
    m = re.search(r"(?m)^" + key + r":\s*(\d+)", open(path).read())
    return int(m.group(1)) if m else None


def run_traffic_probe(args):
    """One C2 TE matrix as the headline step computes it (the one-hot pack
    and MFMA launches, then the persistent launch), then a 256 MiB device
    copy: the program the FETCH_SIZE / WRITE_SIZE counter passes run (the
    copy calibrates the counters)."""
    import torch
    from tenex_amd import _lib, core, onehot, synthetic
    dev = core.require_gpu("cuda:0")
    n, T = 4000, 10000
    if args.continuous:   # the continuous-data variant: discretised on the GPU
        x = torch.from_numpy(synthetic.zero_inflated_lognormal(n, T, seed=args.seed)).to(dev)
        raw, _ = core.fsbw_l_gpu(x, 0.5)
        bins, npv, b_max = core.remap_gpu(raw)
    else:
        bins, npv, b_max, T = core.upload_bins(
            synthetic.mixed_cardinality_bins(n, T, seed=args.seed), dev)
    prep = core.prepare(bins, npv, T, 1)
    oh = core.get_onehot_plan(prep, thr=args.onehot_thr) if args.onehot_thr > 0 else None
    plan = core.get_plan(prep, src_skip=None if oh is None else oh.src_skip)
    out = torch.zeros((n, n), dtype=torch.float32, device=dev)
    targs = core.te_args(prep, plan, out, _lib.TNX_OUT_MATRIX, ld_out=n)
    if oh is not None:
        onehot.run(prep, oh, targs, _lib.stream_handle(dev))
    core.te_run(prep, plan, out, _lib.TNX_OUT_MATRIX, args=targs)
    src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    log("traffic probe done: one C2 matrix, one 256 MiB copy")


def lds_peaks(dev):
    """Measured chip-wide LDS atomic rates (updates/s): conflict-free and
    random keys in 2048 cells (tnx_bench_lds_atomics)."""
    from tenex_amd import core
    return core.lds_atomic_rate(dev, mode=0), core.lds_atomic_rate(dev, mode=1)


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
    from tenex_amd import onehot
    t0 = time.perf_counter()
    oh = (core.get_onehot_plan(prep, thr=args.onehot_thr)
          if core.onehot_enabled() and args.onehot_thr > 0 else None)
    plan = core.get_plan(prep, cap_wave=cw, cap_block=cb,
                         src_skip=None if oh is None else oh.src_skip)
    persistent = core.persistent_ok(prep, plan)
    targs = core.te_args(prep, plan, None, _lib.TNX_OUT_MATRIX, ld_out=n)  # builds the queues
    plan_s = time.perf_counter() - t0
    info = plan.info
    n_pairs = n * (n - 1)
    n_oh = 0 if oh is None else oh.n_pairs
    assert info.n_pairs + n_oh == n_pairs, (info.n_pairs, n_oh, n_pairs)
    out = torch.zeros((n, n), dtype=torch.float32, device=dev)
    targs = core.te_args(prep, plan, out, _lib.TNX_OUT_MATRIX, ld_out=n)
    stream = _lib.stream_handle(dev)

    def run_oh():
        if oh is not None:
            onehot.run(prep, oh, targs, stream)

    def step():
        core.gene_prep_into(prep)
        run_oh()
        core.te_run(prep, plan, out, _lib.TNX_OUT_MATRIX, args=targs)

    if not args.no_graph:  # the same launches captured once in a HIP graph, replayed per step
        graph = core.capture_te_step(prep, plan, out, _lib.TNX_OUT_MATRIX, ld_out=n, oh=oh)
        step = graph.replay  # noqa: F811

    log(f"plan: {n_oh} one-hot pairs in {0 if oh is None else oh.n_tiles} tiles, "
        f"{info.n_wave_items} wave items, {info.n_block_items} block items, "
        f"{'one persistent launch' if persistent else 'two launches'}, {plan_s * 1e3:.1f} ms host")
    dt = timed_steps(step, args.steps, args.warmup)
    reps = max(2, args.steps)
    # the launches alone (HIP events on the stream they run on), and the
    # earlier two-launch kernels on the same plan for comparison
    t_pers = time_launches(lambda: core.te_run(prep, plan, out, 1, args=targs), reps)
    t_oh = time_launches(run_oh, reps) if oh is not None else 0.0
    t_oh_mfma = (time_launches(lambda: _lib.call("tnx_onehot_te", ctypes.byref(oh.args),
                                                 ctypes.byref(targs), stream), reps)
                 if oh is not None else 0.0)
    t_te = t_pers + t_oh
    t_prep = time_launches(lambda: core.gene_prep_into(prep), reps)
    a2 = core.te_args(prep, plan, out, _lib.TNX_OUT_MATRIX, ld_out=n, persistent=False)
    t_two = time_launches(lambda: core.te_run(prep, plan, out, 1, args=a2), reps)
    run_oh()
    core.te_run(prep, plan, out, 1, args=targs)
    log(f"TE launches {t_te * 1e3:.2f} ms (one-hot pack + MFMA {t_oh * 1e3:.2f} ms, "
        f"{'persistent' if persistent else 'two launches'} {t_pers * 1e3:.2f} ms), "
        f"two-launch kernels {t_two * 1e3:.2f} ms, gene_prep {t_prep * 1e3:.3f} ms")
    L = T - tau
    upd = n_pairs * L            # LDS histogram updates of one TE launch
    lds_peak, lds_rand = lds_peaks(dev)
    # algorithmic HBM bytes of one TE launch: each bin row read once (target
    # states 2 B / sample, source 1 B / sample) + 4 B per TE written
    alg_bytes = n * L * 3 + 4 * n_pairs
    te_host = out.cpu().numpy()
    kname = (f"te_persist_kernel<{-(-L // 1024)}, {str(prep.hot[0] > 0).lower()}>" if persistent
             else "te_wave_reg_kernel + te_block_kernel")
    if oh is not None:   # the step's TE launches: the one-hot split, then the histogram launch
        kname = "oh_pack_kernel + oh_mfma_kernel, then " + kname
    result = {
        "metric": "TE gene-pairs/sec at 4k genes x 10k cells (mixed-cardinality, all ordered pairs)",
        "value": n_pairs / dt, "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic (seeded mixed-cardinality bins)",
        "config": {"workload": "C2 mixed-cardinality TE matrix, 4000 genes x 10000 cells",
                   "n_genes": n, "n_cells": T, "tau": tau, "n_pairs": n_pairs,
                   "b_max": int(b_max), "levels": "80% [2,8], 18% [9,32], 2% [33,64]",
                   "seed": args.seed, "persistent_launch": persistent,
                   "arith": "int32 counts, fp32 compensated entropy terms, fp64 final combine"},
        "roofline": {"bound": "lds_atomic", "kernel": kname,
                     "achieved": upd / t_te / 1e9, "peak": lds_peak / 1e9, "unit": "Gupdates/s",
                     "frac": upd / t_te / lds_peak, "updates_per_launch": upd,
                     "peak_random_keys": lds_rand / 1e9,
                     "step_frac": upd / dt / lds_peak,
                     "peak_source": "tnx_bench_lds_atomics mode 0 (conflict-free ds_add_u32, "
                                    "all CUs), measured in this run",
                     "traffic": measured_traffic(), "traffic_source": TRAFFIC_FILE},
        "hbm_roofline": {"bound": "hbm", "achieved": alg_bytes / t_te / 1e9,
                         "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": alg_bytes / t_te / 1e9 / HBM_PEAK_GBS,
                         "algorithmic_bytes": alg_bytes, "traffic": measured_traffic()},
        "kernels_ms": {"te_launches": t_te * 1e3, "persistent_launch": t_pers * 1e3,
                       "onehot_pack_and_mfma": t_oh * 1e3,
                       "te_two_launch_kernels": t_two * 1e3, "gene_prep_kernel": t_prep * 1e3},
        "onehot": onehot_roofline(dev, oh, n_oh, args.onehot_thr, prep.L, t_oh_mfma),
        "plan_host_ms": plan_s * 1e3,
    }
    if tuned is not None:
        result["config"]["autotune"] = tuned
    result["config"]["hip_graph"] = not args.no_graph
    if not args.no_cpu_baseline:
        log("cpu baseline (scatter_add restatement) on a bounded sample ...")
        pairs = sample_pairs(prep.npv_host, args.cpu_sample, args.seed)
        cb, err = cpu_baseline_pairs(bins_np.astype(np.int32), prep.npv_host, pairs, te_host)
        cb["sample"] = "stratified by b_t^2*b_s (60% <=2048, 30% <=16384, 10% larger); " + cb["sample"]
        result["cpu_baseline"] = cb
        result["max_abs_err_vs_fp64_oracle_sample"] = err
    if not args.no_continuous:
        result["continuous_variant"] = run_c2_continuous(args, dev, lds_peak)
    return result


def run_c2_continuous(args, dev, lds_peak):
    """The C2 shape from expression values: 4000 x 10000 zero-inflated
    log-normal floats (30% zeros, mu = 1, sigma = 1) resident in HBM, FSBW-L
    discretisation and dense remap on the GPU, then the TE matrix (one-hot
    split where it applies, persistent launch for the rest, one HIP graph)."""
    import torch
    from tenex_amd import _lib, core, synthetic
    n, T, tau = 4000, 10000, 1
    x_np = synthetic.zero_inflated_lognormal(n, T, seed=args.seed)
    x = torch.from_numpy(x_np).to(dev)

    def disc():
        raw, _ = core.fsbw_l_gpu(x, 0.5)
        return core.remap_gpu(raw)   # reads one status word back

    bins, npv, b_max = disc()
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    reps = 3
    for _ in range(reps):
        bins, npv, b_max = disc()
    torch.cuda.synchronize(dev)
    t_disc = (time.perf_counter() - t0) / reps
    prep = core.prepare(bins, npv, T, tau)
    levels = np.bincount(prep.npv_host, minlength=int(b_max) + 1)
    plan = core.get_te_plan(prep)
    oh = plan.cache.get("onehot")
    out = torch.zeros((n, n), dtype=torch.float32, device=dev)
    graph = core.capture_te_step(prep, plan, out, _lib.TNX_OUT_MATRIX, ld_out=n, oh=oh)
    dt = timed_steps(graph.replay, max(2, min(args.steps, 5)), 1)
    n_pairs = n * (n - 1)
    upd = n_pairs * (T - tau)
    res = {"workload": "C2 shape from continuous data: 4000 x 10000 zero-inflated log-normal "
                       "(30% zeros, mu=1, sigma=1), FSBW-L kp=0.5 + dense remap on the GPU",
           "levels_min_median_max": [int(prep.npv_host.min()), int(np.median(prep.npv_host)),
                                     int(prep.npv_host.max())],
           "level_histogram": {int(b): int(c) for b, c in enumerate(levels) if c},
           "discretize_ms": t_disc * 1e3, "ms_per_step": dt * 1e3, "pairs_per_s": n_pairs / dt,
           "end_to_end_ms": (t_disc + dt) * 1e3,
           "onehot_pairs": 0 if oh is None else oh.n_pairs,
           "wave_items": int(plan.info.n_wave_items), "block_items": int(plan.info.n_block_items),
           "lds_frac": upd / dt / lds_peak}
    log(f"continuous variant: levels {res['levels_min_median_max']}, discretize "
        f"{t_disc * 1e3:.2f} ms, TE step {dt * 1e3:.2f} ms ({res['lds_frac']:.2f} of the LDS peak)")
    if not args.no_cpu_baseline:
        # checked in the CPU-baseline leg: bins against the oracle's
        # discretisation, TE of a sample of pairs against the FP64 oracle
        from oracle import oracle as O
        d, _, npv_o, _ = O.discretize(x_np)
        bins_h = bins[:, :T].cpu().numpy().astype(np.int32)
        res["bins_equal_oracle"] = bool(np.array_equal(bins_h, d)
                                        and np.array_equal(prep.npv_host, npv_o))
        pairs = sample_pairs(prep.npv_host, 2048, args.seed + 1, stratify=False)
        te_o = O.te_pairs(d, npv_o, pairs)
        te_g = out.cpu().numpy()[pairs[:, 0], pairs[:, 1]]   # internal [target, source]
        res["max_abs_err_vs_fp64_oracle_sample"] = float(np.abs(te_g - te_o).max())
        res["oracle_sample_pairs"] = len(pairs)
    return res


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

    def launch(precision=_lib.TNX_B2_AUTO):
        _lib.call("tnx_te_gemm_b2", prep.bins.data_ptr(), n, T, prep.ld, 1, prep.S1.data_ptr(),
                  prep.S2a.data_ptr(), prep.xlogx.data_ptr(), ws.data_ptr(), ws_bytes,
                  out.data_ptr(), n, None, precision, s)

    dt = timed_steps(launch, args.steps, args.warmup)   # the default: FP4 operands (L < 2^24)
    t_k = time_launches(launch, max(3, args.steps))
    te_fp4 = out.cpu().numpy()
    t_8 = time_launches(lambda: launch(_lib.TNX_B2_INT8), max(3, args.steps))
    dmax = float(np.abs(te_fp4 - out.cpu().numpy()).max())
    launch()
    L = T - 1
    ops = 2.0 * 3 * n * n * L
    tops = ops / t_k / 1e12
    result = {"metric": "TE gene-pairs/sec, binary 2000 genes x 5000 cells (GEMM-B2)",
              "value": n * (n - 1) / dt, "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
              "warmup": args.warmup, "ms_per_step": dt * 1e3, "higher_is_better": True,
              "scaling": "weak", "vs_baseline": None,
              "dtype": "fp4 (e2m1) 0/1 operands -> fp32 sums (exact below 2^24)",
              "data": "synthetic Bernoulli with 200 planted lagged pairs (seed 7)",
              "config": {"workload": "C1 GEMM-B2 2000x5000", "n_pairs": n * (n - 1)},
              "roofline": {"bound": "mfma", "kernel": "b2_mfma_kernel<FP4> (+ b2_pack_kernel<FP4>)",
                           "achieved": tops, "peak": MFMA_FP4_PEAK_TOPS, "unit": "TOP/s",
                           "frac": tops / MFMA_FP4_PEAK_TOPS, "ops_per_launch": ops,
                           "frac_of_int8_peak": tops / MFMA_INT8_PEAK_TOPS, "traffic": None},
              "int8_kernel": {"ms": t_8 * 1e3, "tops": ops / t_8 / 1e12,
                              "frac_of_int8_peak": ops / t_8 / 1e12 / MFMA_INT8_PEAK_TOPS,
                              "max_abs_diff_vs_fp4": dmax}}
    fp4_reg = core.mfma_rate(dev, 0)
    result["roofline"]["peak_measured_registers"] = fp4_reg / 1e12
    result["roofline"]["frac_of_measured"] = ops / t_k / fp4_reg
    if not args.no_cpu_baseline:
        d = prep.bins[:, :T].cpu().numpy().astype(np.int32)
        pairs = sample_pairs(prep.npv_host, args.cpu_sample, args.seed, stratify=False)
        cb, err = cpu_baseline_pairs(d, prep.npv_host, pairs, out.cpu().numpy())
        result["cpu_baseline"] = cb
        result["max_abs_err_vs_fp64_oracle_sample"] = err
    return result


def surrogate_cpu_baseline(d, npv, n_pairs_total, n_tests, sample_pairs_n, seed):
    """scatter_add restatement timed on one batch of pairs (a surrogate's TE
    costs what an observed TE costs), extrapolated to n_pairs_total x
    n_tests TE evaluations."""
    from oracle import scatter_add_ref as S
    rng = np.random.default_rng(seed)
    n = d.shape[0]
    tg = rng.integers(0, n, sample_pairs_n * 2)
    sr = rng.integers(0, n, sample_pairs_n * 2)
    keep = tg != sr
    pairs = np.stack([tg[keep], sr[keep]], 1)[:sample_pairs_n].astype(np.int64)
    t_cpu, _ = S.time_pairs(d, int(npv.max()), pairs, tau=1, repeats=1, warmup=0)
    rate = len(pairs) / t_cpu
    aff, threads = host_cores()
    return {"value": n_pairs_total * n_tests / rate, "unit": "s", "cores": int(threads),
            "kind": "port",
            "sample": (f"{len(pairs)} random pairs through the scatter_add restatement "
                       f"({t_cpu:.2f} s on {threads} threads, {aff} cores in affinity), "
                       f"extrapolated to {n_pairs_total} pairs x {n_tests} TE evaluations "
                       f"(observed + surrogates)")}


def surrogate_oracle_err(d, npv, res, n, T, n_surr, bl, seed, dev, n_check, rng_seed):
    """max |mean surrogate TE (GPU) - FP64 oracle| over n_check random tested
    pairs, on the block orders the test drew (same generator and seeds)."""
    from oracle import oracle as O
    from tenex_amd import core
    perm, eff, nb = core.block_perm_device(n, T, "block", bl, 0, n_surr, seed, dev, "cpu")
    rng = np.random.default_rng(rng_seed)
    allp = O.pairfree_pairs(n)
    pairs = allp[rng.choice(len(allp), n_check, replace=False)]       # [target, source]
    obs = res.observed_te[pairs[:, 1], pairs[:, 0]]                     # public [src, tgt]
    s, s2, cg = O.surrogate_stats(d, npv, perm.numpy(), eff, pairs, obs)
    gpu = res.mean_surrogate_te[pairs[:, 1], pairs[:, 0]].astype(np.float64)
    return float(np.abs(gpu - s / n_surr).max())


def run_c3(args):
    import torch
    import tenex_amd as tnx
    from tenex_amd import core, synthetic
    dev = core.require_gpu("cuda:0")
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
    lds_peak, lds_rand = lds_peaks(dev)
    result = {"metric": "surrogate-test seconds, 1000 genes x 3000 cells, 100 block shuffles",
              "value": dt, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
              "ms_per_step": dt * 1e3, "higher_is_better": False, "scaling": "weak",
              "vs_baseline": None, "dtype": "int32",
              "data": "synthetic zero-inflated log-normal (seed 3)",
              "config": {"workload": "C3 plain surrogate test", "b_max": int(res.b_max),
                         "n_surrogates": 100, "block_length": 50, "edges": len(st.grn)},
              "roofline": {"bound": "lds_atomic", "kernel": "whole surrogate test",
                           "achieved": upd / dt / 1e9, "peak": lds_peak / 1e9,
                           "unit": "Gupdates/s", "frac": upd / dt / lds_peak,
                           "updates_per_test": upd, "peak_random_keys": lds_rand / 1e9,
                           "traffic": None}}
    d = res.bin_arrs[:, :, 0].astype(np.int32)
    if not args.no_cpu_baseline:   # the oracle runs in this leg only
        result["max_abs_err_mean_surrogate_vs_fp64_oracle"] = surrogate_oracle_err(
            d, res.n_per_var, st, n, T, 100, 50, 42, dev, 200, args.seed)
        log("cpu baseline (scatter_add restatement, per-surrogate TE) on a bounded sample ...")
        result["cpu_baseline"] = surrogate_cpu_baseline(d, res.n_per_var, n * (n - 1), 101,
                                                        2048, args.seed)
    return result


def run_c4(args):
    import torch
    import tenex_amd as tnx
    from tenex_amd import core, synthetic
    dev = core.require_gpu("cuda:0")
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

    def jaccard(p, q):
        a = set(map(tuple, p.grn.pairs.tolist()))
        b = set(map(tuple, q.grn.pairs.tolist()))
        return len(a & b) / max(1, len(a | b)), len(a), len(b)

    t_plain, plain = timed(lambda: nw.infer("surrogate_test", **kw))
    t_learn, learned = timed(lambda: nw.infer("surrogate_test", learned="mlp", **kw))
    # the configuration's literal rule: re-run pairs with |z - z_thr| < 1.0
    t_lit, literal = timed(lambda: nw.infer("surrogate_test", learned="mlp",
                                            learned_kwargs=dict(band_sigmas=0.0), **kw))
    jac, ea, eb = jaccard(plain, learned)
    jac_lit, _, el = jaccard(plain, literal)
    info = {k: v for k, v in learned.learned_info.items() if k not in ("losses", "rerun_pairs")}
    L = T - 1
    n_pairs = 200 * 2000
    upd = n_pairs * 100 * L
    lds_peak, _ = lds_peaks(dev)
    out = {"metric": "surrogate-test seconds, learned vs plain (6000x3000, 200 TF x 2000 targets)",
           "value": t_learn, "unit": "s", "n_gpus": 1, "steps": args.steps,
           "warmup": max(1, args.warmup),
           "ms_per_step": t_learn * 1e3, "higher_is_better": False, "scaling": "weak",
           "vs_baseline": None, "dtype": "int32 counts; bf16 weights / fp32 accumulate MLP",
           "data": "synthetic zero-inflated log-normal with 100 planted TF->target edges",
           "config": {"workload": "C4 learned-null surrogate test", "n_surrogates": 100,
                      "block_length": 50, "band": "max(1.0, 3.5 x held-out RMS z error)"},
           "roofline": {"bound": "lds_atomic", "kernel": "plain surrogate test (the learned "
                        "test measures a fraction of these histograms)",
                        "achieved": upd / t_plain / 1e9, "peak": lds_peak / 1e9,
                        "unit": "Gupdates/s", "frac": upd / t_plain / lds_peak,
                        "updates_per_test": upd, "traffic": None},
           "plain_s": t_plain, "learned_s": t_learn, "speedup": t_plain / t_learn,
           "edge_jaccard": jac, "edges_plain": ea, "edges_learned": eb,
           "literal_band": {"band": 1.0, "learned_s": t_lit, "speedup": t_plain / t_lit,
                            "edge_jaccard": jac_lit, "edges": el,
                            "n_rerun": int(literal.learned_info["n_rerun"])},
           "learned_info": info}
    d = res.bin_arrs[:, :, 0].astype(np.int32)
    if not args.no_cpu_baseline:
        log("cpu baseline (scatter_add restatement, per-surrogate TE) on a bounded sample ...")
        out["cpu_baseline"] = surrogate_cpu_baseline(d, res.n_per_var, n_pairs, 101, 2048,
                                                     args.seed)
    return out


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


LDS_PEAK_BYTES = 256 * 128 * 2.4e9   # 128 B per clock per CU, 256 CUs, 2.4 GHz


def sort_lds_bytes(L: int) -> tuple[int, int]:
    """(P, LDS bytes per sort) of the in-LDS bitonic network of sort_lds.hpp
    for L keys: stages whose partner is in another wave store and load 16 B
    per key, lane-shuffle stages move 8 B per key (two 4-byte permutes),
    register stages none, plus the 8-byte store of the sorted keys."""
    P = max(1024, 1 << (max(1, L) - 1).bit_length())
    n_lds = n_shfl = 0
    k = 2
    while k <= P:
        j = k >> 1
        while j >= 1:
            if j >= 16 * 64:
                n_lds += 1
            elif j >= 16:
                n_shfl += 1
            j >>= 1
        k <<= 1
    return P, P * (16 * n_lds + 8 * n_shfl + 8)


def run_trace(args):
    """TRACE OutTE / InTE on 1,000 genes x 3,000 cells (the surrogate
    workload's data with 40 hubs driving 5 targets each)."""
    import torch
    import tenex_amd as tnx
    from tenex_amd import core, synthetic
    from tenex_amd.inference.trace import _outgoing, build_causal_network, trace_scores
    dev = core.require_gpu("cuda:0")
    n, T = 1000, 3000
    x = synthetic.zero_inflated_lognormal(n, T, seed=3)
    rng = np.random.default_rng(3)
    hubs = rng.choice(n, 240, replace=False)
    for h in range(40):
        s = hubs[6 * h]
        for t in hubs[6 * h + 1: 6 * h + 6]:
            x[t, 1:] = 0.5 * x[t, 1:] + x[s, :-1] * (0.5 + rng.random(T - 1).astype(np.float32))
    names = np.array([f"g{i}" for i in range(n)])
    res = tnx.TransferEntropyEngine(x, names).compute(accelerator="gpu")
    nw = tnx.NetWeaver(res)

    def step():
        return nw.infer("trace")

    dt = timed_steps(step, args.steps, args.warmup)
    tr = step()
    network = build_causal_network(res.matrix, 100, 2.0, 42, dev)
    outgoing, src, _ = _outgoing(network, n)
    bins, _, _, _ = core.upload_bins(res.bin_arrs, dev)
    t_k = time_launches(lambda: trace_scores(bins, T, 1, network, outgoing), max(3, args.steps))
    L = T - 1
    P, per_sort = sort_lds_bytes(L)
    jobs = sum(1 for i in range(n) if len(outgoing[i])) + sum(1 for i in range(n)
                                                              if len(network[i]))
    lds_bytes = 4 * jobs * per_sort
    d = res.bin_arrs[:, :, 0].astype(np.int32)
    result = {
        "metric": "TRACE seconds (network + OutTE + InTE), 1000 genes x 3000 cells",
        "value": dt, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt * 1e3, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "int64", "data": "synthetic zero-inflated log-normal "
        "(seed 3), 40 hubs x 5 planted targets",
        "config": {"workload": "TRACE OutTE / InTE", "n_genes": n, "n_cells": T,
                   "edges": int(len(src)), "entropy_jobs": 4 * jobs, "sort_size": P},
        "roofline": {"bound": "lds", "kernel": "trace_entropy_kernel",
                     "achieved": lds_bytes / t_k / 1e9, "peak": LDS_PEAK_BYTES / 1e9,
                     "unit": "GB/s", "frac": lds_bytes / t_k / LDS_PEAK_BYTES,
                     "traffic": None, "lds_bytes_per_launch": lds_bytes,
                     "definition": "LDS bytes of the bitonic network (16 B per key per "
                                   "cross-wave stage, 8 B per key per lane-shuffle stage)"},
        "kernels_ms": {"trace_scores": t_k * 1e3},
    }
    if not args.no_cpu_baseline:   # the oracle runs in this leg only
        from oracle import oracle as O
        ids = np.nonzero(tr.outte > 0)[0][:20]
        result["max_abs_err_outte_vs_fp64_oracle_sample"] = max(
            abs(float(tr.outte[i]) - O.trace_te(d, int(i), sorted(outgoing[i]), 0))
            for i in ids) if len(ids) else None
        ids = [i for i in range(n) if len(outgoing[i])][:12]
        t0 = time.perf_counter()
        for i in ids:
            O.trace_te(d, i, sorted(outgoing[i]), 0)
            O.trace_te(d, i, network[i], 1) if len(network[i]) else None
        tc = (time.perf_counter() - t0) * jobs / (2 * len(ids))
        result["cpu_baseline"] = {"value": tc, "unit": "s", "cores": 1, "kind": "port",
                                  "sample": f"the reference's CPU hash + np.unique entropies "
                                            f"(compute_outte_cpu / compute_inte_cpu) for "
                                            f"{len(ids)} variables, extrapolated to {jobs} jobs"}
    return result


def run_multilag(args):
    """The C3 surrogate test at lags 1, 2 and 3, every lag in each launch."""
    import tenex_amd as tnx
    from tenex_amd import core, synthetic
    dev = core.require_gpu("cuda:0")
    n, T = 1000, 3000
    x = synthetic.zero_inflated_lognormal(n, T, seed=3)
    names = np.array([f"g{i}" for i in range(n)])
    eng = tnx.TransferEntropyEngine(x, names)
    pipe = tnx.Pipeline(eng, fdr=0.05)
    taus = [1, 2, 3]

    def step():
        return pipe.infer("surrogate_test", taus=taus, n_surrogates=100, block_length=50)

    dt = timed_steps(step, args.steps, args.warmup)
    r = step()
    upd = n * (n - 1) * 101 * sum(T - t for t in taus)
    lds_peak, lds_rand = lds_peaks(dev)
    result = {
        "metric": "multi-lag surrogate-test seconds, 1000 genes x 3000 cells, taus 1-3, "
                  "100 block shuffles",
        "value": dt, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt * 1e3, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic zero-inflated log-normal (seed 3)",
        "config": {"workload": "multi-lag surrogate test (lag_combine max)", "taus": taus,
                   "n_surrogates": 100, "block_length": 50, "edges": len(r.grn)},
        "roofline": {"bound": "lds_atomic", "kernel": "whole multi-lag test",
                     "achieved": upd / dt / 1e9, "peak": lds_peak / 1e9, "unit": "Gupdates/s",
                     "frac": upd / dt / lds_peak, "updates_per_test": upd,
                     "peak_random_keys": lds_rand / 1e9, "traffic": None},
    }
    if not args.no_cpu_baseline:
        res = eng.compute(accelerator="gpu")
        d = res.bin_arrs[:, :, 0].astype(np.int32)
        cb = surrogate_cpu_baseline(d, res.n_per_var, n * (n - 1), 101 * len(taus), 2048,
                                    args.seed)
        result["cpu_baseline"] = cb
    return result


def run_sortpairs(args):
    """Occupied-key TE: 400 variables x 1,000 cells with 2..120 levels, the
    TE matrix through the engine's matrix path (pairs above the sort
    threshold on te_sort_pair_kernel, the rest on the histogram launch)."""
    import torch
    from tenex_amd import core
    dev = core.require_gpu("cuda:0")
    n, T = 400, 1000
    rng = np.random.default_rng(args.seed)
    levels = rng.integers(2, 121, n)
    d = np.stack([rng.integers(0, b, T) for b in levels]).astype(np.int32)
    for g, b in enumerate(levels):
        d[g, rng.choice(T, b, replace=False)] = np.arange(b)
    b8, b16, npv, _, _ = core.upload_bins_any(d, dev)
    prep = core.prepare(b8, npv, T, 1)
    cap = core.sort_threshold(prep.L)
    pairs = np.stack(np.meshgrid(np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 2)
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    npvh = prep.npv_host.astype(np.int64)
    big = pairs[npvh[pairs[:, 0]] ** 2 * npvh[pairs[:, 1]] > cap]

    def step():
        return core.te_matrix_internal(prep)

    dt = timed_steps(step, args.steps, args.warmup)
    m = step().cpu().numpy()
    out = torch.zeros((n, n), dtype=torch.float32, device=dev)
    t_k = time_launches(lambda: core.te_sort_pairs(prep, big, out, n), max(3, args.steps))
    plan = core.get_plan(prep)
    dense = torch.zeros((n, n), dtype=torch.float32, device=dev)
    t_d = time_launches(lambda: core.te_run(prep, plan, dense, 1, ld_out=n), max(3, args.steps))
    P, per_sort = sort_lds_bytes(T - 1)
    lds_bytes = len(big) * per_sort
    smp = big[rng.choice(len(big), 200, replace=False)]
    result = {
        "metric": "TE matrix seconds, 400 variables x 1000 cells, 2-120 levels",
        "value": dt, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt * 1e3, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "int64", "data": "synthetic uniform levels (seeded)",
        "config": {"workload": "occupied-key routing of large pairs", "n_vars": n,
                   "n_cells": T, "sort_threshold_cells": cap, "sorted_pairs": int(len(big)),
                   "all_pairs": int(len(pairs)), "sort_size": P},
        "roofline": {"bound": "lds", "kernel": "te_sort_pair_kernel",
                     "achieved": lds_bytes / t_k / 1e9, "peak": LDS_PEAK_BYTES / 1e9,
                     "unit": "GB/s", "frac": lds_bytes / t_k / LDS_PEAK_BYTES,
                     "traffic": None, "lds_bytes_per_launch": lds_bytes},
        "kernels_ms": {"te_sort_pairs (large pairs)": t_k * 1e3,
                       "histogram launch, all pairs dense": t_d * 1e3},
    }
    if not args.no_cpu_baseline:   # the oracle runs in this leg only
        from oracle import oracle as O
        t0 = time.perf_counter()
        te64 = O.te_pairs_sparse(d, smp)
        tc = time.perf_counter() - t0
        result["max_abs_err_vs_fp64_oracle_sample"] = float(
            np.abs(m[smp[:, 0], smp[:, 1]].astype(np.float64) - te64).max())
        result["cpu_baseline"] = {"value": len(smp) / tc, "unit": "pairs/s", "cores": 1,
                                  "kind": "port", "sample": "200 large pairs, numpy "
                                  "occupied-state entropies (oracle te_pairs_sparse)"}
    return result


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c2",
                    choices=["c0", "c1", "c2", "c3", "c4", "dpi", "trace", "multilag",
                             "sortpairs"])
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--onehot-thr", type=int, default=None,
                    help="largest (b_x^2 - 1) * b_y a pair may have to be counted on INT8 MFMA "
                         "(C2; 0: every pair in the histogram kernels)")
    ap.add_argument("--cpu-sample", type=int, default=20480)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--continuous", action="store_true",
                    help="with --traffic-probe: the continuous-data variant of C2")
    ap.add_argument("--no-continuous", action="store_true",
                    help="c2: skip the continuous-data variant (discretize + TE matrix)")
    ap.add_argument("--no-graph", action="store_true",
                    help="c2: launch the step eagerly instead of replaying it from a HIP graph")
    ap.add_argument("--autotune", action=argparse.BooleanOptionalAction, default=True,
                    help="c2: time the launch candidates on a sample before the timed "
                         "region (core.autotune; it keeps the defaults unless a candidate "
                         "is 2%% faster)")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="run one C2 matrix and a 256 MiB copy (for rocprofv3 --pmc), no JSON")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build runs one MI355X per process and has no "
              "multi-GPU path (--gpus must be 1)", file=sys.stderr)
        sys.exit(2)
    if args.onehot_thr is None:
        from tenex_amd import onehot
        args.onehot_thr = onehot.default_threshold(10000 - 1)
    if args.traffic_probe:
        run_traffic_probe(args)
        return
    fn = {"c0": run_c0, "c1": run_c1, "c2": run_c2, "c3": run_c3, "c4": run_c4,
          "dpi": run_dpi, "trace": run_trace, "multilag": run_multilag,
          "sortpairs": run_sortpairs}[args.workload]
    res = fn(args)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
