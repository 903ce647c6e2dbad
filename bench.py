#!/usr/bin/env python
"""Benchmark: pairwise transfer entropy on one MI355X.

Default workload (the headline metric): 4,000 genes x 10,000 cells,
mixed-cardinality dense bins (per-gene levels: 80% in [2, 8], 18% in [9, 32],
2% in [33, 64]), tau = 1, all 15,996,000 ordered pairs.  One step = the
per-gene preparation (target states, S1, S2a) plus every pair's joint
histogram and TE, from bins already resident in HBM, into a device TE matrix.

Prints one JSON line (rank 0) with the throughput, the roofline of the
dominant kernel (te_wave_kernel) and the CPU baseline (the torch-CPU
restatement of the reference's scatter_add path on a bounded sample).

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--workload c2|c1|c3]
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


def _events():
    import torch
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def time_launches(fn, reps):
    """Average duration (s) of fn() over reps launches, HIP events on torch's
    current stream (the stream the library launches on)."""
    import torch
    e0, e1 = _events()
    fn()
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    e1.synchronize()
    return e0.elapsed_time(e1) / 1e3 / reps


def cpu_baseline_c2(bins_np, npv, n_sample, seed, gpu_te_internal):
    """scatter_add restatement (oracle/scatter_add_ref.py) on a seeded
    stratified sample of pairs; returns (dict, max |GPU - FP64 oracle|)."""
    import torch
    from oracle import oracle as O
    from oracle import scatter_add_ref as S
    rng = np.random.default_rng(seed)
    n = bins_np.shape[0]
    # stratify by histogram cost class b_t^2 b_s: small / medium / large
    tg = rng.integers(0, n, size=8 * n_sample)
    sr = rng.integers(0, n, size=8 * n_sample)
    keep = tg != sr
    tg, sr = tg[keep], sr[keep]
    cost = npv[tg].astype(np.int64) ** 2 * npv[sr]
    cls = np.digitize(cost, [2048, 16384])
    pick = []
    for c, frac in ((0, 0.6), (1, 0.3), (2, 0.1)):
        idx = np.nonzero(cls == c)[0][: int(n_sample * frac)]
        pick.append(idx)
    pick = np.concatenate(pick)
    pairs = np.stack([tg[pick], sr[pick]], axis=1).astype(np.int64)
    pairs = pairs[np.argsort(pairs[:, 0], kind="stable")]
    b_max = int(npv.max())
    threads = torch.get_num_threads()
    t_cpu, te_cpu = S.time_pairs(bins_np, b_max, pairs, tau=1, repeats=1, warmup=0)
    te64 = O.te_pairs(bins_np, npv, pairs, tau=1)
    gpu = gpu_te_internal[pairs[:, 0], pairs[:, 1]]
    err = float(np.abs(gpu.astype(np.float64) - te64).max())
    err_cpu = float(np.abs(te_cpu.astype(np.float64) - te64).max())
    return {
        "value": len(pairs) / t_cpu, "unit": "pairs/s", "cores": int(threads), "kind": "port",
        "sample": (f"{len(pairs)} pairs stratified by b_t^2*b_s (60% <=2048, 30% <=16384, "
                   f"10% larger), one scatter_add batch at global B={b_max}; "
                   f"{t_cpu:.2f} s; extrapolated pairs/s"),
        "max_abs_err_vs_fp64_oracle": err_cpu,
    }, err


def run_c2(args, torch, core, synthetic):
    from tenex_amd import _lib
    dev = core.require_gpu("cuda:0")
    n, T, tau = 4000, 10000, 1
    log(f"workload C2: {n} genes x {T} cells mixed-cardinality bins, seed {args.seed}")
    bins_np = synthetic.mixed_cardinality_bins(n, T, seed=args.seed)
    bins, npv, b_max, T = core.upload_bins(bins_np, dev)
    prep = core.prepare(bins, npv, T, tau)
    t0 = time.perf_counter()
    plan = core.get_plan(prep)
    plan_s = time.perf_counter() - t0
    info = plan.info
    n_pairs = n * (n - 1)
    assert info.n_pairs == n_pairs
    out = torch.zeros((n, n), dtype=torch.float32, device=dev)
    targs = core.te_args(prep, plan, out, _lib.TNX_OUT_MATRIX, ld_out=n)

    def step():
        core.gene_prep_into(prep)
        core.te_run(prep, plan, out, _lib.TNX_OUT_MATRIX, args=targs)

    log(f"plan: {info.n_wave_items} wave items, {info.n_block_items} block items, "
        f"{plan_s * 1e3:.1f} ms host")
    for i in range(args.warmup):
        step()
        torch.cuda.synchronize()
        log(f"warmup {i + 1}/{args.warmup} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    log(f"timed {args.steps} steps: {dt * 1e3:.2f} ms/step")
    # per-kernel durations
    wave_args = core.te_args(prep, plan, out, _lib.TNX_OUT_MATRIX, ld_out=n)
    wave_args.n_block_items = 0
    block_args = core.te_args(prep, plan, out, _lib.TNX_OUT_MATRIX, ld_out=n)
    block_args.n_wave_items = 0
    block_args.work = plan.work.data_ptr()  # block items sit after the wave items
    # (te_block_kernel indexes work + n_wave_items: keep the offset)
    block_args.n_wave_items = 0
    block_args.work = plan.work.data_ptr() + 16 * info.n_wave_items
    reps = max(2, args.steps)
    t_wave = time_launches(lambda: core.te_run(prep, plan, out, 1, args=wave_args), reps)
    t_block = time_launches(lambda: core.te_run(prep, plan, out, 1, args=block_args), reps)
    t_prep = time_launches(lambda: core.gene_prep_into(prep), reps)
    log(f"te_wave_kernel {t_wave * 1e3:.2f} ms, te_block_kernel {t_block * 1e3:.2f} ms, "
        f"gene_prep {t_prep * 1e3:.3f} ms")
    # wave-kernel pair count
    wh = plan.work_host[: info.n_wave_items]
    wave_pairs = int(np.sum(wh[:, 1].astype(np.int64) * (wh[:, 3] - wh[:, 2])))
    tg_np, sr_np = plan.tgt_host, plan.src_host
    L = T - tau
    # algorithmic HBM bytes of one te_wave_kernel launch: every bin row it
    # reads once (target states 2 B/sample, source 1 B/sample) + 4 B per TE
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
        "scaling": "weak", "vs_baseline": None, "dtype": "int32 counts, fp32 entropy, fp64 combine",
        "data": "synthetic (seeded mixed-cardinality bins)",
        "config": {"workload": "C2 mixed-cardinality TE matrix", "n_genes": n, "n_cells": T,
                   "tau": tau, "n_pairs": n_pairs, "b_max": int(b_max),
                   "levels": "80% [2,8], 18% [9,32], 2% [33,64]", "seed": args.seed},
        "roofline": {"bound": "hbm", "kernel": "te_wave_kernel", "achieved": hbm_achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": hbm_achieved / HBM_PEAK_GBS,
                     "traffic": None,
                     "note": "histogram path is LDS-atomic bound; see lds_roofline"},
        "lds_roofline": {"bound": "lds_atomic", "kernel": "te_wave_kernel",
                         "achieved": upd_rate / 1e9, "peak": lds_peak / 1e9,
                         "peak_random_keys": lds_rand / 1e9, "unit": "Gupdates/s",
                         "frac": upd_rate / lds_peak, "updates_per_launch": wave_pairs * L},
        "kernels_ms": {"te_wave_kernel": t_wave * 1e3, "te_block_kernel": t_block * 1e3,
                       "gene_prep_kernel": t_prep * 1e3},
        "wave_pairs": wave_pairs, "block_pairs": n_pairs - wave_pairs,
        "plan_host_ms": plan_s * 1e3,
    }
    if not args.no_cpu_baseline:
        log("cpu baseline (scatter_add restatement) on a bounded sample ...")
        cb, err = cpu_baseline_c2(bins_np.astype(np.int32), prep.npv_host, args.cpu_sample,
                                  args.seed, te_host)
        result["cpu_baseline"] = cb
        result["max_abs_err_vs_fp64_oracle_sample"] = err
    return result


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c2", choices=["c2"])
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--cpu-sample", type=int, default=2048)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build runs one MI355X per process and has no "
              "multi-GPU path (--gpus must be 1)", file=sys.stderr)
        sys.exit(2)
    import torch
    from tenex_amd import core, synthetic
    res = run_c2(args, torch, core, synthetic)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
