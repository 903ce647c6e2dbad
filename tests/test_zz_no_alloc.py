"""tnx_te_run_noalloc (the persistent launch with its queue take counters in
the library's device array) against tnx_te_run with a caller-provided
counter buffer: identical outputs, bit for bit, on two calls in a row, the
second on a smaller input, and both within 1e-6 of the FP64 oracle."""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tenex_amd import _lib, core, synthetic

pytestmark = pytest.mark.gpu
TOL = 1e-6


def _both(dev, n, T, seed, cap_wave):
    b = synthetic.mixed_cardinality_bins(n, T, seed=seed)
    bins, npv, b_max, T = core.upload_bins(b, dev)
    prep = core.prepare(bins, npv, T, 1)
    plan = core.get_plan(prep, cap_wave=cap_wave)
    assert core.persistent_ok(prep, plan)
    ref = torch.full((n, n), -1.0, dtype=torch.float32, device=dev)
    args = core.te_args(prep, plan, ref, _lib.TNX_OUT_MATRIX, ld_out=n, persistent=True)
    core.te_run(prep, plan, ref, _lib.TNX_OUT_MATRIX, args=args)
    got = torch.full((n, n), -1.0, dtype=torch.float32, device=dev)  # diagonal untouched
    a2 = core.te_args(prep, plan, got, _lib.TNX_OUT_MATRIX, ld_out=n, persistent=True)
    a2.queue_ctr = None
    core.call("tnx_te_run_noalloc", C.byref(a2), core._stream(dev))
    torch.cuda.synchronize(dev)
    return b, prep, plan, ref.cpu().numpy(), got.cpu().numpy()


def test_noalloc_matches_te_run_twice(dev):
    for n, T, seed, cap in ((150, 3000, 31, 1536), (60, 1200, 5, 0)):
        b, prep, plan, ref, got = _both(dev, n, T, seed, cap)
        assert plan.info.n_block_items > 0 and plan.info.n_wave_items > 0
        assert np.array_equal(got, ref), (n, T)
        pairs = O.pairfree_pairs(n)
        te64 = O.te_pairs(b.astype(np.int32), prep.npv_host, pairs)
        assert np.abs(got[pairs[:, 0], pairs[:, 1]].astype(np.float64) - te64).max() < TOL


def test_noalloc_argument_errors(dev):
    b = synthetic.mixed_cardinality_bins(20, 500, seed=1)
    bins, npv, b_max, T = core.upload_bins(b, dev)
    prep = core.prepare(bins, npv, T, 1)
    plan = core.get_plan(prep)
    out = torch.zeros((20, 20), dtype=torch.float32, device=dev)
    lib = _lib.load()
    a = core.te_args(prep, plan, out, _lib.TNX_OUT_MATRIX, ld_out=20, persistent=True)
    assert lib.tnx_te_run_noalloc(C.byref(a), None) == -1       # queue_ctr given
    assert b"queue_ctr must be NULL" in lib.tnx_last_error()
    a = core.te_args(prep, plan, out, _lib.TNX_OUT_MATRIX, ld_out=20, persistent=False)
    assert lib.tnx_te_run_noalloc(C.byref(a), None) == -1       # no persistent plan
    assert b"persistent plan" in lib.tnx_last_error()
    assert lib.tnx_te_run_noalloc(None, None) == -1
