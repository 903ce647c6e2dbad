"""The C-ABI library loads and exports every entry point include/*.h declares,
and the ctypes structs match the C layouts (no GPU calls)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from tenex_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tenex_hip.h")


def declared_symbols():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"TNX_API\s+[\w\s\*]+?\b(tnx_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    syms = declared_symbols()
    assert len(syms) >= 15
    lib = C.CDLL(_lib.LIB_PATH)
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    assert set(syms) == set(_lib.SIGNATURES)


def test_abi_version_and_error_reporting():
    lib = _lib.load()
    assert lib.tnx_abi_version() == _lib.ABI_VERSION == 13
    info = _lib.PlanInfo()
    rc = lib.tnx_te_plan(None, 3, 10, None, 0, None, 0, 0, 0, None, None, None, None, 0,
                         C.byref(info))
    assert rc == -1
    assert b"null pointer" in lib.tnx_last_error()


def test_struct_layouts_match_header(tmp_path):
    prog = tmp_path / "sz.c"
    prog.write_text(f'''#include <stdio.h>
#include <stddef.h>
#include "{HEADER}"
int main(void) {{
  printf("%zu %zu %zu %zu %zu %zu %zu\\n", sizeof(tnx_device_props_t), sizeof(tnx_work_t),
         sizeof(tnx_plan_info_t), sizeof(tnx_te_args_t), sizeof(tnx_surr_args_t),
         sizeof(tnx_mlp_t), sizeof(tnx_adam_t));
  printf("%zu %zu %zu %zu %zu\\n", offsetof(tnx_te_args_t, out), offsetof(tnx_surr_args_t, sum_te),
         offsetof(tnx_plan_info_t, est_cost_wave), offsetof(tnx_te_args_t, queue_items),
         offsetof(tnx_te_args_t, counts_ld));
  printf("%zu %zu %zu\\n", sizeof(tnx_onehot_args_t), offsetof(tnx_onehot_args_t, tiles),
         offsetof(tnx_onehot_args_t, n_mvars));
  return 0;
}}''')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", str(prog), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    got = [C.sizeof(_lib.DeviceProps), C.sizeof(_lib.Work), C.sizeof(_lib.PlanInfo),
           C.sizeof(_lib.TeArgs), C.sizeof(_lib.SurrArgs), C.sizeof(_lib.Mlp),
           C.sizeof(_lib.Adam), _lib.TeArgs.out.offset, _lib.SurrArgs.sum_te.offset, _lib.PlanInfo.est_cost_wave.offset,
           _lib.TeArgs.queue_items.offset, _lib.TeArgs.counts_ld.offset,
           C.sizeof(_lib.OnehotArgs), _lib.OnehotArgs.tiles.offset, _lib.OnehotArgs.n_mvars.offset]
    assert [int(v) for v in out] == got


def _dummy_args(lds):
    """TeArgs with non-null placeholder pointers: tnx_te_run validates them
    before any HIP call."""
    a = _lib.TeArgs()
    for f in ("bins", "states", "n_per_var", "S1", "S2a", "xlogx", "tgt_order", "src_order",
              "work", "out"):
        setattr(a, f, 64)
    a.n_vars, a.T, a.ld, a.lds, a.tau = 4, 100, 128, lds, 1
    a.out_mode, a.cap_wave = _lib.TNX_OUT_MATRIX, 3136
    return a


def test_te_run_rejects_state_stride_not_multiple_of_16():
    lib = _lib.load()
    a = _dummy_args(104)  # >= L = 99 and a multiple of 8, not of 16
    assert lib.tnx_te_run(C.byref(a), None) == -1
    assert b"multiple of 16" in lib.tnx_last_error()


def test_te_run_rejects_counts_hook_without_queue():
    lib = _lib.load()
    a = _dummy_args(112)
    a.counts_map = 64
    a.counts_dbg = 64
    a.counts_ld = 8
    assert lib.tnx_te_run(C.byref(a), None) == -1
    assert b"persistent" in lib.tnx_last_error()


def test_te_queue_partitions_every_entry():
    """tnx_te_queue (host): every wave slot (or, with item_per_wave, wave
    item) and every block item is queued exactly once, wave entries in the
    first eight queues and block entries in the next eight, and the entries
    of one source chunk share an XCD."""
    import numpy as np
    rng = np.random.default_rng(5)
    n, L = 700, 2999
    npv = rng.integers(2, 40, size=n).astype(np.int32)
    info = _lib.PlanInfo()
    args = (npv.ctypes.data, n, L, None, 0, None, 0, 0, 0, None)
    assert _lib.call("tnx_te_plan", *args, None, None, None, 0, C.byref(info)) == 0
    nw, nb = info.n_wave_items, info.n_block_items
    assert nw > 0 and nb > 0
    tg = np.empty(info.n_tgt, np.int32)
    sr = np.empty(info.n_src, np.int32)
    work = np.empty((nw + nb, 4), np.int32)
    _lib.call("tnx_te_plan", *args, tg.ctypes.data, sr.ctypes.data, work.ctypes.data, nw + nb,
              C.byref(info))
    npair = work[:, 1].astype(np.int64) * (work[:, 3] - work[:, 2])
    for per_wave in (0, 1):
        nwe = C.c_int64(0)
        qa = (work.ctypes.data, nw, nb, npv.ctypes.data, n, tg.ctypes.data, sr.ctypes.data, L,
              0, per_wave)
        _lib.call("tnx_te_queue", *qa, None, 0, None, C.byref(nwe))
        if per_wave:
            exp_wave = np.arange(nw)
        else:
            exp_wave = np.concatenate([4 * i + np.arange(min(4, npair[i])) for i in range(nw)])
        assert nwe.value == len(exp_wave)
        n_ent = nwe.value + nb
        items = np.empty(n_ent, np.int32)
        off = np.zeros(17, np.int32)
        _lib.call("tnx_te_queue", *qa, items.ctypes.data, n_ent, off.ctypes.data, C.byref(nwe))
        assert off[0] == 0 and off[8] == nwe.value and off[16] == n_ent
        assert np.all(np.diff(off) >= 0)
        wave_part, block_part = items[: off[8]], items[off[8]:]
        assert np.array_equal(np.sort(wave_part), np.sort(exp_wave))
        assert np.array_equal(np.sort(block_part), np.arange(nb))
        # the XCD of each entry, by the source chunk (64 sources) of its item
        q_of = np.repeat(np.arange(16) % 8, np.diff(off))
        item = np.concatenate([wave_part if per_wave else wave_part // 4, nw + block_part])
        chunk = work[item, 2] // 64
        for c in np.unique(chunk):
            assert len(np.unique(q_of[chunk == c])) == 1
        lib = _lib.load()
        assert lib.tnx_te_queue(*qa, items.ctypes.data, n_ent - 1, off.ctypes.data,
                                C.byref(nwe)) == -4
