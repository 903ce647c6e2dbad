"""ctypes binding of libtenex_hip.so (include/tenex_hip.h).

The product path has no fallback: if the library is missing or a GPU is not
available, the calls below raise.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libtenex_hip.so")
ABI_VERSION = 13

TNX_OUT_PAIRFREE = 0
TNX_OUT_MATRIX = 1
TNX_OUT_STATS = 2
TNX_OUT_STAGE = 3

TNX_B2_AUTO = 0
TNX_B2_INT8 = 1
TNX_B2_FP4 = 2

_lock = threading.Lock()
_lib = None

p_ = C.c_void_p
i64 = C.c_int64
i32 = C.c_int32


class DeviceProps(C.Structure):
    _fields_ = [("lds_bytes_per_block", i64), ("n_cu", i32), ("wave_size", i32),
                ("clock_khz", i32), ("gcn_arch", i32), ("name", C.c_char * 64)]


class Work(C.Structure):
    _fields_ = [("tgt_beg", i32), ("tgt_cnt", i32), ("src_beg", i32), ("src_end", i32)]


class PlanInfo(C.Structure):
    _fields_ = [("n_wave_items", i64), ("n_block_items", i64), ("n_pairs", i64),
                ("cap_wave", i32), ("cap_block", i32), ("max_block_cols", i32),
                ("n_tgt", i32), ("n_src", i32),
                ("est_cost_wave", C.c_double), ("est_cost_block", C.c_double)]


class TeArgs(C.Structure):
    _fields_ = [("bins", p_), ("states", p_), ("n_per_var", p_), ("S1", p_), ("S2a", p_),
                ("xlogx", p_), ("tgt_order", p_), ("src_order", p_), ("work", p_),
                ("n_vars", i64), ("T", i64), ("ld", i64), ("lds", i64),
                ("tau", i32), ("out_mode", i32),
                ("n_wave_items", i64), ("n_block_items", i64),
                ("cap_wave", i32), ("cap_block", i32), ("max_block_cols", i32),
                ("out", p_), ("ld_out", i64), ("pair_offset", i64), ("n_pairs_local", i64),
                ("sum_te", p_), ("sumsq_te", p_), ("count_ge", p_), ("observed", p_),
                ("hot_zero", C.c_int32), ("stage_vars", C.c_int32), ("stage_stride", C.c_int64),
                ("item_per_wave", C.c_int32),
                ("queue_items", p_), ("queue_off", p_), ("queue_ctr", p_),
                ("counts_map", p_), ("counts_dbg", p_), ("counts_ld", C.c_int64),
                ("xlogx_fx", p_), ("fx_shift", C.c_int32), ("var_L", p_),
                ("cap_sort", i64), ("hot_cw", C.c_int32), ("hot_maxb", C.c_int32)]


class OnehotArgs(C.Structure):
    _fields_ = [("A", p_), ("B", p_), ("Kp", i64), ("rows_a", i64), ("rows_b", i64),
                ("arow", p_), ("brow", p_), ("marg", p_), ("moff", p_),
                ("tiles", p_), ("n_tiles", i64),
                ("rg_off", p_), ("t_var", p_), ("t_row", p_),
                ("cg_off", p_), ("s_var", p_), ("s_col", p_),
                ("mvars", p_), ("n_mvars", i64), ("src_ie", i32), ("pad0", i32)]


TNX_QUEUES = 8
TNX_QUEUE_CTR_INTS = 256
TNX_NOALLOC_SCRATCH_INTS = 256


class SurrArgs(C.Structure):
    _fields_ = [("bins", p_), ("n_per_var", p_), ("xlogx", p_), ("perm", p_),
                ("tgt_order", p_), ("src_order", p_), ("work", p_), ("observed", p_),
                ("n_vars", i64), ("T", i64), ("ld", i64),
                ("tau", i32), ("n_surr", i32), ("n_blocks", i32), ("block_length", i32),
                ("n_wave_items", i64), ("n_block_items", i64),
                ("cap_wave", i32), ("cap_block", i32), ("max_block_cols", i32),
                ("sum_te", p_), ("sumsq_te", p_), ("count_ge", p_), ("te_k_dbg", p_),
                ("max_levels", i32), ("max_tgt_cnt", i32)]


class Mlp(C.Structure):
    _fields_ = [("W1", p_), ("W2", p_), ("W3", p_), ("b1", p_), ("b2", p_), ("b3", p_),
                ("D", i32), ("H", i32), ("O", i32)]


class Adam(C.Structure):
    _fields_ = [("lr", C.c_float), ("lr_min", C.c_float), ("beta1", C.c_float),
                ("beta2", C.c_float), ("eps", C.c_float), ("steps", i32), ("t0", i32),
                ("batch", i32)]


# name -> (restype, argtypes)
SIGNATURES = {
    "tnx_abi_version": (C.c_int, []),
    "tnx_last_error": (C.c_char_p, []),
    "tnx_device_props": (C.c_int, [C.c_int, C.POINTER(DeviceProps)]),
    "tnx_fsbw_l": (C.c_int, [p_, i64, i64, C.c_float, p_, p_, p_, p_, p_]),
    "tnx_remap_dense": (C.c_int, [p_, i64, i64, p_, i64, p_, p_, p_]),
    "tnx_coarsen_bins": (C.c_int, [p_, i64, i64, i64, p_, i32, p_]),
    "tnx_xlogx_table": (C.c_int, [p_, i64, p_]),
    "tnx_xlogx_table_fx": (C.c_int, [p_, i64, i32, p_]),
    "tnx_gene_prep": (C.c_int, [p_, i64, i64, i64, i32, p_, p_, i64, p_, p_, p_, i32, p_]),
    "tnx_te_plan": (C.c_int, [p_, i64, i64, p_, i64, p_, i64, i32, i32, p_, p_, p_, p_, i64,
                              C.POINTER(PlanInfo)]),
    "tnx_onehot_pack": (C.c_int, [C.POINTER(OnehotArgs), p_, i64, i64, i32, p_, p_]),
    "tnx_onehot_te": (C.c_int, [C.POINTER(OnehotArgs), C.POINTER(TeArgs), p_]),
    "tnx_te_run": (C.c_int, [C.POINTER(TeArgs), p_]),
    "tnx_te_run_noalloc": (C.c_int, [C.POINTER(TeArgs), p_]),
    "tnx_remap_dense16": (C.c_int, [p_, i64, i64, p_, i64, p_, p_, p_]),
    "tnx_coarsen_bins16": (C.c_int, [p_, i64, i64, i64, p_, i32, p_, p_]),
    "tnx_te_sort_pairs": (C.c_int, [p_, i64, i64, i64, i32, p_, i64, p_, p_, p_, i32, p_, i64,
                                    p_]),
    "tnx_gene_prep_sort": (C.c_int, [p_, i64, i64, i64, i32, p_, i64, p_, i32, p_, p_, p_]),
    "tnx_binarize": (C.c_int, [p_, i64, i64, i64, i32, C.c_float, p_, i64, p_, p_]),
    "tnx_trace_entropy": (C.c_int, [p_, i64, i64, i64, i32, i32, p_, p_, p_, i64, p_, i32, p_,
                                    p_, p_]),
    "tnx_te_queue": (C.c_int, [p_, i64, i64, p_, i64, p_, p_, i64, i32, i32, p_, i64, p_,
                               C.POINTER(C.c_int64)]),
    "tnx_te_counts_debug": (C.c_int, [p_, p_, p_, p_, p_, p_, i64, i64, i64, i64, i32, p_, i64,
                                      i32, p_, p_, p_]),
    "tnx_gemm_b2_workspace": (i64, [i64, i64, i32]),
    "tnx_te_gemm_b2": (C.c_int, [p_, i64, i64, i64, i32, p_, p_, p_, p_, i64, p_, i64, p_, i32,
                                 p_]),
    "tnx_surrogate_run": (C.c_int, [C.POINTER(SurrArgs), p_]),
    "tnx_block_shuffle": (C.c_int, [p_, i64, i64, i64, p_, i32, i32, p_, p_]),
    "tnx_block_shuffle_copies": (C.c_int, [p_, i64, i32, i64, i64, p_, i32, i32, p_, p_]),
    "tnx_stats_accumulate": (C.c_int, [p_, i32, i64, p_, i64, p_, p_, p_, p_, p_]),
    "tnx_dpi_maxmin": (C.c_int, [p_, i64, i64, i64, p_, i64, i64, p_, i64, C.c_float, p_, i64,
                                 p_]),
    "tnx_mlp_forward": (C.c_int, [C.POINTER(Mlp), p_, i64, p_, p_]),
    "tnx_mlp_workspace_bytes": (i64, [C.c_int32, C.c_int32, i64]),
    "tnx_mlp_backward": (C.c_int, [C.POINTER(Mlp), p_, p_, p_, i64, p_, p_, p_, i64, p_]),
    "tnx_mlp_fit": (C.c_int, [C.POINTER(Mlp), p_, p_, p_, i64, p_, p_, i64, C.POINTER(Adam), p_,
                              p_, i64, p_, p_]),
    "tnx_mlp_adam": (C.c_int, [C.POINTER(Mlp), p_, p_, p_, p_, i64, C.c_float, C.c_float,
                               C.c_float, C.c_float, i32, p_]),
    "tnx_bench_lds_atomics": (C.c_int, [i32, i64, i32, p_, p_]),
    "tnx_bench_mfma": (C.c_int, [i32, i64, i32, p_, p_]),
}


class TnxError(RuntimeError):
    """A libtenex_hip call returned an error status."""


def load():
    """Load libtenex_hip.so (raises if it is absent or has another ABI)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"libtenex_hip.so not found at {LIB_PATH}; build it with "
                f"`python -m tenex_amd._build` (hipcc, gfx950)")
        lib = C.CDLL(LIB_PATH)
        missing = [n for n in SIGNATURES if not hasattr(lib, n)]
        if missing:
            raise RuntimeError(f"libtenex_hip.so lacks exports {missing}; rebuild it")
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        v = lib.tnx_abi_version()
        if v != ABI_VERSION:
            raise RuntimeError(f"libtenex_hip ABI {v} != expected {ABI_VERSION}")
        _lib = lib
    return _lib


def call(name: str, *args) -> int:
    """Call an entry point; raise TnxError with tnx_last_error() on failure."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.tnx_last_error().decode(errors="replace")
        if rc == -1:
            raise ValueError(f"{name}: {msg}")
        raise TnxError(f"{name} failed ({rc}): {msg}")
    return rc


def ptr(t) -> int | None:
    """Raw device (or host) address of a torch tensor / numpy array, or None."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        return t.data_ptr()
    return t.ctypes.data


def stream_handle(device=None) -> int:
    import torch
    return torch.cuda.current_stream(device).cuda_stream
