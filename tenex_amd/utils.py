"""Data loading (tenex/utils.py:13-100): expression CSV and 1-D time files
with .npy caches next to the source file (or under ~/.cache/tenex when its
folder is read-only), as the reference writes and reads them."""
from __future__ import annotations

import hashlib
import os
import os.path as osp

import numpy as np
import torch


def load_exp_data(dpath: str, make_binary: bool = False):
    """Raw (unaligned) expression matrix from a CSV whose first row holds
    the gene names and first column the cell labels: returns (node_name
    (n_genes,) str, exp_data (n_genes, n_cells) float32).  make_binary:
    values > 0 -> 1.  Caches *_raw.npy / *_raw_node_name.npy (the legacy
    names without _raw are read too)."""
    suffix = "_binary" if make_binary else ""
    ddir = osp.dirname(osp.abspath(dpath))
    base, _ = osp.splitext(dpath)
    if os.access(ddir, os.W_OK):
        f_bin, f_name = base + suffix + "_raw.npy", base + suffix + "_raw_node_name.npy"
        f_bin_old, f_name_old = base + suffix + ".npy", base + suffix + "_node_name.npy"
    else:
        cache = osp.expanduser("~/.cache/tenex")
        os.makedirs(cache, exist_ok=True)
        key = hashlib.md5((dpath + suffix).encode()).hexdigest()[:12]
        f_bin, f_name = osp.join(cache, key + "_raw.npy"), osp.join(cache, key + "_raw_node_name.npy")
        f_bin_old, f_name_old = osp.join(cache, key + ".npy"), osp.join(cache, key + "_node_name.npy")
    for fb, fn in ((f_bin, f_name), (f_bin_old, f_name_old)):
        if osp.isfile(fb) and osp.isfile(fn):
            return np.load(fn, allow_pickle=True), np.load(fb)
    raw = np.loadtxt(dpath, delimiter=",", dtype=str)
    node_name = raw[0, 1:]
    exp_data = raw[1:, 1:].T.astype(np.float32)
    if make_binary:
        exp_data = (exp_data > 0).astype(np.float32)
    np.save(f_bin, exp_data)
    np.save(f_name, node_name)
    return node_name, exp_data


def load_time_data(dpath: str, dtype=np.float32) -> np.ndarray:
    """A 1-D text file (pseudotime, branch labels) with a .npy cache keyed on
    the dtype (<base>_<dtype>.npy)."""
    base, _ = osp.splitext(dpath)
    tag = np.dtype(dtype).str.lstrip("<>=|")
    f_npy = f"{base}_{tag}.npy"
    if osp.isfile(f_npy):
        return np.load(f_npy)
    data = np.loadtxt(dpath, dtype=dtype)
    if os.access(osp.dirname(osp.abspath(dpath)), os.W_OK):
        np.save(f_npy, data)
    return data


def get_device_list() -> list[int]:
    return list(range(torch.cuda.device_count()))
