"""scRNA-seq ingest (tenex/io.py:17-136): expression, pseudotime and branch
labels (files or arrays) -> cells of one branch in pseudotime order."""
from __future__ import annotations

import os
import os.path as osp
from dataclasses import dataclass
from typing import Optional

import numpy as np

from .preprocess import align_data
from .utils import load_exp_data, load_time_data


@dataclass
class ScRnaData:
    """data (n_genes, T) float32 aligned; gene_names (n_genes,); sources
    (optional source-variable names)."""
    data: np.ndarray
    gene_names: np.ndarray
    sources: Optional[np.ndarray] = None


def _load_str_with_npy_cache(path: str) -> np.ndarray:
    base, _ = osp.splitext(path)
    f_npy = base + ".npy"
    if osp.isfile(f_npy):
        return np.load(f_npy, allow_pickle=True)
    data = np.loadtxt(path, dtype=str)
    if os.access(osp.dirname(osp.abspath(path)), os.W_OK):
        np.save(f_npy, data)
    return data


def load_scrna(expression, pseudotime, branch, gene_names=None, branch_id: int = 1,
               sources=None, make_binary: bool = False) -> ScRnaData:
    """Load and align scRNA-seq data (tenex/io.py:58-136): file paths or
    arrays; cells of branch_id sorted by pseudotime."""
    if isinstance(expression, str):
        names, exp_data = load_exp_data(expression, make_binary)
        if gene_names is None:
            gene_names = names
    else:
        exp_data = np.asarray(expression, dtype=np.float32)
        if gene_names is None:
            raise ValueError("gene_names is required when expression is an ndarray")
        if make_binary:
            exp_data = (exp_data > 0).astype(np.float32)
    pseudotime = (load_time_data(pseudotime, dtype=np.float32) if isinstance(pseudotime, str)
                  else np.asarray(pseudotime, dtype=np.float32))
    branch = (load_time_data(branch, dtype=np.int32) if isinstance(branch, str)
              else np.asarray(branch, dtype=np.int32))
    if isinstance(sources, str):
        sources = _load_str_with_npy_cache(sources)
    elif sources is not None:
        sources = np.asarray(sources)
    aligned = align_data(exp_data, pseudotime, branch, branch_id=branch_id)
    return ScRnaData(data=aligned, gene_names=np.asarray(gene_names), sources=sources)
