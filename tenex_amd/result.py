"""TransferEntropyResult: the (n, n) TE matrix plus metadata (tenex/result.py:12-105)."""
from __future__ import annotations

import numpy as np


class TransferEntropyResult:
    """Result of TransferEntropyEngine.compute().

    matrix[i, j] = TE(var_i -> var_j) (float32); bin_arrs (n, T, 1) int8;
    n_per_var int32; b_max; tau; kernel name; timings (profile=True); sources.
    Behaves like the matrix for numpy calls and indexing.
    """

    __slots__ = ("matrix", "variable_names", "bin_arrs", "n_per_var", "b_max", "tau", "kernel",
                 "timings", "sources")

    def __init__(self, matrix, variable_names, bin_arrs=None, n_per_var=None, b_max=0, tau=1,
                 kernel="", timings=None, sources=None):
        self.matrix = matrix
        self.variable_names = np.asarray(variable_names)
        self.bin_arrs = bin_arrs
        self.n_per_var = n_per_var
        self.b_max = b_max
        self.tau = tau
        self.kernel = kernel
        self.timings = timings
        self.sources = sources

    def __array__(self, dtype=None, copy=None):
        arr = self.matrix if dtype is None else self.matrix.astype(dtype)
        if copy:
            arr = arr.copy()
        return arr

    def __getitem__(self, key):
        return self.matrix[key]

    @property
    def shape(self):
        return self.matrix.shape

    @property
    def dtype(self):
        return self.matrix.dtype

    @property
    def ndim(self):
        return self.matrix.ndim

    @property
    def size(self):
        return self.matrix.size

    @property
    def T(self):
        return self.matrix.T

    def __len__(self):
        return len(self.matrix)

    @property
    def n_vars(self) -> int:
        return self.matrix.shape[0]

    def __repr__(self):
        return (f"TransferEntropyResult(n_vars={self.n_vars}, kernel='{self.kernel}', "
                f"tau={self.tau}, b_max={self.b_max})")
