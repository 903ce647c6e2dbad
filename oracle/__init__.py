"""CPU oracle for the transfer-entropy hot path.

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by the cpu_baseline leg of bench.py, always as the checker or the CPU
baseline, never by the product package (tenex_amd/).

Contents
--------
oracle.py          ctypes bindings to libtnx_oracle.so (tnx_oracle.c, FP64,
                   exact integer counts) plus numpy restatements of the
                   reference's host arithmetic (FSBW-L, remap, pair order,
                   surrogate statistics, BH).
scatter_add_ref.py torch-CPU restatement of the reference's scatter_add TE
                   batch (tenex/kernels/scatter_add.py:33-119), op for op:
                   the CPU baseline.

Parity status of every function is stated in its docstring; none is
"parity unpinned" except where explicitly said.
"""
