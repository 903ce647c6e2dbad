"""GRN edge list (tenex/inference/grn.py:11-27)."""
from dataclasses import dataclass

import numpy as np


@dataclass
class GRN:
    """Gene regulatory network edges."""
    source: np.ndarray  # (n_edges,) source names
    target: np.ndarray  # (n_edges,) target names
    te: np.ndarray      # (n_edges,) float32 scores
    pairs: np.ndarray   # (n_edges, 2) (source_idx, target_idx)

    def __len__(self):
        return len(self.te)

    def to_sif(self) -> np.ndarray:
        return np.stack((self.source, self.te.astype(str), self.target), axis=1)

    def to_edge_list(self) -> list:
        return [(s, t, float(v)) for s, t, v in zip(self.source, self.target, self.te)]
