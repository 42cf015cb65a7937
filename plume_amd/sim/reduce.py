"""Deterministic sums of grid fields (sim/reduce.py of the reference): one thread per (j, k) line sums over
i in a fixed order; the line totals are then added in float64."""

import numpy as np
import torch

from plume_amd._lib import call, int3, ptr, stream


class TripleSum:
    """Totals of three same-shaped fields, with one copy to the host (sim/reduce.py:28-38)."""

    def __init__(self, shape, device: str):
        self.shape = tuple(shape)
        self.lines = torch.zeros((3, shape[1], shape[2]), dtype=torch.float32, device=device)

    def __call__(self, a: torch.Tensor, b: torch.Tensor, c: torch.Tensor) -> tuple[float, float, float]:
        call("plm_line_sums", int3(self.shape), ptr(a), ptr(b), ptr(c), ptr(self.lines), stream())
        totals = self.lines.cpu().numpy().astype(np.float64).sum(axis=(1, 2))
        return float(totals[0]), float(totals[1]), float(totals[2])
