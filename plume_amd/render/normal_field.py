"""The smoke surface direction field (render/normal_field.py of the reference): the gradient of a
Gaussian-blurred copy of the smoke, three separable blur passes then a central-difference gradient."""

import math

import numpy as np
import torch

from plume_amd._lib import call, ptr, shape3, stream


def gaussian_weights(sigma: float, radius: int) -> np.ndarray:
    """render/normal_field.py:64-69."""
    if radius == 0:
        return np.ones(1, dtype=np.float32)
    offsets = np.arange(-radius, radius + 1, dtype=np.float64)
    weights = np.exp(-offsets**2 / (2.0 * sigma**2))
    return (weights / weights.sum()).astype(np.float32)


class NormalField:
    """Owns the scratch arrays and computes the blurred smoke gradient once per frame
    (render/normal_field.py:41-61)."""

    def __init__(self, cells, blur_cells: float, device: str):
        self.device = device
        cells = tuple(cells)
        self.blurred = torch.zeros(cells, dtype=torch.float32, device=device)
        self.scratch = torch.zeros(cells, dtype=torch.float32, device=device)
        self.gradient = torch.zeros(cells + (3,), dtype=torch.float32, device=device)
        self.radius = math.ceil(3.0 * blur_cells) if blur_cells > 0 else 0
        self.weights = torch.from_numpy(gaussian_weights(blur_cells, self.radius)).to(device)

    def compute(self, smoke: torch.Tensor) -> torch.Tensor:
        """Returns the gradient (per cell) of the blurred smoke."""
        source = smoke
        for axis, target in enumerate((self.blurred, self.scratch, self.blurred)):
            call("plm_blur_axis", shape3(smoke), ptr(source), ptr(target), ptr(self.weights), self.radius, axis,
                 stream())
            source = target
        call("plm_gradient", shape3(smoke), ptr(source), ptr(self.gradient), stream())
        return self.gradient
