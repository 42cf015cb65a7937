"""The light volume (render/light_volume.py of the reference): the share of the key light reaching each cell,
exp(-opacity_per_density * column density toward the light), marched per cell in half-cell steps."""

import numpy as np
import torch

from plume_amd._lib import call, float3, ptr, shape3, stream


class LightVolume:
    """Owns the light array and fills it once per rendered frame (render/light_volume.py:36-51).
    max_steps > 0 fixes the march at that many steps per cell (a bench mode); 0 is the reference's."""

    def __init__(self, cells, toward_light_world, opacity_per_density: float, device: str, max_steps: int = 0):
        self.device = device
        direction = np.asarray(toward_light_world, dtype=np.float64)
        self.toward_light = tuple(float(x) for x in direction / np.linalg.norm(direction))
        self.opacity_per_density = opacity_per_density
        self.max_steps = int(max_steps)
        self.light = torch.zeros(tuple(cells), dtype=torch.float32, device=device)

    def compute(self, smoke: torch.Tensor, cell_m: float) -> torch.Tensor:
        """The light reaching each cell of a smoke field (cells of `cell_m`), 0 to 1."""
        call("plm_light_volume", shape3(smoke), ptr(smoke), float3(self.toward_light), cell_m,
             self.opacity_per_density, self.max_steps, ptr(self.light), stream())
        return self.light
