"""What a live run notes per frame (view/live_run.py of the reference): which open faces, and the ceiling, the
smoke reached, read from the six boundary layers of the smoke (plm_boundary_layers). The run loop, sheet
preview and window state of the reference's LiveRun are not part of this package."""

import math

import numpy as np
import torch

from plume_amd._lib import call, int3, ptr, stream

FACE_NAMES = ("left", "right", "bottom", "top", "back", "front")  # x low, x high, y low, y high, z low, z high
CEILING = "top"


def boundary_layers(smoke: torch.Tensor, layers: torch.Tensor) -> torch.Tensor:
    """Copies the smoke's six boundary layers into `layers` (6, A, B): face in FACE_NAMES order, a and b the
    face's other two axes in x, y, z order, 0 past a face's size (view/live_run.py:41-57)."""
    if layers.dim() != 3 or layers.shape[0] != len(FACE_NAMES):
        raise ValueError(f"layers must be (6, A, B); got {tuple(layers.shape)}")
    call("plm_boundary_layers", int3(smoke.shape), ptr(smoke), ptr(layers), int(layers.shape[1]),
         int(layers.shape[2]), stream())
    return layers


def visible_density(recipe) -> float:
    """Smoke density at which one cell's opacity reaches one byte step, 1/255 (view/live_run.py:60-62)."""
    return -math.log(1.0 - 1.0 / 255.0) / (recipe.render.opacity_per_density * recipe.domain.cell_m)


def faces_reached(layers: np.ndarray, recipe) -> tuple[str, ...]:
    """The open faces, and the ceiling, whose boundary layer holds visible smoke, in FACE_NAMES order
    (view/live_run.py:65-73)."""
    largest = np.asarray(layers).reshape(len(FACE_NAMES), -1).max(axis=1)
    threshold = visible_density(recipe)
    faces = recipe.domain.faces
    return tuple(name for name, value in zip(FACE_NAMES, largest, strict=True)
                 if value >= threshold and (name == CEILING or getattr(faces, name) == "open"))


class EdgeWatch:
    """Per-run layers scratch (6, side, side), side = the longest cell count (view/live_run.py:167-172)."""

    def __init__(self, recipe, cells, device: str):
        self.recipe = recipe
        side = max(cells)
        self.layers = torch.zeros((len(FACE_NAMES), side, side), dtype=torch.float32, device=device)

    def faces_reached(self, smoke: torch.Tensor) -> tuple[str, ...]:
        boundary_layers(smoke, self.layers)
        return faces_reached(self.layers.cpu().numpy(), self.recipe)
