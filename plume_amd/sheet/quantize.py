"""The sheet's four 8-bit channels from a finished float frame (sheet/quantize.py of the reference).

R = heat / heat divisor and G = thickness in both layouts. Shade layout: B = shade / shade divisor and A =
255 (unused). Normal layout: B and A = camera-space normal x and y mapped from -1..1 to 0..255, and exactly
128 where R and G are both 0. Every value maps to floor(v * 255 + 0.5) clamped to 0..255 (round half up).
"""

import numpy as np

from plume_amd.recipe.schema import SHADE_LAYOUT
from plume_amd.render.raw_frame import HEAT, NORMAL_X, NORMAL_Y, SHADE, THICKNESS

FACING_BYTE = 128
UNUSED_BYTE = 255


def to_byte(values: np.ndarray) -> np.ndarray:
    """floor(v * 255 + 0.5) clamped to 0..255, as uint8."""
    return np.clip(np.floor(np.asarray(values) * 255.0 + 0.5), 0.0, 255.0).astype(np.uint8)


def quantize(frame: np.ndarray, heat_divisor: float, channels: str, shade_divisor: float = 1.0) -> np.ndarray:
    """The (size, size, 4) uint8 RGBA frame of a (size, size, 5) float frame (sheet/quantize.py:21-35)."""
    size = frame.shape[:2]
    rgba = np.empty((*size, 4), dtype=np.uint8)
    rgba[..., 0] = to_byte(frame[..., HEAT] / heat_divisor) if heat_divisor > 0.0 else 0
    rgba[..., 1] = to_byte(frame[..., THICKNESS])
    if channels == SHADE_LAYOUT:
        rgba[..., 2] = to_byte(frame[..., SHADE] / shade_divisor)
        rgba[..., 3] = UNUSED_BYTE
        return rgba
    rgba[..., 2] = to_byte((frame[..., NORMAL_X] + 1.0) * 0.5)
    rgba[..., 3] = to_byte((frame[..., NORMAL_Y] + 1.0) * 0.5)
    empty = (rgba[..., 0] == 0) & (rgba[..., 1] == 0)
    rgba[empty, 2:] = FACING_BYTE
    return rgba
