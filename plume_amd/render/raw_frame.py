"""The raw frame image (render/raw_frame.py of the reference): (height, width, 5) float32 holding heat,
thickness, then shade * thickness and two zeros (shade layout) or the camera-space normal (normal layout)."""

import numpy as np

from plume_amd.recipe.schema import NORMAL_LAYOUT

HEAT = 0
THICKNESS = 1
NORMAL_X = 2
NORMAL_Y = 3
NORMAL_Z = 4
SHADE = 2
CHANNELS = 5


def empty_pixel(channels: str) -> np.ndarray:
    """The raw value of empty space (render/raw_frame.py:24-29)."""
    pixel = np.zeros(CHANNELS, dtype=np.float32)
    if channels == NORMAL_LAYOUT:
        pixel[NORMAL_Z] = 1.0
    return pixel
