"""Resampling a frame's crop square to the sheet's frame size (sheet/resample.py of the reference): Pillow's
bicubic resize over the crop box, which widens its filter when it shrinks, on every channel of the raw
frame set on a canvas of empty space large enough that no filter tap leaves it."""

import math

import numpy as np
from PIL import Image

from plume_amd.recipe.schema import SHADE_LAYOUT
from plume_amd.render.raw_frame import CHANNELS, HEAT, SHADE, THICKNESS, empty_pixel
from plume_amd.sheet.crop import CropSquare

BICUBIC_SUPPORT = 2.0  # Pillow's bicubic filter reaches 2 source pixels, times the shrink factor when shrinking


def resample_frame(raw: np.ndarray, crop: CropSquare, frame_px: int, channels: str) -> np.ndarray:
    """The crop square of a (height, width, 5) raw frame as a (frame_px, frame_px, 5) float64 frame, with
    thickness clamped to [0, 1], heat to >= 0 and (shade layout) shade times thickness to [0, thickness]
    (sheet/resample.py:16-30)."""
    height, width = raw.shape[:2]
    left, top, right, bottom = crop.box()
    shrink = max(1.0, crop.side / frame_px)
    # every tap of every output pixel stays on the canvas, so the canvas size does not change a value
    overhang = max(0.0, -left, -top, right - width, bottom - height)
    reach = math.ceil(overhang + BICUBIC_SUPPORT * shrink) + 2
    canvas = np.empty((height + 2 * reach, width + 2 * reach, CHANNELS), dtype=np.float32)
    canvas[...] = empty_pixel(channels)
    canvas[reach:reach + height, reach:reach + width] = raw
    box = (left + reach, top + reach, right + reach, bottom + reach)
    frame = np.empty((frame_px, frame_px, CHANNELS), dtype=np.float64)
    for channel in range(CHANNELS):
        plane = Image.fromarray(np.ascontiguousarray(canvas[..., channel]))  # float32: mode "F"
        frame[..., channel] = np.asarray(plane.resize((frame_px, frame_px), Image.Resampling.BICUBIC, box=box))
    np.clip(frame[..., THICKNESS], 0.0, 1.0, out=frame[..., THICKNESS])
    np.maximum(frame[..., HEAT], 0.0, out=frame[..., HEAT])
    if channels == SHADE_LAYOUT:  # bicubic ringing may leave shade * thickness outside [0, thickness]
        frame[..., SHADE] = np.minimum(np.maximum(frame[..., SHADE], 0.0), frame[..., THICKNESS])
    return frame
