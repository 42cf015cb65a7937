"""The sheet's frames from a run's raw frames (sheet/assemble.py of the reference), in order: the union
content box, its square crop, bicubic resampling, the soften blur, (shade layout) thin-pixel easing and the
shade divisor, the heat divisor, quantization, and the gutter check, which grows the margin and starts over
from the crop while some frame's content reaches its 2-pixel edge ring."""

from dataclasses import dataclass

import numpy as np

from plume_amd.recipe.schema import SHADE_LAYOUT
from plume_amd.render.raw_frame import HEAT
from plume_amd.sheet.crop import CropSquare, content_box, square_crop
from plume_amd.sheet.quantize import quantize
from plume_amd.sheet.resample import resample_frame
from plume_amd.sheet.shade import ease_thin_pixels, shade_divisor
from plume_amd.sheet.soften import soften

GUTTER_PX = 2
MARGIN_GROWTH = 0.02
MAX_MARGIN_TRIES = 8


@dataclass(frozen=True)
class AssembledFrames:
    """The quantized frames and the facts the sheet's JSON and the reports use (sheet/assemble.py:25-34)."""

    frames: list[np.ndarray]  # (frame_px, frame_px, 4) uint8 each
    crop: CropSquare
    margin: float
    frame_m: float  # the crop's side in metres
    seat: tuple[float, float]  # the seat point in frame fractions
    heat_divisor: float  # K m: the softened heat that maps to R = 255
    shade_divisor: float | None  # the eased shade that maps to B = 255; None in the normal layout


class GutterError(Exception):
    """Content still reaches a frame's edge after the largest margin growth."""


def gutter_is_empty(frame: np.ndarray) -> bool:
    """True when R and G are 0 on the frame's outer GUTTER_PX rows and columns."""
    g = GUTTER_PX
    edges = (frame[:g, :, :2], frame[-g:, :, :2], frame[:, :g, :2], frame[:, -g:, :2])
    return not any(edge.any() for edge in edges)


def finish(raw_frames: list[np.ndarray], crop: CropSquare, frame_px: int, render, channels: str):
    """(quantized frames, heat divisor, shade divisor or None) of the raw frames through one crop."""
    frames = [soften(resample_frame(raw, crop, frame_px, channels), render.soften_px, channels)
              for raw in raw_frames]
    divisor = None
    if channels == SHADE_LAYOUT:
        frames = [ease_thin_pixels(frame, render.shade_fade_thickness, render.shade_neighborhood_px)
                  for frame in frames]
        divisor = shade_divisor(frames, render.shade_fade_thickness)
    heat_divisor = max(float(frame[..., HEAT].max()) for frame in frames)
    return [quantize(frame, heat_divisor, channels, divisor or 1.0) for frame in frames], heat_divisor, divisor


def assemble(raw_frames: list[np.ndarray], view, seat_world_m: np.ndarray, frame_px: int, render,
             channels: str) -> AssembledFrames:
    """The sheet frames of a run (sheet/assemble.py:41-65)."""
    box = content_box(raw_frames)
    margin = render.crop_margin
    for _ in range(MAX_MARGIN_TRIES + 1):
        crop = square_crop(box, margin)
        frames, heat_divisor, divisor = finish(raw_frames, crop, frame_px, render, channels)
        if all(gutter_is_empty(frame) for frame in frames):
            x, y = view.to_pixels(seat_world_m)
            return AssembledFrames(frames=frames, crop=crop, margin=margin, frame_m=crop.side * view.pixel_m,
                                   seat=((x - crop.left) / crop.side, (y - crop.top) / crop.side),
                                   heat_divisor=heat_divisor, shade_divisor=divisor)
        margin += MARGIN_GROWTH
    raise GutterError(f"content still touches the frame edge with crop_margin {margin:.2f}")


def seat_point(recipe) -> np.ndarray:
    """The world point the sheet's seat marks: the centre of the emitter `camera.seat` names
    (sheet/assemble.py:68-71)."""
    emitter = next(e for e in recipe.emitter if e.name == recipe.camera.seat)
    return np.array(emitter.center_m, dtype=np.float64)
