"""The crop square (sheet/crop.py of the reference): the square, in domain-view pixels, around every pixel
that any frame of the run shows, grown by a margin."""

from dataclasses import dataclass

import numpy as np

from plume_amd.render.raw_frame import HEAT, THICKNESS

# The smallest value that rounds to a byte of 1 (round half up of v * 255), so a pixel below it on every
# channel quantizes to nothing (sheet/crop.py:10).
CONTENT_LEVEL = 0.5 / 255.0


@dataclass(frozen=True)
class CropSquare:
    """A square in continuous domain-view pixel coordinates; it may reach past the view (sheet/crop.py:13-22)."""

    left: float
    top: float
    side: float

    def box(self) -> tuple[float, float, float, float]:
        return (self.left, self.top, self.left + self.side, self.top + self.side)


class NothingVisibleError(Exception):
    """No frame shows any smoke or heat, so there is no content to crop to."""


def content_box(raw_frames: list[np.ndarray]) -> tuple[int, int, int, int]:
    """(left, top, right, bottom) in pixels, right and bottom exclusive, of the pixels any frame shows: thickness
    of at least CONTENT_LEVEL, or heat of at least CONTENT_LEVEL of the run's hottest pixel
    (sheet/crop.py:29-42)."""
    hottest = max(float(np.max(frame[..., HEAT])) for frame in raw_frames)
    shown = np.zeros(raw_frames[0].shape[:2], dtype=bool)
    for frame in raw_frames:
        np.logical_or(shown, frame[..., THICKNESS] >= CONTENT_LEVEL, out=shown)
        if hottest > 0.0:
            np.logical_or(shown, frame[..., HEAT] >= CONTENT_LEVEL * hottest, out=shown)
    rows = np.flatnonzero(shown.any(axis=1))
    if rows.size == 0:
        raise NothingVisibleError("no frame shows any smoke or heat; check the emitters and the camera")
    columns = np.flatnonzero(shown.any(axis=0))
    return int(columns[0]), int(rows[0]), int(columns[-1]) + 1, int(rows[-1]) + 1


def square_crop(box: tuple[int, int, int, int], margin: float) -> CropSquare:
    """The box's longer side times (1 + 2 margin), as a square on the box's centre (sheet/crop.py:45-50)."""
    left, top, right, bottom = box
    side = max(right - left, bottom - top) * (1.0 + 2.0 * margin)
    return CropSquare(left=(left + right) / 2.0 - side / 2.0, top=(top + bottom) / 2.0 - side / 2.0, side=side)
