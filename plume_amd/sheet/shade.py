"""The shade layout's B channel (sheet/shade.py of the reference): thin pixels eased into the shade of the
smoke around them, then the divisor that maps the run's brightest well-covered smoke to 1.

A thin pixel's own shade rests on little smoke and an empty pixel has none; a game that filters or mipmaps
the sheet would otherwise mix their speckle or a dark outline into the smoke's edge.
"""

import numpy as np

from plume_amd.render.raw_frame import SHADE, THICKNESS
from plume_amd.sheet.soften import blur

# Thickness that the frame's mean shade counts as in every neighbourhood: where the blurred smoke is well
# under one G step (1/255) the neighbourhood's shade leans to the frame's mean (sheet/shade.py:17).
FRAME_MEAN_WEIGHT = 1.0 / 255.0
# Thickness at or below this has no shade of its own (sheet/shade.py:19).
NO_SMOKE = 1e-12


def ease_thin_pixels(frame: np.ndarray, fade_thickness: float, neighborhood_px: float) -> np.ndarray:
    """The frame with its SHADE slot turned from shade * thickness into shade (sheet/shade.py:22-45):

        mean = sum(shade * thickness) / sum(thickness)          (0 for a frame without smoke)
        neighbourhood = (blur(shade * thickness) + m mean) / (blur(thickness) + m),  m = FRAME_MEAN_WEIGHT
        own = clamp(shade * thickness / thickness, 0, 1)        (0 where thickness <= NO_SMOKE)
        shade = neighbourhood + clamp(thickness / fade_thickness, 0, 1) (own - neighbourhood)

    with the blur a Gaussian of sigma `neighborhood_px`."""
    thickness = frame[..., THICKNESS]
    weighted = frame[..., SHADE]
    total = float(np.sum(thickness))
    mean = float(np.sum(weighted)) / total if total > 0.0 else 0.0
    neighborhood = ((blur(weighted, neighborhood_px) + FRAME_MEAN_WEIGHT * mean)
                    / (blur(thickness, neighborhood_px) + FRAME_MEAN_WEIGHT))
    smoky = thickness > NO_SMOKE
    own = np.zeros_like(thickness)
    np.divide(weighted, thickness, out=own, where=smoky)
    own = np.clip(own, 0.0, 1.0)
    weight = np.clip(thickness / fade_thickness, 0.0, 1.0)
    eased = frame.copy()
    eased[..., SHADE] = neighborhood + weight * (own - neighborhood)
    return eased


def shade_divisor(frames: list[np.ndarray], fade_thickness: float) -> float:
    """The largest (eased) shade over the run's pixels at least `fade_thickness` thick; failing any, over
    pixels with any smoke; failing that, 1 (sheet/shade.py:48-55)."""
    for floor in (fade_thickness, np.finfo(np.float64).tiny):
        largest = 0.0
        for frame in frames:
            covered = frame[..., THICKNESS] >= floor
            if covered.any():
                largest = max(largest, float(frame[..., SHADE][covered].max()))
        if largest > 0.0:
            return largest
    return 1.0
