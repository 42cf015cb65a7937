"""Orthographic camera and the sheet's domain view (the interface of render/camera.py:16-78), host-side in float64.

The camera's axes are the columns of one rotation, R = R_y(azimuth) · R_x(-elevation): column 0 is the
camera's right, column 1 its up and column 2 the direction from the scene toward the camera (camera-space +z,
OpenGL style). At elevation 0 and azimuth 0 the camera sits on the +z side looking along -z; elevation tilts it
to look down, azimuth orbits it counterclockwise seen from above.

The domain view is the smallest pixel rectangle on the camera plane that holds the projection of the whole
domain box: the box's eight corners are projected on right and up, the rectangle spans their extremes, and its
size in pixels rounds up (with a 1e-6 pixel allowance, so a span of an exact pixel count is not rounded up one
more by float error). Pixel (0, 0) is the top-left one.
"""

import math
from dataclasses import dataclass

import numpy as np


def _rotation(elevation_deg: float, azimuth_deg: float) -> np.ndarray:
    """R_y(azimuth) · R_x(-elevation), written out: columns right, up, toward."""
    ce, se = math.cos(math.radians(elevation_deg)), math.sin(math.radians(elevation_deg))
    ca, sa = math.cos(math.radians(azimuth_deg)), math.sin(math.radians(azimuth_deg))
    # + 0.0 turns the -0.0 of a zero product into +0.0
    return np.array([[ca, -se * sa, ce * sa],
                     [0.0, ce, se],
                     [-sa, -se * ca, ce * ca]]) + 0.0


@dataclass(frozen=True)
class Camera:
    """World-space unit axes of the camera."""

    right: np.ndarray
    up: np.ndarray
    toward: np.ndarray  # from the scene toward the camera: camera-space +z

    @property
    def basis(self) -> np.ndarray:
        """3×3 matrix whose rows are right, up and toward (world → camera space)."""
        return np.stack([self.right, self.up, self.toward])

    def to_camera_space(self, world_vector: np.ndarray) -> np.ndarray:
        return self.basis @ np.asarray(world_vector, dtype=np.float64)


def make_camera(elevation_deg: float, azimuth_deg: float) -> Camera:
    """The camera for an elevation and an azimuth in degrees (render/camera.py:26-34)."""
    rotation = _rotation(elevation_deg, azimuth_deg)
    return Camera(right=rotation[:, 0].copy(), up=rotation[:, 1].copy(), toward=rotation[:, 2].copy())


@dataclass(frozen=True)
class DomainView:
    """A pixel grid on the camera plane: its top-left corner (left_m, top_m) in camera-plane metres, the pixel
    pitch and the size in pixels (render/camera.py:37-52)."""

    camera: Camera
    left_m: float
    top_m: float
    pixel_m: float
    width_px: int
    height_px: int

    def to_pixels(self, world_point: np.ndarray) -> tuple[float, float]:
        """Continuous pixel coordinates (x to the right, y down) of a world point's projection."""
        plane_x, plane_y = np.asarray(world_point, dtype=np.float64) @ np.stack([self.camera.right, self.camera.up]).T
        return float((plane_x - self.left_m) / self.pixel_m), float((self.top_m - plane_y) / self.pixel_m)


def _pixels_spanning(extent_m: float, pixel_m: float) -> int:
    return math.ceil(extent_m / pixel_m - 1e-6)


def make_domain_view(camera: Camera, domain_min_m, domain_size_m, cell_m: float, px_per_cell: float) -> DomainView:
    """The smallest view at `px_per_cell` pixels per cell covering the projected domain box
    (render/camera.py:55-71)."""
    low = np.asarray(domain_min_m, dtype=np.float64)
    size = np.asarray(domain_size_m, dtype=np.float64)
    picks = np.array([[a, b, c] for a in (0.0, 1.0) for b in (0.0, 1.0) for c in (0.0, 1.0)])
    projected = np.array([[(low + size * pick) @ axis for axis in (camera.right, camera.up)] for pick in picks])
    lo, hi = projected.min(axis=0), projected.max(axis=0)
    pixel_m = cell_m / px_per_cell
    return DomainView(camera=camera, left_m=float(lo[0]), top_m=float(hi[1]), pixel_m=pixel_m,
                      width_px=_pixels_spanning(hi[0] - lo[0], pixel_m),
                      height_px=_pixels_spanning(hi[1] - lo[1], pixel_m))


def square_domain_view(camera: Camera, domain_min_m, domain_size_m, pixels: int) -> DomainView:
    """A `pixels` x `pixels` view centred on the projected domain box, with the pixel pitch that fits the box's
    longer projected side (no reference counterpart: the benchmark's fixed 128² and 256² images)."""
    low = np.asarray(domain_min_m, dtype=np.float64)
    size = np.asarray(domain_size_m, dtype=np.float64)
    picks = np.array([[a, b, c] for a in (0.0, 1.0) for b in (0.0, 1.0) for c in (0.0, 1.0)])
    projected = np.array([[(low + size * pick) @ axis for axis in (camera.right, camera.up)] for pick in picks])
    lo, hi = projected.min(axis=0), projected.max(axis=0)
    span = float(max(hi - lo))
    pixel_m = span / pixels
    center = (lo + hi) / 2.0
    return DomainView(camera=camera, left_m=float(center[0] - span / 2.0), top_m=float(center[1] + span / 2.0),
                      pixel_m=pixel_m, width_px=pixels, height_px=pixels)


def sheet_domain_view(recipe, derived) -> DomainView:
    """The view `plume run` renders every frame over: the recipe's camera around the whole domain
    (render/camera.py:74-78)."""
    return make_domain_view(make_camera(recipe.camera.elevation_deg, recipe.camera.azimuth_deg),
                            derived.domain_min_m, derived.size_m, recipe.domain.cell_m, recipe.render.px_per_cell)
