"""The soften pass (sheet/soften.py of the reference): a separable Gaussian blur of every channel with the
edge pixels extended, then, in the normal layout, normals made unit length again (facing the camera)."""

import math

import numpy as np

from plume_amd.recipe.schema import NORMAL_LAYOUT
from plume_amd.render.raw_frame import NORMAL_X, NORMAL_Z

MIN_NORMAL_LENGTH = 1e-6


def soften(frame: np.ndarray, sigma_px: float, channels: str) -> np.ndarray:
    """A (size, size, 5) frame blurred by `sigma_px`, normals renormalized in the normal layout
    (sheet/soften.py:12-16)."""
    softened = blur(frame, sigma_px)
    return renormalize(softened) if channels == NORMAL_LAYOUT else softened


def gaussian_weights(sigma_px: float) -> np.ndarray:
    """exp(-d^2 / (2 sigma^2)) for d = -ceil(3 sigma) .. ceil(3 sigma), summing to 1 (sheet/soften.py:27-31)."""
    radius = math.ceil(3.0 * sigma_px)
    d = np.arange(-radius, radius + 1, dtype=np.float64)
    w = np.exp(-(d * d) / (2.0 * sigma_px * sigma_px))
    return w / w.sum()


def edge_clamped_operator(size: int, weights: np.ndarray) -> np.ndarray:
    """The (size, size) matrix of a 1-D convolution with `weights` whose taps past either end read the end
    pixel: row i holds weight w[t] at column clamp(i + t - radius, 0, size - 1), taps that land on the same
    column summed."""
    radius = len(weights) // 2
    operator = np.zeros((size, size), dtype=np.float64)
    rows = np.arange(size)
    for tap, weight in enumerate(weights):
        np.add.at(operator, (rows, np.clip(rows + tap - radius, 0, size - 1)), weight)
    return operator


def blur(image: np.ndarray, sigma_px: float) -> np.ndarray:
    """A Gaussian blur over the first two axes of `image` (edges extended); sigma 0 returns `image`
    (sheet/soften.py:19-24)."""
    if sigma_px <= 0.0:
        return image
    weights = gaussian_weights(sigma_px)
    down = edge_clamped_operator(image.shape[0], weights)
    across = edge_clamped_operator(image.shape[1], weights)
    blurred = np.tensordot(down, image, axes=(1, 0))  # along rows (axis 0)
    return np.moveaxis(np.tensordot(across, blurred, axes=(1, 1)), 0, 1)  # then along columns (axis 1)


def renormalize(frame: np.ndarray) -> np.ndarray:
    """The frame with its camera-space normal turned toward the camera (|z|) and scaled to unit length;
    a normal shorter than 1e-6 becomes (0, 0, 1) (sheet/soften.py:46-53)."""
    result = frame.copy()
    normal = result[..., NORMAL_X:NORMAL_Z + 1]
    normal[..., 2] = np.abs(normal[..., 2])
    length = np.sqrt(np.sum(normal * normal, axis=-1, keepdims=True))
    unit = normal / np.maximum(length, MIN_NORMAL_LENGTH)
    result[..., NORMAL_X:NORMAL_Z + 1] = np.where(length > MIN_NORMAL_LENGTH, unit, np.array([0.0, 0.0, 1.0]))
    return result
