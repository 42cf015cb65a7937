"""Test infrastructure: a numpy restatement of the viewer's kernels (view/ of the reference), the checker for
libplume_hip's plm_narrow_f16 / plm_widen_f16 / plm_narrow_center_velocity / plm_boundary_layers. Only the
tests may import it.

Pins (tests/test_view_host.py): binary16 rounding against hand-derived IEEE bit patterns (ties to even at
1 + 2^-11 and at the subnormal 2^-25, the largest finite 65504 and the overflow tie 65520, signed zero);
the centre of affine face fields in closed form; the boundary layers against the reference's own
arange case (tests/view/test_live_run.py:91-107).
"""

import numpy as np


def narrow(values: np.ndarray) -> np.ndarray:
    """narrow_kernel, view/frame_cache.py:26-29: FP32 to IEEE binary16, round to nearest even (values past
    the largest finite half round to infinity, as IEEE says)."""
    with np.errstate(over="ignore"):
        return np.asarray(values, dtype=np.float32).astype(np.float16)


def widen(values: np.ndarray) -> np.ndarray:
    """widen_kernel / widen_vector_kernel, view/frame_cache.py:32-35,45-49 (exact)."""
    return np.asarray(values, dtype=np.float16).astype(np.float32)


def center_velocity_half(u: np.ndarray, v: np.ndarray, w: np.ndarray) -> np.ndarray:
    """center_velocity_kernel (sim/vorticity.py:17-22) in FP32, then narrow_vector_kernel
    (view/frame_cache.py:38-42): (nx, ny, nz, 3) binary16."""
    half = np.float32(0.5)
    centre = np.stack([half * (u[:-1] + u[1:]), half * (v[:, :-1] + v[:, 1:]), half * (w[:, :, :-1] + w[:, :, 1:])],
                      axis=-1).astype(np.float32)
    return narrow(centre)


def boundary_layers(smoke: np.ndarray, extent_a: int, extent_b: int) -> np.ndarray:
    """boundary_layer_kernel, view/live_run.py:41-57: (6, extent_a, extent_b), zero past each face."""
    layers = np.zeros((6, extent_a, extent_b), dtype=np.float32)
    faces = (smoke[0], smoke[-1], smoke[:, 0], smoke[:, -1], smoke[:, :, 0], smoke[:, :, -1])
    for index, face in enumerate(faces):
        a, b = min(face.shape[0], extent_a), min(face.shape[1], extent_b)
        layers[index, :a, :b] = face[:a, :b]
    return layers
