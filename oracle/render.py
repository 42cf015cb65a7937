"""The oracle's frame render: light volume or normal field, then the march (render/frame_renderer.py:16-94,
render/light_volume.py:36-51, render/normal_field.py:41-69 of the reference), over numpy arrays.

TEST INFRASTRUCTURE ONLY (see `oracle/__init__.py`). `view` is a DomainView-like object (camera basis in
float64, left_m, top_m, pixel_m, width_px, height_px); `settings` a RenderSettings-like object.
"""

import math

import numpy as np

SHADE_LAYOUT = "heat,thickness,shade,unused"


def toward_light_world(camera, light_dir):
    x, y, z = light_dir
    return x * camera.right + y * camera.up + z * camera.toward


def gaussian_weights(sigma: float, radius: int) -> np.ndarray:
    """render/normal_field.py:64-69."""
    if radius == 0:
        return np.ones(1, dtype=np.float32)
    offsets = np.arange(-radius, radius + 1, dtype=np.float64)
    weights = np.exp(-offsets**2 / (2.0 * sigma**2))
    return (weights / weights.sum()).astype(np.float32)


def light_volume(o, smoke, toward_light_world_dir, opacity_per_density, cell_m, max_steps=0):
    direction = np.asarray(toward_light_world_dir, dtype=np.float64)
    unit = direction / np.linalg.norm(direction)
    light = np.zeros(smoke.shape, dtype=o.dtype)
    o.light_volume(smoke, unit, cell_m, opacity_per_density, light, max_steps)
    return light


def normal_gradient(o, smoke, blur_cells):
    radius = math.ceil(3.0 * blur_cells) if blur_cells > 0 else 0
    weights = gaussian_weights(blur_cells, radius).astype(o.dtype)
    blurred = np.zeros_like(smoke)
    scratch = np.zeros_like(smoke)
    source = smoke
    for axis, target in enumerate((blurred, scratch, blurred)):
        o.blur_axis(source, target, weights, radius, axis)
        source = target
    gradient = np.zeros(smoke.shape + (3,), dtype=o.dtype)
    o.gradient(source, gradient)
    return gradient


def view_params(o, view):
    """domain_view_params, render/frame_renderer.py:72-83."""
    camera = view.camera
    return o.view(camera.right, camera.up, camera.toward, view.left_m, view.top_m, view.pixel_m, 0, (0.0, 0.0, 0.0))


def render_oracle(o, grid, view, settings, channels, params=None, height=None, width=None, max_steps=0,
                  light_steps=0):
    """The raw frame (height, width, 5) of grid.smoke / grid.temperature (FrameRenderer.render). max_steps and
    light_steps > 0 are the benchmark's fixed sample counts per ray and per light path."""
    shade = channels == SHADE_LAYOUT
    smoke, temperature = grid.smoke.data, grid.temperature.data
    if shade:
        light = light_volume(o, smoke, toward_light_world(view.camera, settings.light_dir),
                             settings.opacity_per_density, grid.info.cell_m, light_steps)
        gradient = np.zeros((1, 1, 1, 3), dtype=o.dtype)
    else:
        gradient = normal_gradient(o, smoke, settings.normal_blur_cells)
        light = np.zeros((1, 1, 1), dtype=o.dtype)
    look = o.look(settings.opacity_per_density, settings.glow_start_k, settings.smoke_hides_heat,
                  settings.normal_fade_thickness, 1 if shade else 0, settings.shade_fill)
    params = view_params(o, view) if params is None else params
    height = view.height_px if height is None else height
    width = view.width_px if width is None else width
    image = np.zeros((height, width, 5), dtype=o.dtype)
    o.raymarch(grid.info, smoke, temperature, gradient, light, params, look, image, max_steps)
    return image
