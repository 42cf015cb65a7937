"""Renders smoke and temperature into a raw frame image (render/frame_renderer.py of the reference)."""

import numpy as np
import torch

from plume_amd._lib import LookParams, ViewParams, float3
from plume_amd.recipe.schema import SHADE_LAYOUT
from plume_amd.render.light_volume import LightVolume
from plume_amd.render.normal_field import NormalField
from plume_amd.render.raw_frame import CHANNELS
from plume_amd.render.raymarch import march as launch_march


class FrameRenderer:
    """Per-run render state (view, settings, scratch arrays); renders one frame at a time
    (render/frame_renderer.py:16-36). The layout picks the volume pass; the other gets a 1-cell stand-in."""

    def __init__(self, view, settings, cells, channels: str, device: str, max_steps: int = 0,
                 light_steps: int = 0):
        self.device = device
        self.shade_layout = channels == SHADE_LAYOUT
        self.view_params = domain_view_params(view)
        self.look_params = look_params(settings, self.shade_layout)
        self.max_steps = int(max_steps)
        if self.shade_layout:
            light_world = toward_light_world(view.camera, settings.light_dir)
            self.light_volume = LightVolume(cells, light_world, settings.opacity_per_density, device, light_steps)
            self.gradient = torch.zeros((1, 1, 1, 3), dtype=torch.float32, device=device)
            self.light = self.light_volume.light
        else:
            self.normal_field = NormalField(cells, settings.normal_blur_cells, device)
            self.light = torch.zeros((1, 1, 1), dtype=torch.float32, device=device)
            self.gradient = self.normal_field.gradient
        self.image = torch.zeros((view.height_px, view.width_px, CHANNELS), dtype=torch.float32, device=device)

    def render(self, grid) -> np.ndarray:
        """The raw frame of the grid's current smoke and temperature, copied to the host
        (render/frame_renderer.py:38-41)."""
        return self.render_fields(grid.smoke.data, grid.temperature.data, grid.info)

    def render_fields(self, smoke: torch.Tensor, temperature: torch.Tensor, info) -> np.ndarray:
        """render/frame_renderer.py:43-48."""
        self.prepare(smoke, info.cell_m)
        self.march(smoke, temperature, info, self.view_params, self.image)
        return self.image.cpu().numpy().copy()

    def prepare(self, smoke: torch.Tensor, cell_m: float) -> None:
        """Runs the layout's volume pass on a smoke field (render/frame_renderer.py:50-56)."""
        if self.shade_layout:
            self.light = self.light_volume.compute(smoke, cell_m)
        else:
            self.gradient = self.normal_field.compute(smoke)

    def march(self, smoke, temperature, info, view: ViewParams, image: torch.Tensor) -> None:
        """Marches every pixel of `image` for `view` with the last `prepare`'s volume pass
        (render/frame_renderer.py:58-63)."""
        launch_march(smoke, temperature, self.gradient, self.light, info, view, self.look_params, image,
                     self.max_steps)


def toward_light_world(camera, light_dir) -> np.ndarray:
    """The world direction toward the key light; render.light_dir is in camera space
    (render/frame_renderer.py:66-69)."""
    x, y, z = light_dir
    return x * camera.right + y * camera.up + z * camera.toward


def domain_view_params(view) -> ViewParams:
    """The march's orthographic view over a domain view (render/frame_renderer.py:72-83)."""
    params = ViewParams()
    params.right = float3(view.camera.right)
    params.up = float3(view.camera.up)
    params.toward = float3(view.camera.toward)
    params.left_m = view.left_m
    params.top_m = view.top_m
    params.pixel_m = view.pixel_m
    params.perspective = 0
    params.eye = float3((0.0, 0.0, 0.0))
    return params


def look_params(settings, shade_layout: bool) -> LookParams:
    """render/frame_renderer.py:86-94."""
    params = LookParams()
    params.opacity_per_density = settings.opacity_per_density
    params.glow_start_k = settings.glow_start_k
    params.smoke_hides_heat = settings.smoke_hides_heat
    params.normal_fade_thickness = settings.normal_fade_thickness
    params.shade_layout = 1 if shade_layout else 0
    params.shade_fill = settings.shade_fill
    return params
