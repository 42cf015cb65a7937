"""The ray march (render/raymarch.py of the reference): one ray per view pixel, front to back, in fixed
half-cell steps, gathering thickness (Beer-Lambert), heat hidden by smoke in front, and the shade or the
normal. The kernel is plm_raymarch; this module holds its parameter structs."""

import ctypes

import torch

from plume_amd._lib import LookParams, ViewParams, call, int3, ptr, stream

FAR_M = 1.0e4  # how far toward the camera rays start (render/raymarch.py:21)

__all__ = ["FAR_M", "LookParams", "ViewParams", "march"]


def march(smoke: torch.Tensor, temperature: torch.Tensor, gradient: torch.Tensor, light: torch.Tensor, info,
          view: ViewParams, look: LookParams, image: torch.Tensor, max_steps: int = 0) -> None:
    """raymarch_kernel over every pixel of `image` ((height, width, 5) float32 on the GPU). max_steps > 0
    fixes every ray at that many samples (a bench mode); 0 is the reference's half-cell stepping."""
    height, width = image.shape[:2]
    call("plm_raymarch", ctypes.byref(info), ptr(smoke), ptr(temperature), ptr(gradient), int3(gradient.shape[:3]),
         ptr(light), int3(light.shape[:3]), ctypes.byref(view), ctypes.byref(look), ptr(image), height, width,
         int(max_steps), stream())
