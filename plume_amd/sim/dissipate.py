"""Dissipation: smoke fades and temperature relaxes toward ambient (sim/dissipate.py of the reference)."""

import math

from plume_amd._lib import call, ptr, stream


def dissipate(grid, smoke_per_s: float, cooling_per_s: float, ambient_k: float, dt: float) -> None:
    """One step of exponential fading; nothing when both rates are 0 (sim/dissipate.py:18-24)."""
    if smoke_per_s == 0.0 and cooling_per_s == 0.0:
        return
    call("plm_dissipate", grid.smoke.data.numel(), ptr(grid.smoke.data), ptr(grid.temperature.data),
         math.exp(-smoke_per_s * dt), math.exp(-cooling_per_s * dt), ambient_k, stream())
