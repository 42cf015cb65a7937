"""Body forces: buoyancy and drag toward the wind (sim/forces.py of the reference)."""

import ctypes
import math

from plume_amd._lib import call, float3, ptr, stream


def apply_buoyancy(grid, gravity: float, heat_buoyancy: float, smoke_weight: float, ambient_k: float,
                   dt: float) -> None:
    """v += gravity * (heat_buoyancy * (T - ambient) / ambient - smoke_weight * smoke) * dt
    (sim/forces.py:26-31)."""
    call("plm_buoyancy", ctypes.byref(grid.info), ptr(grid.v.data), ptr(grid.smoke.data),
         ptr(grid.temperature.data), gravity, heat_buoyancy, smoke_weight, ambient_k, dt, stream())


def apply_drag(grid, wind_mps, drag_per_s: float, dt: float) -> None:
    """Relaxes every face velocity toward the wind by exp(-drag * dt); nothing when drag is 0
    (sim/forces.py:40-48)."""
    if drag_per_s == 0.0:
        return
    keep = math.exp(-drag_per_s * dt)
    u, v, w = (f.data for f in grid.velocity_fields())
    call("plm_drag", ctypes.byref(grid.info), ptr(u), ptr(v), ptr(w), float3(wind_mps), keep, stream())
