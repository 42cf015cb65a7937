"""Advection: clamped MacCormack (Selle et al. 2008), sim/advect.py of the reference.

MacCormack runs semi-Lagrangian advection forward, then backward, and corrects the first result by half the
round trip's error, clamped to the eight values the forward trace blended. dt / cell_m is formed in double
and rounded once, as the reference's Python does (sim/advect.py:62).
"""

import ctypes

from plume_amd._lib import call, float3, ptr, shape3, stream


def _dt_over_cell(dt: float, grid) -> float:
    return float(dt) / grid.info.cell_m


def advect_semi_lagrangian(field, velocity, grid, dt: float) -> None:
    """Plain semi-Lagrangian advection into `field.next` (sim/advect.py:53-56)."""
    u, v, w = velocity
    call("plm_advect_sl", ctypes.byref(grid.info), ptr(field.data), ptr(field.next), shape3(field.data),
         float3(field.offset), field.outside, field.uses_outside, ptr(u), ptr(v), ptr(w), _dt_over_cell(dt, grid),
         stream())


def advect_maccormack(field, velocity, grid, dt: float, fuse_back: bool = True) -> None:
    """Clamped MacCormack advection of `field.data` into `field.next` (sim/advect.py:59-68). With
    fuse_back the backward pass runs inside the correction (same values; `field.back` is not written)."""
    u, v, w = velocity
    call("plm_advect_maccormack", ctypes.byref(grid.info), ptr(field.data), ptr(field.hat),
         ptr(field.back), ptr(field.next), shape3(field.data), float3(field.offset), field.outside,
         field.uses_outside, ptr(u), ptr(v), ptr(w), _dt_over_cell(dt, grid), 1 if fuse_back else 0, stream())


def advect_velocity(grid, dt: float) -> None:
    """MacCormack of u, v and w by the pre-advection velocity into their `next` arrays, in two launches
    (the loop of sim/stepper.py:46-50, before the swaps)."""
    u, v, w = grid.velocity_fields()
    call("plm_advect_velocity", ctypes.byref(grid.info), ptr(u.data), ptr(v.data), ptr(w.data), ptr(u.hat),
         ptr(v.hat), ptr(w.hat), ptr(u.next), ptr(v.next), ptr(w.next), _dt_over_cell(dt, grid), stream())
