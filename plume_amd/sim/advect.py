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


def advect_velocity(grid, dt: float, carry_bounds: bool = True) -> None:
    """MacCormack of u, v and w by the pre-advection velocity into their `next` arrays, in two launches
    (the loop of sim/stepper.py:46-50, before the swaps). carry_bounds: the forward pass hands the clamp
    range to the correction (same values, one trace fewer per sample)."""
    u, v, w = grid.velocity_fields()
    if carry_bounds:
        call("plm_advect_velocity_bounded", ctypes.byref(grid.info), ptr(u.data), ptr(v.data), ptr(w.data),
             ptr(u.hat), ptr(v.hat), ptr(w.hat), ptr(u.bounds), ptr(v.bounds), ptr(w.bounds), ptr(u.next),
             ptr(v.next), ptr(w.next), _dt_over_cell(dt, grid), stream())
        return
    call("plm_advect_velocity", ctypes.byref(grid.info), ptr(u.data), ptr(v.data), ptr(w.data), ptr(u.hat),
         ptr(v.hat), ptr(w.hat), ptr(u.next), ptr(v.next), ptr(w.next), _dt_over_cell(dt, grid), stream())


def advect_scalars(grid, velocity, dt: float, fields=None) -> None:
    """MacCormack of cell-centred fields (default: smoke and temperature) by one velocity, one shared trace
    per cell, into their `next` arrays (sim/stepper.py:55-59 without the budget and the swaps)."""
    fields = (grid.smoke, grid.temperature) if fields is None else tuple(fields)
    if not 1 <= len(fields) <= 2 or any(f.offset != (0.5, 0.5, 0.5) or f.uses_outside != 1 for f in fields):
        raise ValueError("advect_scalars takes one or two cell-centred fields that use their outside value")
    a = fields[0]
    b = fields[1] if len(fields) == 2 else None
    u, v, w = velocity
    call("plm_advect_scalars", ctypes.byref(grid.info), ptr(a.data), ptr(a.hat), ptr(a.bounds), ptr(a.next),
         a.outside, ptr(b.data) if b else None, ptr(b.hat) if b else None, ptr(b.bounds) if b else None,
         ptr(b.next) if b else None, b.outside if b else 0.0, ptr(u), ptr(v), ptr(w), _dt_over_cell(dt, grid),
         stream())
