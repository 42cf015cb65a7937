"""Curl-noise turbulence (sim/turbulence.py of the reference): the discrete curl of a seeded value-noise
potential, blended between time slices, added to every face velocity."""

import ctypes

from plume_amd._lib import call, ptr, stream

TURBULENCE_STREAM = 1 << 24


def apply_turbulence(grid, seed: int, scale_m: float, strength_mps2: float, speed_mps: float, time_s: float,
                     dt: float) -> None:
    """Adds one step of the curl-noise push at `time_s`; nothing when strength is 0
    (sim/turbulence.py:59-67). pattern_time is formed in double, as the reference does."""
    if strength_mps2 == 0.0:
        return
    pattern_time = time_s * speed_mps / scale_m
    u, v, w = (f.data for f in grid.velocity_fields())
    call("plm_turbulence", ctypes.byref(grid.info), ptr(u), ptr(v), ptr(w), int(seed), scale_m, strength_mps2,
         pattern_time, dt, stream())
