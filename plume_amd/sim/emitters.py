"""Emitters: sources that add smoke and hold temperature and velocity inside a soft-edged shape
(sim/emitters.py of the reference). Every emitter kernel gathers over the emitter list."""

import ctypes
import math

import numpy as np
import torch

from plume_amd._lib import EmitterParams, call, float3, ptr, stream

SPHERE = 0
BOX = 1
BLEND_REFERENCE_S = 1.0 / 80.0


class EmitterArray:
    """The recipe's emitters as a device array of plm_emitter structs (one H2D copy per run)."""

    def __init__(self, params: list[EmitterParams], device: str):
        self.count = len(params)
        raw = (EmitterParams * max(1, self.count))(*params)
        host = torch.frombuffer(bytearray(bytes(raw)), dtype=torch.uint8)
        self.buffer = host.to(device)
        self.params = params


def emitter_struct(settings, index: int) -> EmitterParams:
    """One emitter's struct, with noise stream = its index (sim/emitters.py:124-144)."""
    p = EmitterParams()
    p.shape = SPHERE if settings.shape == "sphere" else BOX
    p.center_m = float3(settings.center_m)
    p.radius_m = settings.radius_m
    p.half_size_m = float3(size / 2 for size in settings.size_m)
    p.edge_softness = settings.edge_softness
    p.start_s = settings.start_s
    p.end_s = settings.end_s
    p.smoke_per_s = settings.smoke_per_s
    p.temperature_k = settings.temperature_k
    p.velocity_mps = float3(settings.velocity_mps)
    p.outward = 1 if settings.velocity_mode == "outward" else 0
    p.noise_scale_m = settings.noise.scale_m
    p.noise_strength = settings.noise.strength
    p.stream = index
    return p


def emitter_params(emitters, device: str) -> EmitterArray:
    """The recipe's emitters as a device array, each with its own noise stream (sim/emitters.py:124-144)."""
    return EmitterArray([emitter_struct(e, i) for i, e in enumerate(emitters)], device)


def emit_box(info, params: list[EmitterParams], time_s: float):
    """The box of cells [lo, hi) that holds the support of every emitter active at `time_s` (is_active,
    sim/emitters.py:80-82, compared in float32 as the kernels compare), or None when no active emitter reaches
    the grid. Outside an emitter's centre +- extent (radius, or half size per axis) its coverage is 0
    (soft_edge, sim/emitters.py:41-50), so cells and faces outside the box are left as they are; the box is
    widened by two cells on each side against rounding."""
    t = np.float32(time_s)
    lo, hi = [int(info.dims[a]) for a in range(3)], [0, 0, 0]
    for p in params:
        if not (np.float32(p.start_s) <= t < np.float32(p.end_s)):
            continue
        extent = [p.radius_m] * 3 if p.shape == SPHERE else list(p.half_size_m)
        for a in range(3):
            first = math.floor((p.center_m[a] - extent[a] - info.origin_m[a]) / info.cell_m) - 2
            last = math.ceil((p.center_m[a] + extent[a] - info.origin_m[a]) / info.cell_m) + 2
            lo[a] = min(lo[a], max(0, first))
            hi[a] = max(hi[a], min(int(info.dims[a]), last))
    if any(lo[a] >= hi[a] for a in range(3)):
        return None
    return lo, hi


def emit(grid, emitters: EmitterArray, seed: int, time_s: float, dt: float) -> None:
    """Applies every emitter active at `time_s` for one step of length `dt` (sim/emitters.py:147-153), launched
    over the active emitters' box of cells (the same result as over the whole grid)."""
    box = emit_box(grid.info, emitters.params, time_s)
    if box is None:
        return
    u, v, w = (f.data for f in grid.velocity_fields())
    lo, hi = ((ctypes.c_int32 * 3)(*b) for b in box)
    call("plm_emit_box", ctypes.byref(grid.info), ptr(grid.smoke.data), ptr(grid.temperature.data), ptr(u), ptr(v),
         ptr(w), ptr(emitters.buffer, torch.uint8), emitters.count, int(seed), time_s, dt, lo, hi, stream())


def emit_whole_grid(grid, emitters: EmitterArray, seed: int, time_s: float, dt: float) -> None:
    """emit over every cell and face of the grid (plm_emit), the reference's launch shape."""
    u, v, w = (f.data for f in grid.velocity_fields())
    call("plm_emit", ctypes.byref(grid.info), ptr(grid.smoke.data), ptr(grid.temperature.data), ptr(u), ptr(v),
         ptr(w), ptr(emitters.buffer, torch.uint8), emitters.count, int(seed), time_s, dt, stream())
