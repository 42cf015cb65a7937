"""Emitters: sources that add smoke and hold temperature and velocity inside a soft-edged shape
(sim/emitters.py of the reference). Every emitter kernel gathers over the emitter list."""

import ctypes

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


def emit(grid, emitters: EmitterArray, seed: int, time_s: float, dt: float) -> None:
    """Applies every emitter active at `time_s` for one step of length `dt` (sim/emitters.py:147-153)."""
    u, v, w = (f.data for f in grid.velocity_fields())
    call("plm_emit", ctypes.byref(grid.info), ptr(grid.smoke.data), ptr(grid.temperature.data), ptr(u), ptr(v),
         ptr(w), ptr(emitters.buffer, torch.uint8), emitters.count, int(seed), time_s, dt, stream())
