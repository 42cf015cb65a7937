"""The CFL cap (sim/step_cap.py of the reference): enough steps that the fastest gas, measured at the start
of each block, crosses at most `max_cells_per_step` cells per step.

The reference copies u, v and w to the host (38 MB at 128 x 192 x 128) and reduces them in NumPy; here the
maximum |face velocity| is reduced on the device and one float crosses to the host. The maximum of float32
magnitudes is exact, so the step count is the same.
"""

import ctypes
import math
from dataclasses import dataclass

import numpy as np
import torch

from plume_amd._lib import call, ptr, stream
from plume_amd.recipe.errors import RecipeError


@dataclass(frozen=True)
class StepPlan:
    steps: int
    speed_mps: float


class StepCap:
    def __init__(self, recipe):
        self.cell_m = recipe.domain.cell_m
        self.max_cells = recipe.solver.max_cells_per_step
        self.max_substeps = recipe.solver.max_substeps
        self.emitters = [(e.start_s, e.end_s, _emitter_speed(e.velocity_mps, e.velocity_mode))
                         for e in recipe.emitter]
        self._speed = None

    def plan(self, grid, start_s: float, length_s: float, base_steps: int, label: str) -> StepPlan:
        """Steps for the block from `start_s` lasting `length_s`, at least `base_steps`
        (sim/step_cap.py:38-48)."""
        grid_speed = self.grid_speed(grid)
        if not math.isfinite(grid_speed):
            # the reference's math.ceil of a NaN or infinite speed raises here too (sim/step_cap.py:38-48)
            raise ValueError(f"{label}: the fastest face velocity is {grid_speed}: the simulation has blown up")
        speed = max(grid_speed, self.emitter_speed(start_s, start_s + length_s))
        needed = math.ceil(speed * length_s / (self.max_cells * self.cell_m))
        steps = max(base_steps, needed)
        if steps > max(self.max_substeps, base_steps):
            raise RecipeError(
                f"{label} needs {steps} steps to keep the fastest gas ({speed:.3f} m/s) within "
                f"solver.max_cells_per_step ({self.max_cells}); solver.max_substeps is {self.max_substeps}. "
                f"Raise it, or slow the gas (emitter velocity, buoyancy, turbulence, drag)")
        return StepPlan(steps, speed)

    def grid_speed(self, grid) -> float:
        """The fastest face velocity component, in m/s (sim/step_cap.py:50-55)."""
        if self._speed is None or self._speed.device != grid.pressure.device:
            self._speed = torch.zeros(1, dtype=torch.float32, device=grid.pressure.device)
        u, v, w = (f.data for f in grid.velocity_fields())
        call("plm_max_abs3", ctypes.byref(grid.info), ptr(u), ptr(v), ptr(w), ptr(self._speed), stream())
        return float(self._speed.item())

    def emitter_speed(self, start_s: float, end_s: float) -> float:
        speeds = [speed for begin, finish, speed in self.emitters if begin < end_s and finish > start_s]
        return max(speeds, default=0.0)


def _emitter_speed(velocity_mps, mode: str) -> float:
    velocity = np.asarray(velocity_mps, dtype=np.float64)
    if mode == "outward":
        return float(np.linalg.norm(velocity))
    return float(np.abs(velocity).max())
