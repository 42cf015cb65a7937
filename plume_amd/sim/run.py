"""The frame loop (sim/run.py of the reference): warmup, then each frame's steps followed by a capture.

Time runs in blocks of the base step count; a block's start is always `base steps so far * dt`, and the CFL
cap may split a block into more, shorter steps. Frame i is captured after warmup_s + (i + 1) / fps.
"""

from collections.abc import Callable, Generator
from dataclasses import dataclass

from plume_amd.sim.grid import Grid, make_grid
from plume_amd.sim.step_cap import StepCap
from plume_amd.sim.stepper import Stepper

Capture = Callable[[int, Grid], None]


@dataclass(frozen=True)
class StepLog:
    """How many steps each part of the run took, after the CFL cap (sim/run.py:22-39)."""

    warmup_steps: int
    steps_per_frame: tuple[int, ...]
    base_steps_per_frame: int
    base_warmup_steps: int

    @property
    def total_steps(self) -> int:
        return self.warmup_steps + sum(self.steps_per_frame)

    @property
    def added_steps(self) -> int:
        base = self.base_warmup_steps + self.base_steps_per_frame * len(self.steps_per_frame)
        return self.total_steps - base


_TUNED_GRIDS: set[tuple[int, int, int]] = set()


def tune_sweep_tiles(solve, cells) -> None:
    """Pins the fastest tile shape of the two-colour sweep for each multigrid level of this grid size, once
    per process (a few milliseconds; every shape gives bit-identical results, so only the speed changes)."""
    key = tuple(int(c) for c in cells)
    if key in _TUNED_GRIDS or not hasattr(solve, "autotune"):
        return
    solve.autotune()
    _TUNED_GRIDS.add(key)


def simulate(recipe, derived, device: str, capture: Capture) -> StepLog:
    """Runs the whole recipe, calling `capture(frame_index, grid)` once per frame (sim/run.py:42-50)."""
    frames = simulate_frames(recipe, derived, device)
    while True:
        try:
            frame, grid = next(frames)
        except StopIteration as finished:
            return finished.value
        capture(frame, grid)


def simulate_frames(recipe, derived, device: str) -> Generator[tuple[int, Grid], None, StepLog]:
    """Yields (frame_index, grid) after each frame's steps and returns the step log (sim/run.py:53-85)."""
    grid = make_grid(derived, recipe.domain.cell_m, recipe.domain.faces, recipe.domain.ambient_k, device)
    stepper = Stepper(grid, recipe)
    tune_sweep_tiles(stepper.pressure_solve, derived.cells)
    cap = StepCap(recipe)
    base_done = 0

    def run_block(base_steps: int, label: str) -> int:
        nonlocal base_done
        start_s = base_done * derived.dt_s
        length_s = base_steps * derived.dt_s
        steps = cap.plan(grid, start_s, length_s, base_steps, label).steps
        for index in range(steps):
            if steps == base_steps:  # the cap did not bind: exactly the uncapped loop's times
                stepper.step((base_done + index) * derived.dt_s, derived.dt_s)
            else:
                stepper.step(start_s + index * length_s / steps, length_s / steps)
        base_done += base_steps
        return steps

    warmup_steps = 0
    remaining = derived.warmup_steps
    while remaining > 0:
        block = min(remaining, derived.steps_per_frame)
        warmup_steps += run_block(block, "warmup")
        remaining -= block
    frame_steps = []
    for frame in range(recipe.time.frames):
        frame_steps.append(run_block(derived.steps_per_frame, f"frame {frame}"))
        yield frame, grid
    return StepLog(warmup_steps, tuple(frame_steps), derived.steps_per_frame, derived.warmup_steps)
