"""The frame loop (the interface of sim/run.py:22-85): warmup, then each frame's steps followed by a capture.

The run is a schedule of blocks, each `base` steps of the recipe's step length long: the warmup in blocks of
one frame's steps (the last one shorter when the warmup is not a whole number of frames), then one block per
frame. A block's start time is the number of base steps before it times dt (never a running sum, so the clock
does not drift). At the start of each block the CFL cap measures the fastest gas and may ask for more steps
than `base`; the block is then split into that many equal steps over the same interval. Frame i is yielded once
its block is done, i.e. after warmup_s + (i + 1) / fps of simulated time.
"""

from collections.abc import Callable, Generator, Iterator
from dataclasses import dataclass

from plume_amd.sim.grid import Grid, make_grid
from plume_amd.sim.step_cap import StepCap
from plume_amd.sim.stepper import Stepper

Capture = Callable[[int, Grid], None]


@dataclass(frozen=True)
class StepLog:
    """Steps taken by the warmup and by each frame once the CFL cap has had its say, beside the recipe's base
    counts (sim/run.py:22-39)."""

    warmup_steps: int
    steps_per_frame: tuple[int, ...]
    base_steps_per_frame: int
    base_warmup_steps: int

    @property
    def total_steps(self) -> int:
        return self.warmup_steps + sum(self.steps_per_frame)

    @property
    def added_steps(self) -> int:
        """How many steps the cap added to the recipe's own, warmup included."""
        return self.total_steps - (self.base_warmup_steps + self.base_steps_per_frame * len(self.steps_per_frame))


@dataclass(frozen=True)
class _Block:
    label: str
    base: int  # steps of the recipe's dt the block stands for
    frame: int | None  # the frame captured after it; None for warmup


def _schedule(warmup_steps: int, steps_per_frame: int, frames: int) -> Iterator[_Block]:
    whole, rest = divmod(warmup_steps, steps_per_frame)
    for _ in range(whole):
        yield _Block("warmup", steps_per_frame, None)
    if rest:
        yield _Block("warmup", rest, None)
    for frame in range(frames):
        yield _Block(f"frame {frame}", steps_per_frame, frame)


def _step_times(base_before: int, base: int, steps: int, dt: float) -> list[tuple[float, float]]:
    """(start time, length) of each step of a block that begins after `base_before` base steps. Unsplit, step
    n starts at (base_before + n) dt exactly as a loop without the cap would; split, the block's length
    base * dt is shared by `steps` equal steps."""
    if steps == base:
        return [((base_before + n) * dt, dt) for n in range(steps)]
    start, length = base_before * dt, base * dt
    return [(start + n * length / steps, length / steps) for n in range(steps)]


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
    """The callback form of `simulate_frames`: `capture(frame_index, grid)` after each frame; returns the step
    log (sim/run.py:42-50)."""
    frames = simulate_frames(recipe, derived, device)
    try:
        while True:
            capture(*next(frames))
    except StopIteration as done:
        return done.value


def simulate_frames(recipe, derived, device: str) -> Generator[tuple[int, Grid], None, StepLog]:
    """Yields (frame_index, live grid) after each frame's block and returns the StepLog (sim/run.py:53-85)."""
    grid = make_grid(derived, recipe.domain.cell_m, recipe.domain.faces, recipe.domain.ambient_k, device)
    stepper = Stepper(grid, recipe)
    tune_sweep_tiles(stepper.pressure_solve, derived.cells)
    cap = StepCap(recipe)
    dt = derived.dt_s
    base_before = 0
    warmup_taken, frame_taken = 0, []
    for block in _schedule(derived.warmup_steps, derived.steps_per_frame, recipe.time.frames):
        steps = cap.plan(grid, base_before * dt, block.base * dt, block.base, block.label).steps
        for time_s, length_s in _step_times(base_before, block.base, steps, dt):
            stepper.step(time_s, length_s)
        base_before += block.base
        if block.frame is None:
            warmup_taken += steps
            continue
        frame_taken.append(steps)
        yield block.frame, grid
    return StepLog(warmup_taken, tuple(frame_taken), derived.steps_per_frame, derived.warmup_steps)
