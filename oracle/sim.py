"""The oracle's simulation loop: the reference's orchestration of its kernels, over numpy arrays.

TEST INFRASTRUCTURE ONLY (see `oracle/__init__.py`). Follows, in order and in arithmetic:
sim/grid.py (make_grid, AdvectedField), sim/advect.py:59-68, sim/forces.py, sim/vorticity.py:87-117,
sim/emitters.py:124-153, sim/turbulence.py:59-67, sim/pressure.py:124-162, sim/multigrid.py:80-147,
sim/reduce.py:28-38, sim/conserve.py:33-52, sim/dissipate.py:18-24, sim/step_cap.py, sim/stepper.py,
sim/run.py. Recipes are read by attribute (recipe.domain.cell_m, ...), as the reference's `Recipe`.
"""

import math
from dataclasses import dataclass

import numpy as np

from oracle import Oracle

CENTER = (0.5, 0.5, 0.5)
U_FACE = (0.0, 0.5, 0.5)
V_FACE = (0.5, 0.0, 0.5)
W_FACE = (0.5, 0.5, 0.0)
FACE_OFFSETS = (U_FACE, V_FACE, W_FACE)

MIN_LEVEL_CELLS = 2
MAX_LEVELS = 8
COARSEST_SWEEPS = 20
COARSEST_SWEEPS_PER_CELL = 4
RELAXATION = 1.5


class Field:
    """AdvectedField, sim/grid.py:36-51."""

    def __init__(self, shape, offset, outside, uses_outside, dtype, initial=0.0):
        self.offset = offset
        self.outside = outside
        self.uses_outside = 1 if uses_outside else 0
        self.data = np.full(shape, initial, dtype=dtype)
        self.hat = np.zeros(shape, dtype=dtype)
        self.back = np.zeros(shape, dtype=dtype)
        self.next = np.zeros(shape, dtype=dtype)

    def swap(self):
        self.data, self.next = self.next, self.data


class Grid:
    """Grid and make_grid, sim/grid.py:54-98. `open_lo`/`open_hi` are 0/1 per axis."""

    def __init__(self, o: Oracle, cells, cell_m, origin_m, open_lo, open_hi, ambient_k):
        nx, ny, nz = cells
        dt = o.dtype
        self.o = o
        self.cells = tuple(cells)
        self.cell_m = cell_m
        self.origin_m = tuple(origin_m)
        self.open_lo = tuple(open_lo)
        self.open_hi = tuple(open_hi)
        self.info = o.info(cells, cell_m, origin_m, open_lo, open_hi)
        self.u = Field((nx + 1, ny, nz), U_FACE, 0.0, False, dt)
        self.v = Field((nx, ny + 1, nz), V_FACE, 0.0, False, dt)
        self.w = Field((nx, ny, nz + 1), W_FACE, 0.0, False, dt)
        self.smoke = Field((nx, ny, nz), CENTER, 0.0, True, dt)
        self.temperature = Field((nx, ny, nz), CENTER, ambient_k, True, dt, initial=ambient_k)
        self.pressure = np.zeros((nx, ny, nz), dtype=dt)
        self.divergence = np.zeros((nx, ny, nz), dtype=dt)

    def velocity_fields(self):
        return self.u, self.v, self.w


def face_flags(faces):
    """open_lo, open_hi from a FaceSettings-like object (sim/grid.py:78-79)."""
    is_open = lambda kind: 1 if kind == "open" else 0  # noqa: E731
    return ((is_open(faces.left), is_open(faces.bottom), is_open(faces.back)),
            (is_open(faces.right), is_open(faces.top), is_open(faces.front)))


def make_grid(o: Oracle, derived, cell_m, faces, ambient_k) -> Grid:
    open_lo, open_hi = face_flags(faces)
    return Grid(o, derived.cells, cell_m, derived.domain_min_m, open_lo, open_hi, ambient_k)


# ---------------------------------------------------------------------- operators

def advect_maccormack(field: Field, velocity, grid: Grid, dt: float) -> None:
    """sim/advect.py:59-68."""
    u, v, w = velocity
    grid.o.advect_maccormack(grid.info, field.data, field.hat, field.back, field.next, field.offset,
                             field.outside, field.uses_outside, u, v, w, dt)


def advect_semi_lagrangian(field: Field, velocity, grid: Grid, dt: float) -> None:
    """sim/advect.py:53-56."""
    u, v, w = velocity
    grid.o.advect_sl(grid.info, field.data, field.next, field.offset, field.outside, field.uses_outside, u, v, w,
                     dt / grid.info.cell_m)


def apply_buoyancy(grid: Grid, gravity, heat_buoyancy, smoke_weight, ambient_k, dt) -> None:
    grid.o.buoyancy(grid.info, grid.v.data, grid.smoke.data, grid.temperature.data, gravity, heat_buoyancy,
                    smoke_weight, ambient_k, dt)


def apply_drag(grid: Grid, wind_mps, drag_per_s, dt) -> None:
    """sim/forces.py:40-48."""
    if drag_per_s == 0.0:
        return
    keep = math.exp(-drag_per_s * dt)
    for axis, field in enumerate(grid.velocity_fields()):
        grid.o.drag(field.data, float(wind_mps[axis]), keep)


class VorticityConfinement:
    """sim/vorticity.py:87-117."""

    def __init__(self, grid: Grid):
        shape = grid.cells + (3,)
        self.center = np.zeros(shape, dtype=grid.o.dtype)
        self.curl = np.zeros(shape, dtype=grid.o.dtype)
        self.push = np.zeros(shape, dtype=grid.o.dtype)

    def apply(self, grid: Grid, strength: float, dt: float) -> None:
        if strength == 0.0:
            return
        o = grid.o
        u, v, w = (f.data for f in grid.velocity_fields())
        o.center_velocity(grid.info, u, v, w, self.center)
        o.curl(grid.info, self.center, self.curl)
        o.confinement_push(grid.info, self.curl, self.push, strength)
        for axis, component in enumerate((u, v, w)):
            o.add_face_push(grid.info, component, self.push, axis, dt)

    def measure_curl(self, grid: Grid):
        u, v, w = (f.data for f in grid.velocity_fields())
        grid.o.center_velocity(grid.info, u, v, w, self.center)
        grid.o.curl(grid.info, self.center, self.curl)
        return self.curl


def emitter_params(emitters) -> list[dict]:
    """sim/emitters.py:124-144: each emitter's struct fields, stream = its index."""
    params = []
    for index, e in enumerate(emitters):
        params.append(dict(
            shape=0 if e.shape == "sphere" else 1, center_m=tuple(e.center_m), radius_m=e.radius_m,
            half_size_m=tuple(s / 2 for s in e.size_m), edge_softness=e.edge_softness, start_s=e.start_s,
            end_s=e.end_s, smoke_per_s=e.smoke_per_s, temperature_k=e.temperature_k,
            velocity_mps=tuple(e.velocity_mps), outward=1 if e.velocity_mode == "outward" else 0,
            noise_scale_m=e.noise.scale_m, noise_strength=e.noise.strength, stream=index))
    return params


def emit(grid: Grid, params: list[dict], seed: int, time_s: float, dt: float) -> None:
    """sim/emitters.py:147-153."""
    o = grid.o
    array = o.emitters(params)
    o.emit_scalars(grid.info, grid.smoke.data, grid.temperature.data, array, len(params), seed, time_s, dt)
    for axis, field in enumerate(grid.velocity_fields()):
        o.emit_velocity(grid.info, field.data, axis, array, len(params), seed, time_s, dt)


def apply_turbulence(grid: Grid, seed, scale_m, strength_mps2, speed_mps, time_s, dt) -> None:
    """sim/turbulence.py:59-67."""
    if strength_mps2 == 0.0:
        return
    pattern_time = time_s * speed_mps / scale_m
    for axis, field in enumerate(grid.velocity_fields()):
        grid.o.turbulence(grid.info, field.data, axis, seed, scale_m, strength_mps2, pattern_time, dt)


def dissipate(grid: Grid, smoke_per_s, cooling_per_s, ambient_k, dt) -> None:
    """sim/dissipate.py:18-24."""
    if smoke_per_s == 0.0 and cooling_per_s == 0.0:
        return
    grid.o.dissipate(grid.smoke.data, grid.temperature.data, math.exp(-smoke_per_s * dt),
                     math.exp(-cooling_per_s * dt), ambient_k)


# ---------------------------------------------------------------------- pressure

def compute_divergence(grid: Grid):
    u, v, w = (f.data for f in grid.velocity_fields())
    grid.o.divergence(grid.info, u, v, w, grid.divergence)
    return grid.divergence


def project(grid: Grid, solve) -> None:
    """sim/pressure.py:144-154."""
    for axis, field in enumerate(grid.velocity_fields()):
        grid.o.zero_closed_faces(grid.info, field.data, axis)
    compute_divergence(grid)
    solve(grid)
    for axis, field in enumerate(grid.velocity_fields()):
        grid.o.subtract_gradient(grid.info, field.data, grid.pressure, axis)


class GaussSeidel:
    """sim/pressure.py:124-132."""

    def __init__(self, iterations: int):
        self.iterations = iterations

    def __call__(self, grid: Grid) -> None:
        grid.o.rb_sweeps(grid.info, grid.pressure, grid.divergence, self.iterations, RELAXATION)


def coarse_cells(cells):
    return tuple((n + 1) // 2 for n in cells)


def level_infos(o: Oracle, cells, cell_m, origin_m, open_lo, open_hi, max_levels=MAX_LEVELS):
    """The level list of Multigrid._build_levels, sim/multigrid.py:84-122: (cells, info) per level."""
    levels = [(tuple(cells), o.info(cells, cell_m, origin_m, open_lo, open_hi))]
    level_cell_m = float(o.real(cell_m).value)
    while min(levels[-1][0]) > MIN_LEVEL_CELLS and len(levels) < max_levels:
        index = len(levels)
        coarse = coarse_cells(levels[-1][0])
        level_cell_m = float(o.real(level_cell_m * 2.0).value)
        scale = 2.0 ** index
        fill = [top_n / scale - (n - 1) for top_n, n in zip(cells, coarse)]
        levels.append((coarse, o.info(coarse, level_cell_m, origin_m, open_lo, open_hi, fill, 0.5 / scale)))
    return levels


class Multigrid:
    """sim/multigrid.py:97-147 (the cycle runs in C, orc_multigrid).

    tolerance > 0 is the build's converge-to-a-residual mode (the recipe keys multigrid_tolerance and
    multigrid_max_cycles, which the reference lacks): cycles run one at a time, from none, while the FP64
    relative residual |rhs - A q| / |rhs| of the top level is above `tolerance`, at most `max_cycles`. Each
    cycle is the reference's V-cycle, so n single cycles equal one call of n cycles. The cycle counts of the
    solves are kept in `cycles_log`."""

    def __init__(self, grid: Grid, cycles: int, sweeps: int, tolerance: float = 0.0, max_cycles: int = 16):
        self.cycles = cycles
        self.sweeps = sweeps
        self.tolerance = float(tolerance)
        self.max_cycles = int(max_cycles)
        self.cycles_log: list[int] = []
        self.levels = level_infos(grid.o, grid.cells, grid.cell_m, grid.origin_m, grid.open_lo, grid.open_hi)
        dt = grid.o.dtype
        self.q = [None] + [np.zeros(c, dtype=dt) for c, _ in self.levels[1:]]
        self.rhs = [None] + [np.zeros(c, dtype=dt) for c, _ in self.levels[1:]]
        self.residual = [np.zeros(c, dtype=dt) for c, _ in self.levels]

    def _run(self, grid: Grid, cycles: int) -> None:
        grid.o.multigrid([info for _, info in self.levels], self.q, self.rhs, self.residual, cycles, self.sweeps)

    def __call__(self, grid: Grid) -> None:
        self.q[0] = grid.pressure
        self.rhs[0] = grid.divergence
        if self.tolerance <= 0.0:
            self._run(grid, self.cycles)
            return
        ran = 0
        while ran < self.max_cycles and relative_residual(grid) > self.tolerance:
            self._run(grid, 1)
            ran += 1
        self.cycles_log.append(ran)


def relative_residual(grid: Grid) -> float:
    """|rhs - A q| / |rhs| of the top level in FP64 (the residual itself in the oracle's precision); a zero
    rhs gives 0 when q solves it and inf otherwise."""
    r = np.zeros_like(grid.pressure)
    grid.o.residual(grid.info, grid.pressure, grid.divergence, r)
    r2 = float(np.dot(r.ravel().astype(np.float64), r.ravel().astype(np.float64)))
    b = grid.divergence.ravel().astype(np.float64)
    b2 = float(np.dot(b, b))
    if b2 > 0.0:
        return math.sqrt(r2) / math.sqrt(b2)
    return math.inf if r2 > 0.0 else 0.0


def poisson_residual(grid: Grid, q=None, rhs=None) -> float:
    """‖rhs − A q‖ / ‖rhs‖ on the top level (tests/sim/test_multigrid.py:46-49 of the reference)."""
    q = grid.pressure if q is None else q
    rhs = grid.divergence if rhs is None else rhs
    r = np.zeros_like(q)
    grid.o.residual(grid.info, q, rhs, r)
    return float(np.linalg.norm(r.astype(np.float64)) / np.linalg.norm(rhs.astype(np.float64)))


def make_pressure_solve(grid: Grid, solver):
    """sim/pressure_solver.py:9-13."""
    if solver.pressure_solver == "gauss_seidel":
        return GaussSeidel(solver.pressure_iterations)
    return Multigrid(grid, solver.multigrid_cycles, solver.multigrid_sweeps,
                     getattr(solver, "multigrid_tolerance", 0.0), getattr(solver, "multigrid_max_cycles", 16))


# ---------------------------------------------------------------------- budget

def held_total(before: float, semi_lagrangian: float, maccormack: float) -> float:
    """sim/conserve.py:33-35."""
    return max(semi_lagrangian, min(maccormack, before))


class SmokeBudget:
    """sim/conserve.py:38-52 with TripleSum, sim/reduce.py:28-38."""

    def __init__(self, grid: Grid):
        self.lines = np.zeros((3, grid.cells[1], grid.cells[2]), dtype=grid.o.dtype)
        self.last_factor = 1.0
        self.last_totals = (0.0, 0.0, 0.0)

    def totals(self, grid: Grid, a, b, c):
        grid.o.line_sums(a, b, c, self.lines)
        totals = self.lines.astype(np.float64).sum(axis=(1, 2))
        return float(totals[0]), float(totals[1]), float(totals[2])

    def hold(self, grid: Grid) -> None:
        smoke = grid.smoke
        before, semi_lagrangian, maccormack = self.totals(grid, smoke.data, smoke.hat, smoke.next)
        self.last_totals = (before, semi_lagrangian, maccormack)
        target = held_total(before, semi_lagrangian, maccormack)
        self.last_factor = target / maccormack if maccormack > 0.0 else 1.0
        if self.last_factor != 1.0:
            grid.o.scale(smoke.next, self.last_factor)


# ---------------------------------------------------------------------- loop

class Stepper:
    """sim/stepper.py:17-60."""

    def __init__(self, grid: Grid, recipe):
        self.grid = grid
        self.recipe = recipe
        self.emitters = emitter_params(recipe.emitter)
        self.pressure_solve = make_pressure_solve(grid, recipe.solver)
        self.smoke_budget = SmokeBudget(grid)
        self.confinement = VorticityConfinement(grid) if recipe.forces.vorticity_confinement > 0 else None

    def step(self, time_s: float, dt: float) -> None:
        grid, recipe = self.grid, self.recipe
        ambient = recipe.domain.ambient_k
        emit(grid, self.emitters, recipe.seed, time_s, dt)
        forces = recipe.forces
        apply_buoyancy(grid, forces.gravity_mps2, forces.heat_buoyancy, forces.smoke_weight, ambient, dt)
        if self.confinement is not None:
            self.confinement.apply(grid, forces.vorticity_confinement, dt)
        turbulence = forces.turbulence
        apply_turbulence(grid, recipe.seed, turbulence.scale_m, turbulence.strength_mps2, turbulence.speed_mps,
                         time_s, dt)
        apply_drag(grid, forces.wind_mps, forces.drag_per_s, dt)
        fields = grid.velocity_fields()
        velocity = tuple(f.data for f in fields)
        for f in fields:
            advect_maccormack(f, velocity, grid, dt)
        for f in fields:
            f.swap()
        project(grid, self.pressure_solve)
        velocity = tuple(f.data for f in grid.velocity_fields())
        advect_maccormack(grid.smoke, velocity, grid, dt)
        self.smoke_budget.hold(grid)
        grid.smoke.swap()
        advect_maccormack(grid.temperature, velocity, grid, dt)
        grid.temperature.swap()
        dissipate(grid, recipe.dissipation.smoke_per_s, recipe.dissipation.cooling_per_s, ambient, dt)


class StepCapError(Exception):
    pass


def _emitter_speed(velocity_mps, mode):
    velocity = np.asarray(velocity_mps, dtype=np.float64)
    if mode == "outward":
        return float(np.linalg.norm(velocity))
    return float(np.abs(velocity).max())


class StepCap:
    """sim/step_cap.py:29-59."""

    def __init__(self, recipe):
        self.cell_m = recipe.domain.cell_m
        self.max_cells = recipe.solver.max_cells_per_step
        self.max_substeps = recipe.solver.max_substeps
        self.emitters = [(e.start_s, e.end_s, _emitter_speed(e.velocity_mps, e.velocity_mode))
                         for e in recipe.emitter]

    def plan(self, grid: Grid, start_s, length_s, base_steps, label) -> int:
        speed = max(max(float(np.abs(f.data).max()) for f in grid.velocity_fields()),
                    max([s for b, e, s in self.emitters if b < start_s + length_s and e > start_s], default=0.0))
        needed = math.ceil(speed * length_s / (self.max_cells * self.cell_m))
        steps = max(base_steps, needed)
        if steps > max(self.max_substeps, base_steps):
            raise StepCapError(f"{label} needs {steps} steps; max_substeps is {self.max_substeps}")
        return steps


@dataclass(frozen=True)
class StepLog:
    warmup_steps: int
    steps_per_frame: tuple


def simulate_frames(o: Oracle, recipe, derived):
    """sim/run.py:53-85: yields (frame, grid) after each frame; returns the StepLog."""
    grid = make_grid(o, derived, recipe.domain.cell_m, recipe.domain.faces, recipe.domain.ambient_k)
    stepper = Stepper(grid, recipe)
    cap = StepCap(recipe)
    base_done = 0

    def run_block(base_steps, label):
        nonlocal base_done
        start_s = base_done * derived.dt_s
        length_s = base_steps * derived.dt_s
        steps = cap.plan(grid, start_s, length_s, base_steps, label)
        for index in range(steps):
            if steps == base_steps:
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
    return StepLog(warmup_steps, tuple(frame_steps))
