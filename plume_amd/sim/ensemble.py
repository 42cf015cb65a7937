"""A batched ensemble of independent smoke instances advanced by one captured HIP graph per substep pair.

The reference runs one grid per process (sim/run.py:57-58); this is an addition with no reference counterpart.
Every member is an ordinary grid + Stepper (each with its own recipe: grid size, emitter, buoyancy,
confinement), so each member's step is exactly the reference's; what the ensemble changes is how the launches
reach the GPU:

- after the emitting frames, two substeps of every member are captured once into one HIP graph (through
  torch.cuda.CUDAGraph) and replayed: one launch from the host per two substeps of the whole ensemble,
  instead of ~100 host-side launches per member and substep;
- members are spread over side streams inside the graph (8 by default, the largest members first onto the
  least loaded stream by cells), so the graph's branches run concurrently: the members' latency-bound kernels
  (coarse multigrid levels, the one-workgroup tails, reductions) overlap with each other and with the large
  members' bandwidth-bound ones;
- the pressure solves run their tolerance mode with a fixed number of device-gated V-cycles (the most any
  member needed in the eager frames, plus one), so nothing in the graph reads back to the host. Each solve in
  the graph adds 1 to its member's device counter when it ends with the residual still above tolerance; the
  frame loop reads the counters once per replayed frame (beside the CFL plan's read). A frame in which any
  solve stayed open is undone from a snapshot of the fields taken at its start and run again eagerly, where
  every solve continues to its target, and the graph for that step length is captured again with more cycles.
  A replayed frame therefore equals the eager frame bit for bit.

Two substeps are captured because AdvectedField.swap exchanges buffers by reference: after two substeps every
field is back in the buffer the graph was captured with. The substep count of a frame is therefore common to
all members (the largest any member's CFL cap asks for) and rounded up to an even number; more, shorter
substeps keep every member within its cap. Frames in which an emitter is active or any member has turbulence
(their launches depend on the time, which a replay cannot change) run eagerly, member by member.
"""

from dataclasses import dataclass

import numpy as np
import torch

from plume_amd.recipe.derived import derive
from plume_amd.recipe.load import fill_dependent_defaults, parse_recipe
from plume_amd.sim.grid import make_grid
from plume_amd.sim.multigrid import Multigrid
from plume_amd.sim.scenes import puff_table
from plume_amd.sim.step_cap import StepCap
from plume_amd.sim.stepper import Stepper

ENSEMBLE_SIZES = ((48, 48, 48), (64, 96, 64), (96, 128, 96), (128, 192, 128))


@dataclass
class Member:
    recipe: object
    derived: object
    grid: object
    stepper: Stepper
    cap: StepCap

    @property
    def cells(self) -> int:
        return int(np.prod(self.grid.cells))


def ensemble_tables(count: int, seed: int, frames: int = 48, sizes=ENSEMBLE_SIZES) -> list[dict]:
    """`count` puff recipes drawn from `seed`: a grid size from `sizes`, an emitter radius of 4-12 cells,
    buoyancy alpha in [0.05, 0.2] and confinement epsilon in [0, 0.5] each."""
    rng = np.random.default_rng(seed)
    tables = []
    for index in range(count):
        cells = sizes[int(rng.integers(len(sizes)))]
        radius = float(rng.uniform(4.0, 12.0))
        alpha = float(rng.uniform(0.05, 0.2))
        epsilon = float(rng.uniform(0.0, 0.5))
        tables.append(puff_table(cells, frames=frames, radius_cells=radius, alpha=alpha, confinement=epsilon,
                                 seed=seed + index))
    return tables


def balanced_streams(cells: list[int], streams: int) -> list[int]:
    """A stream per member, the largest members first onto the least loaded stream (longest processing time
    first, by cells), so the branches of a replayed graph finish close together."""
    load = [0] * streams
    out = [0] * len(cells)
    for index in sorted(range(len(cells)), key=lambda i: -cells[i]):
        k = min(range(streams), key=lambda s: (load[s], s))
        out[index] = k
        load[k] += cells[index]
    return out


class Ensemble:
    def __init__(self, tables: list[dict], device: str, streams: int = 8):
        self.device = device
        self.members = []
        for table in tables:
            recipe = fill_dependent_defaults(parse_recipe(table), "ensemble")
            derived = derive(recipe)
            grid = make_grid(derived, recipe.domain.cell_m, recipe.domain.faces, recipe.domain.ambient_k, device)
            self.members.append(Member(recipe, derived, grid, Stepper(grid, recipe), StepCap(recipe)))
        frame_dt = {round(m.derived.dt_s * m.derived.steps_per_frame, 12) for m in self.members}
        base = {m.derived.steps_per_frame for m in self.members}
        if len(frame_dt) != 1 or len(base) != 1:
            raise ValueError("ensemble members must share the frame rate and the substeps per frame")
        self.frame_s = frame_dt.pop()
        self.base_steps = base.pop()
        self.streams = [torch.cuda.Stream(device=device) for _ in range(max(1, streams))]
        self.stream_of = balanced_streams([m.cells for m in self.members], len(self.streams))
        self.graphs: dict[float, torch.cuda.CUDAGraph] = {}
        self.frames_done = 0
        self.extra_cycles = 1  # device-gated V-cycles captured beyond the most any eager solve needed
        self.log: list[tuple[int, int, bool]] = []  # (frame, substeps, replayed)
        self.redone: list[int] = []  # frames replayed with an open gate, then run again eagerly
        # per member: solves in replayed graphs that ended above tolerance (device, cleared each frame)
        self.open_solves = torch.zeros(len(self.members), dtype=torch.int32, device=device)
        # per member: the most cycles any solve ran in the replays of the current frame (device)
        self.frame_cycles = torch.zeros(len(self.members), dtype=torch.int32, device=device)
        self.budgets: dict[float, list[int | None]] = {}  # per graph: each member's captured cycle budget
        self._snapshot: list[list[torch.Tensor]] | None = None

    @property
    def cells(self) -> int:
        return sum(m.cells for m in self.members)

    def autotune(self) -> None:
        """Picks the sweep tile shape of every distinct level size once (process-wide)."""
        seen = set()
        for m in self.members:
            solve = m.stepper.pressure_solve
            if isinstance(solve, Multigrid) and m.grid.cells not in seen:
                seen.add(m.grid.cells)
                solve.autotune()

    def frame_steps(self, start_s: float) -> int:
        """Substeps of the frame from `start_s` for every member: the largest any member's CFL cap asks for
        (sim/step_cap.py:38-48), rounded up to an even count."""
        steps = max(m.cap.plan(m.grid, start_s, self.frame_s, self.base_steps, f"frame {self.frames_done}").steps
                    for m in self.members)
        return steps + (steps % 2)

    def emitting(self, start_s: float, end_s: float) -> bool:
        return any(np.float32(p.start_s) < np.float32(end_s) and np.float32(p.end_s) > np.float32(start_s)
                   for m in self.members for p in m.stepper.emitters.params)

    @property
    def turbulent(self) -> bool:
        """Some member's turbulence is on: its noise pattern moves with the time (sim/turbulence.py:59-67),
        which a replayed graph would freeze at the capture's time."""
        return any(m.recipe.forces.turbulence.strength_mps2 != 0.0 for m in self.members)

    def _state(self, m: Member) -> list[torch.Tensor]:
        """What a member carries from one frame to the next: the fields' data and the pressure (the warm
        start). Everything else is rewritten every substep."""
        return [f.data for f in m.grid.velocity_fields()] + [m.grid.smoke.data, m.grid.temperature.data,
                                                              m.grid.pressure]

    def _take_snapshot(self) -> None:
        if self._snapshot is None:
            self._snapshot = [[torch.empty_like(t) for t in self._state(m)] for m in self.members]
        for m, saved in zip(self.members, self._snapshot):
            for dst, src in zip(saved, self._state(m)):
                dst.copy_(src)

    def _restore_snapshot(self) -> None:
        for m, saved in zip(self.members, self._snapshot):
            for dst, src in zip(self._state(m), saved):
                dst.copy_(src)

    def step_eager(self, time_s: float, dt: float) -> None:
        for m in self.members:
            m.stepper.step(time_s, dt)

    def capture(self, time_s: float, dt: float, extra_cycles: int = 1) -> torch.cuda.CUDAGraph:
        """Captures two substeps of length `dt` of every member, from a time at which no emitter is active (the
        captured launches do not depend on the time then). Call after at least one eager frame, which sets each
        solve's cycle count and allocates every scratch array."""
        if self.emitting(time_s, time_s + 2.0 * dt):
            raise ValueError("an emitter is active: its launches depend on the time and cannot be replayed")
        if self.turbulent:
            raise ValueError("turbulence is on: its pattern depends on the time and cannot be replayed")
        solves = []
        budgets: list[int | None] = []
        for m in self.members:
            solve = m.stepper.pressure_solve
            budget = None
            if isinstance(solve, Multigrid) and solve.tolerance > 0.0:
                solve.settle()
                budget = solve.graph_cycles = max(1, solve.recent_max + extra_cycles)
                solves.append(solve)
            budgets.append(budget)
        torch.cuda.synchronize(self.device)
        graph = torch.cuda.CUDAGraph()
        try:
            with torch.cuda.graph(graph):
                main = torch.cuda.current_stream()
                for s in self.streams:
                    s.wait_stream(main)
                for index in sorted(range(len(self.members)), key=lambda i: -self.members[i].cells):
                    m = self.members[index]
                    with torch.cuda.stream(self.streams[self.stream_of[index]]):
                        for substep in range(2):
                            m.stepper.step(time_s + substep * dt, dt)
                            solve = m.stepper.pressure_solve
                            if solve in solves:  # its gate: 1 while above tolerance (plm_vcycle)
                                count = self.open_solves[index:index + 1]
                                count.add_(solve.norm_scratch[1:2].ne(0.0).to(torch.int32))
                                most = self.frame_cycles[index:index + 1]
                                torch.maximum(most, solve.cycles_run[0:1], out=most)
                for s in self.streams:
                    main.wait_stream(s)
        finally:
            for solve in solves:
                solve.graph_cycles = None
        self.graphs[dt] = graph
        self.budgets[dt] = budgets
        return graph

    def run_frame(self, replay: bool = True) -> int:
        """Advances every member one frame; returns the substeps taken."""
        start_s = self.frames_done * self.frame_s
        steps = self.frame_steps(start_s)
        dt = self.frame_s / steps
        quiet = not self.emitting(start_s, start_s + self.frame_s) and not self.turbulent
        replayed = False
        if replay and quiet and self.frames_done > 0:
            graph = self.graphs.get(dt) or self.capture(start_s, dt, self.extra_cycles)
            self._take_snapshot()
            self.open_solves.zero_()
            self.frame_cycles.zero_()
            for _ in range(steps // 2):
                graph.replay()
            replayed = True
            opened, most = torch.stack([self.open_solves, self.frame_cycles]).tolist()
            if not any(opened):
                # the cycles the replayed solves ran set the members' predictions; a budget two or more cycles
                # above what a member now needs (plus the margin) wastes gated launches: capture again
                stale = False
                for m, count, budget in zip(self.members, most, self.budgets[dt]):
                    if budget is not None:
                        m.stepper.pressure_solve.note_cycles(count)
                        stale = stale or budget > m.stepper.pressure_solve.recent_max + self.extra_cycles
                if stale:
                    self.graphs.pop(dt, None)
            if any(opened):
                # some solve ended above tolerance: undo the frame, run it eagerly (each solve continues to its
                # target) and give the next capture the cycle counts the eager solves needed
                self._restore_snapshot()
                for index in range(steps):
                    self.step_eager(start_s + index * dt, dt)
                self.graphs.pop(dt, None)
                self.redone.append(self.frames_done)
                replayed = False
        else:
            for index in range(steps):
                self.step_eager(start_s + index * dt, dt)
        self.log.append((self.frames_done, steps, replayed))
        self.frames_done += 1
        return steps


def run_alone(table: dict, schedule: list[int], device: str):
    """One ensemble member's recipe stepped on its own, eagerly, with the ensemble's substep counts per frame
    (for checking that batching changes nothing)."""
    recipe = fill_dependent_defaults(parse_recipe(table), "ensemble")
    derived = derive(recipe)
    grid = make_grid(derived, recipe.domain.cell_m, recipe.domain.faces, recipe.domain.ambient_k, device)
    stepper = Stepper(grid, recipe)
    frame_s = derived.dt_s * derived.steps_per_frame
    for frame, steps in enumerate(schedule):
        dt = frame_s / steps
        for index in range(steps):
            stepper.step(frame * frame_s + index * dt, dt)
    return grid

