"""A batched ensemble of independent smoke instances advanced by one captured HIP graph per substep pair.

The reference runs one grid per process (sim/run.py:57-58); this is an addition with no reference counterpart.
Every member is an ordinary grid + Stepper (each with its own recipe: grid size, emitter, buoyancy,
confinement), so each member's step is exactly the reference's; what the ensemble changes is how the launches
reach the GPU:

- after the emitting frames, two substeps of every member are captured once into one HIP graph (through
  torch.cuda.CUDAGraph) and replayed: one launch from the host per two substeps of the whole ensemble,
  instead of ~100 host-side launches per member and substep;
- the pressure solves of the members up to `batch_cells` cells run as one batched multigrid
  (plm_mg_batch, multigrid.BatchedMultigrid): each stage of a V-cycle is one launch for all of them, persistent
  blocks walking the members' tiles (one workgroup per member for the LDS tails), with tile shapes from the
  batch's own autotuner. Those members' steps split around it: up to the divergence on the side streams, the
  batch on the capture stream, the gradient and the rest on the side streams again. Larger members step whole
  on their side streams, beside the batch (their solves in tolerance mode with a fixed number of device-gated
  V-cycles each);
- members are spread over side streams inside the graph (8 by default, the largest members first onto the
  least loaded stream by cells), so the graph's branches run concurrently;
- nothing in the graph reads back to the host: every tolerance-mode solve, batched or not, runs a fixed budget
  of gated cycles (the most its members needed lately, plus one) and adds 1 to its member's device counter when
  it ends with the residual still above tolerance. The frame loop reads the counters once per replayed frame
  (beside the CFL plan's read). A frame in which any solve stayed open is undone from a snapshot of the fields
  taken at its start and run again eagerly, where every solve continues to its target, and the graph is captured
  again with more cycles. A replayed frame therefore equals the eager frame bit for bit.

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
from plume_amd.sim.multigrid import BatchedMultigrid, Multigrid
from plume_amd.sim.pressure import close_faces_and_divergence, subtract_gradient
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
    def __init__(self, tables: list[dict], device: str, streams: int = 8, batch_cells: int | None = 1_500_000):
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
        # the members whose pressure solves run as one batched multigrid inside the graph (tolerance-mode
        # multigrid with common sweeps and tolerance, at most `batch_cells` cells: None batches every size), and
        # that batch (built on the first capture). Larger members step whole on the side streams, beside it.
        self.batch_cells = batch_cells
        self.batched = self._batchable()
        self.batch: BatchedMultigrid | None = None
        self.batch_tiles: list[tuple[int, list[float]]] = []  # the batch autotuner's shape and times per depth

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

    def _batchable(self) -> list[int]:
        solves = [m.stepper.pressure_solve for m in self.members]
        chosen = [i for i, solve in enumerate(solves) if isinstance(solve, Multigrid) and solve.tolerance > 0.0
                  and (self.batch_cells is None or self.members[i].cells <= self.batch_cells)]
        keys = {(solves[i].sweeps, solves[i].tolerance) for i in chosen}
        if len(keys) != 1 or next(iter(keys))[0] % 2 != 0:
            return []
        return chosen

    def step_eager(self, time_s: float, dt: float) -> None:
        for m in self.members:
            m.stepper.step(time_s, dt)

    def capture(self, time_s: float, dt: float, extra_cycles: int = 1) -> torch.cuda.CUDAGraph:
        """Captures two substeps of length `dt` of every member, from a time at which no emitter is active (the
        captured launches do not depend on the time then). Call after at least one eager frame, which sets each
        solve's cycle count and allocates every scratch array.

        Members whose solves batch (self.batched) run each substep in three parts: their steps up to the
        projection's divergence on the side streams, one batched multigrid solve of them all on the capture stream
        (every stage of a V-cycle one launch for all members), then the gradient and the rest of their steps on
        the side streams again. The batch runs as many cycles as the most any of them needed lately plus the
        margin, each member stopping on its own gate. Other members step whole on the side streams."""
        if self.emitting(time_s, time_s + 2.0 * dt):
            raise ValueError("an emitter is active: its launches depend on the time and cannot be replayed")
        if self.turbulent:
            raise ValueError("turbulence is on: its pattern depends on the time and cannot be replayed")
        solves = []
        budgets: list[int | None] = []
        for index, m in enumerate(self.members):
            solve = m.stepper.pressure_solve
            budget = None
            if isinstance(solve, Multigrid) and solve.tolerance > 0.0:
                solve.settle()
                budget = max(1, solve.recent_max + extra_cycles)
                if index not in self.batched:
                    solve.graph_cycles = budget
                solves.append(solve)
            budgets.append(budget)
        batch_cycles = max((budgets[i] for i in self.batched), default=0)
        for i in self.batched:
            budgets[i] = batch_cycles
        if self.batched and self.batch is None:
            self.batch = BatchedMultigrid([self.members[i].stepper.pressure_solve for i in self.batched],
                                          [self.members[i].grid for i in self.batched])
            self.batch_tiles = self.batch.autotune()
        torch.cuda.synchronize(self.device)
        order = sorted(range(len(self.members)), key=lambda i: -self.members[i].cells)
        graph = torch.cuda.CUDAGraph()

        def note(index: int) -> None:
            solve = self.members[index].stepper.pressure_solve
            if solve in solves:  # its gate: 1 while above tolerance
                count = self.open_solves[index:index + 1]
                count.add_(solve.norm_scratch[1:2].ne(0.0).to(torch.int32))
                most = self.frame_cycles[index:index + 1]
                torch.maximum(most, solve.cycles_run[0:1], out=most)

        try:
            with torch.cuda.graph(graph):
                main = torch.cuda.current_stream()
                for substep in range(2):
                    t = time_s + substep * dt
                    for s in self.streams:
                        s.wait_stream(main)
                    for index in order:
                        m = self.members[index]
                        with torch.cuda.stream(self.streams[self.stream_of[index]]):
                            if index in self.batched:
                                m.stepper.before_projection(t, dt)
                                close_faces_and_divergence(m.grid)
                            else:
                                m.stepper.step(t, dt)
                                note(index)
                    if self.batched:
                        own = [self.streams[k] for k in sorted({self.stream_of[i] for i in self.batched})]
                        for s in own:
                            main.wait_stream(s)
                        self.batch(batch_cycles)
                        for index in self.batched:
                            note(index)
                        for s in own:
                            s.wait_stream(main)
                        for index in order:
                            if index in self.batched:
                                m = self.members[index]
                                with torch.cuda.stream(self.streams[self.stream_of[index]]):
                                    subtract_gradient(m.grid)
                                    m.stepper.after_projection(dt)
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
                for index, (m, count, budget) in enumerate(zip(self.members, most, self.budgets[dt])):
                    if budget is not None:
                        m.stepper.pressure_solve.note_cycles(count)
                        if index not in self.batched:
                            stale = stale or budget > m.stepper.pressure_solve.recent_max + self.extra_cycles
                if self.batched:  # the batch's budget follows the member that needs the most cycles
                    need = max(self.members[i].stepper.pressure_solve.recent_max for i in self.batched)
                    stale = stale or self.budgets[dt][self.batched[0]] > need + self.extra_cycles
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

