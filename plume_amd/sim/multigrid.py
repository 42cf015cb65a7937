"""Geometric multigrid for the pressure solve (sim/multigrid.py of the reference).

Same hierarchy as the reference: each level halves the cell count per axis (rounding up) and doubles the
cell size, down to 2 cells on the shortest axis or 8 levels; a coarse level whose axis overhangs the domain
weighs its last cell by the share inside (`coarse_geometry`). One V(sweeps, sweeps) cycle: pre-smooth,
residual, restrict, coarse q = 0, recurse, prolong-add, post-smooth; the coarsest level gets
max(20, 4 * longest) sweeps at 1.5. Warm start from `grid.pressure`.

The whole solve is one call into libplume_hip (plm_vcycle): fused two-colour sweeps, a fused
residual + restrict, and every level of at most a few thousand cells solved by one workgroup in LDS. With
`tolerance` > 0 the solve stops once |rhs - A q| / |rhs| (FP64, reduced on the device) is at most
`tolerance`, capped at `max_cycles` cycles, with no host round trip.
"""

import ctypes
from collections import deque
from dataclasses import dataclass

import torch

from plume_amd._lib import LIB, GridInfo, call, ptr, stream
from plume_amd.sim.grid import grid_info

MIN_LEVEL_CELLS = 2
MAX_LEVELS = 8
PLM_LDS_CELLS = 8192  # include/plume_hip.h: levels this small (together) are solved in one workgroup's LDS
COARSEST_SWEEPS = 20
COARSEST_SWEEPS_PER_CELL = 4
COARSEST_RELAXATION = 1.5
RECENT_SOLVES = 8  # the solves whose cycle counts set a replayed graph's budget
SMOOTHING_RELAXATION = 1.0


@dataclass
class Level:
    info: GridInfo
    cells: tuple[int, int, int]
    q: torch.Tensor
    rhs: torch.Tensor
    residual: torch.Tensor


def coarse_cells(cells) -> tuple[int, int, int]:
    return tuple((n + 1) // 2 for n in cells)


def coarse_geometry(top_cells, cells, level: int) -> tuple[tuple[float, float, float], float]:
    """fill_hi and ghost_gap of a level `level` steps down (sim/multigrid.py:84-94)."""
    scale = 2.0 ** level
    fill = tuple(top_n / scale - (n - 1) for top_n, n in zip(top_cells, cells))
    return fill, 0.5 / scale


def build_levels(grid, max_levels: int = MAX_LEVELS) -> list[Level]:
    """The level list of Multigrid._build_levels (sim/multigrid.py:106-122); level 0 aliases the grid's
    pressure and divergence."""
    device = grid.device
    top = Level(grid.info, tuple(grid.cells), grid.pressure, grid.divergence,
                torch.zeros(grid.cells, dtype=torch.float32, device=device))
    levels = [top]
    while min(levels[-1].cells) > MIN_LEVEL_CELLS and len(levels) < max_levels:
        above = levels[-1]
        cells = coarse_cells(above.cells)
        fill, ghost_gap = coarse_geometry(grid.cells, cells, len(levels))
        info = grid_info(cells, above.info.cell_m * 2.0, tuple(above.info.origin_m), tuple(above.info.open_lo),
                         tuple(above.info.open_hi), fill, ghost_gap)
        levels.append(Level(info, cells, *(torch.zeros(cells, dtype=torch.float32, device=device) for _ in range(3))))
    return levels


class Multigrid:
    """`cycles` V-cycles per solve with `sweeps` smoothing sweeps a side (sim/multigrid.py:97-104).
    tolerance > 0: stop at that FP64 relative residual, at most `max_cycles` cycles."""

    def __init__(self, grid, cycles: int, sweeps: int, tolerance: float = 0.0, max_cycles: int | None = None,
                 max_levels: int = MAX_LEVELS):
        self.cycles = cycles
        self.sweeps = sweeps
        self.tolerance = float(tolerance)
        self.max_cycles = int(max_cycles if max_cycles is not None else cycles)
        self.device = grid.device
        self.levels = build_levels(grid, max_levels)
        n = len(self.levels)
        self._infos = (GridInfo * n)(*(level.info for level in self.levels))
        doubles = LIB.plm_norm_scratch_doubles(ctypes.byref(self.levels[0].info))
        self.norm_scratch = torch.zeros(doubles, dtype=torch.float64, device=self.device)
        # [0] cycles of the last solve, [1] running total (tolerance mode)
        self.cycles_run = torch.zeros(2, dtype=torch.int32, device=self.device)
        self._predicted = 1
        self.max_cycles_seen = 0
        self.recent = deque(maxlen=RECENT_SOLVES)  # cycle counts of the latest solves
        # the last solve's cycle count, copied to pinned host memory without waiting and read by the next
        # solve (one host round trip per solve: the gate's)
        self._count_host = torch.zeros(1, dtype=torch.int32, pin_memory=True)
        self._count_event = torch.cuda.Event()
        self._count_ready = None
        # Set for graph capture: every tolerance-mode solve then launches exactly this many gated cycles and
        # reads nothing back (cycles past convergence exit on the device gate).
        self.graph_cycles = None

    def _pointers(self):
        array = ctypes.c_void_p * len(self.levels)
        return (array(*(ptr(level.q) for level in self.levels)), array(*(ptr(level.rhs) for level in self.levels)),
                array(*(ptr(level.residual) for level in self.levels)))

    def __call__(self, grid) -> None:
        self.levels[0].q = grid.pressure
        self.levels[0].rhs = grid.divergence
        if self.tolerance <= 0.0:
            self._cycles(self.cycles, resume=False)
            return
        if self.graph_cycles is not None:
            self._cycles(min(self.graph_cycles, self.max_cycles), resume=False)
            return
        # Tolerance mode: run the cycle count the last solve needed (cycles past convergence exit at once
        # on the device gate), then add single cycles while the gate says the target is not met. One host
        # read per solve in the usual case (the gate), instead of a launch per cycle up to the cap; the
        # count of cycles that ran comes back without a wait and sets the next solve's first launch.
        if self._count_ready is not None and self._count_ready.query():
            self._note_cycles(int(self._count_host[0]))
            self._count_ready = None
        predicted = max(1, min(self._predicted, self.max_cycles))
        self._cycles(predicted, resume=False)
        ran = predicted
        while ran < self.max_cycles and float(self.norm_scratch[1].item()) != 0.0:
            self._cycles(1, resume=True)
            ran += 1
        if ran > predicted:  # every added cycle ran: the count is known here
            self._note_cycles(ran)
        self._count_host.copy_(self.cycles_run[:1], non_blocking=True)
        self._count_event.record()
        self._count_ready = self._count_event

    def settle(self) -> None:
        """Waits for the last solve's cycle count and takes it into the prediction (max_cycles_seen)."""
        if self._count_ready is not None:
            self._count_ready.synchronize()
            self._note_cycles(int(self._count_host[0]))
            self._count_ready = None

    def _note_cycles(self, count: int) -> None:
        self._predicted = max(1, count)
        self.max_cycles_seen = max(self.max_cycles_seen, self._predicted)
        self.recent.append(self._predicted)

    def note_cycles(self, count: int) -> None:
        """Takes the cycle count of a solve run elsewhere (a replayed graph) into the prediction."""
        self._note_cycles(int(count))

    @property
    def recent_max(self) -> int:
        """The most cycles any of the latest solves ran (the most ever, before any solve)."""
        return max(self.recent) if self.recent else max(1, self.max_cycles_seen)

    def _cycles(self, count: int, resume: bool) -> None:
        q, rhs, residual = self._pointers()
        tolerance_mode = self.tolerance > 0.0
        call("plm_vcycle", len(self.levels), self._infos, q, rhs, residual, int(count), self.sweeps, self.tolerance,
             1 if resume else 0, ptr(self.norm_scratch, torch.float64) if tolerance_mode else None,
             ptr(self.cycles_run, torch.int32) if tolerance_mode else None, stream())

    def autotune(self, repeats: int = 5) -> dict:
        """Times every tile shape of the two-colour sweep on each level that runs it (those above the LDS
        tail) and pins the fastest per level size, process-wide. Uses scratch tensors, not the solve's
        state; synchronises. Returns {level cells: (chosen shape, microseconds per shape)}."""
        shapes = int(LIB.plm_sweep_tile_shapes(None, None, None, 0))
        tail_cells = 0
        tuned = {}
        for level in reversed(self.levels):
            tail_cells += int(level.q.numel())
            if tail_cells <= PLM_LDS_CELLS:
                continue
            q = torch.zeros_like(level.q)
            rhs = torch.randn(level.cells, generator=torch.Generator().manual_seed(1)).to(self.device)
            scratch = torch.empty_like(level.q)
            times = (ctypes.c_float * shapes)()
            chosen = ctypes.c_int32()
            call("plm_autotune_sweep", ctypes.byref(level.info), ptr(q), ptr(rhs), ptr(scratch), int(repeats), times,
                 ctypes.byref(chosen), stream())
            tuned[level.cells] = (chosen.value, [round(t, 2) for t in times])
        return tuned

    def last_relative_residual(self) -> float:
        """The FP64 |rhs - A q| / |rhs| after the last solve in tolerance mode (synchronises)."""
        return float(self.norm_scratch[0].item())

    def last_cycles(self) -> int:
        """V-cycles the last tolerance-mode solve ran (synchronises)."""
        return int(self.cycles_run[0].item())

    def total_cycles(self) -> int:
        """V-cycles run by every tolerance-mode solve so far (synchronises)."""
        return int(self.cycles_run[1].item())


class BatchedMultigrid:
    """The tolerance-mode solves of many grids (ensemble members of uneven size) advanced together by
    plm_mg_batch: each stage of a V-cycle is one launch for every member (a persistent grid of blocks over all
    members' tiles; one workgroup per member for the LDS tails), each member stopping on its own gate or its
    max_cycles. A member's pressure ends bit for bit as its own `Multigrid` would leave it, given enough cycles.
    The members' level arrays (and each grid's pressure and divergence) must keep their storage while the batch
    lives: the device table holds their addresses."""

    def __init__(self, solves: list[Multigrid], grids):
        if not solves:
            raise ValueError("an empty batch")
        sweeps = {s.sweeps for s in solves}
        if len(sweeps) != 1:
            raise ValueError("batched members must share the smoothing sweeps")
        if any(s.tolerance <= 0.0 for s in solves):
            raise ValueError("the batched solve runs in tolerance mode")
        self.solves = list(solves)
        self.grids = list(grids)
        tolerance = {s.tolerance for s in solves}
        if len(tolerance) != 1:
            raise ValueError("batched members must share the tolerance")
        self.tolerance = tolerance.pop()
        count = len(solves)
        self._keep = []  # ctypes arrays the table was built from (the device table copies their values)
        infos, qs, rhss, residuals = [], [], [], []
        for solve, grid in zip(solves, grids):
            solve.levels[0].q = grid.pressure
            solve.levels[0].rhs = grid.divergence
            q, rhs, residual = solve._pointers()
            infos.append(ctypes.cast(solve._infos, ctypes.c_void_p))
            qs.append(ctypes.cast(q, ctypes.c_void_p))
            rhss.append(ctypes.cast(rhs, ctypes.c_void_p))
            residuals.append(ctypes.cast(residual, ctypes.c_void_p))
            self._keep += [q, rhs, residual]
        array = ctypes.c_void_p * count
        n_levels = (ctypes.c_int32 * count)(*(len(s.levels) for s in solves))
        caps = (ctypes.c_int32 * count)(*(s.max_cycles for s in solves))
        scratch = array(*(ptr(s.norm_scratch, torch.float64) for s in solves))
        cycles = array(*(ptr(s.cycles_run, torch.int32) for s in solves))
        handle = ctypes.c_void_p()
        call("plm_mg_batch_create", count, n_levels, array(*infos), array(*qs), array(*rhss), array(*residuals),
             int(sweeps.pop()), self.tolerance, caps, scratch, cycles, ctypes.byref(handle))
        self._handle = handle

    def __call__(self, cycles: int) -> None:
        """Checks every member's residual, then runs up to `cycles` lockstep V-cycles on the current stream."""
        call("plm_mg_batch_solve", self._handle, int(cycles), stream())

    def autotune(self, repeats: int = 3) -> list[tuple[int, list[float]]]:
        """Picks the tile shape of each depth's sweep launches by timing every shape over all members (writes
        only the members' scratch levels; synchronises). Returns (chosen shape, microseconds per shape) per
        depth."""
        depth = int(LIB.plm_mg_batch_depth(self._handle))
        shapes = int(LIB.plm_sweep_tile_shapes(None, None, None, 0))
        times = (ctypes.c_float * max(1, depth * shapes))()
        call("plm_mg_batch_autotune", self._handle, int(repeats), times, stream())
        out = []
        for d in range(depth):
            row = [round(times[d * shapes + v], 2) for v in range(shapes)]
            out.append((min(range(shapes), key=lambda v: row[v]), row))
        return out

    def close(self) -> None:
        if self._handle is not None and self._handle.value:
            call("plm_mg_batch_destroy", self._handle)
        self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown: the library may be gone
            pass


def residual(level: Level) -> torch.Tensor:
    """residual_kernel (sim/multigrid.py:36-43) into `level.residual`."""
    call("plm_residual", ctypes.byref(level.info), ptr(level.q), ptr(level.rhs), ptr(level.residual), stream())
    return level.residual


def relative_residual(info, q: torch.Tensor, rhs: torch.Tensor) -> float:
    """|rhs - A q| / |rhs| in FP64 on the device (synchronises)."""
    scratch = torch.zeros(LIB.plm_norm_scratch_doubles(ctypes.byref(info)), dtype=torch.float64, device=q.device)
    out = torch.zeros(2, dtype=torch.float64, device=q.device)
    call("plm_residual_norms", ctypes.byref(info), ptr(q), ptr(rhs), ptr(scratch, torch.float64),
         ptr(out, torch.float64), stream())
    r2, b2 = out.tolist()
    return (r2 ** 0.5) / (b2 ** 0.5) if b2 > 0 else float("inf") if r2 > 0 else 0.0
