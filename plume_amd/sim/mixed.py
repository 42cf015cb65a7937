"""Mixed-precision pressure solve: iterative refinement around a low-precision multigrid.

The reference solves the pressure Poisson equation in FP32 only (sim/multigrid.py:97-147); this is an addition
with no reference counterpart, for large grids where the solve is HBM bound. The FP32 pressure x is refined
until |b - A x| / |b| <= tolerance, the norms reduced in FP64 on the device. Each round

1. forms r = b - A x in FP32 (residual_kernel, sim/multigrid.py:36-43);
2. scales r by a power of two (exact) that brings its RMS near 1, so FP16 neither overflows nor underflows,
   and rounds it into the storage type (BF16 or FP16, round to nearest even);
3. runs `inner_cycles` V(sweeps, sweeps) cycles of the reference's V-cycle on A e = r from e = 0, every
   level's q and rhs stored in that type, FP32 arithmetic in the sweeps (form "f16x2": the flat levels' sweeps
   in packed FP16, two cells per lane), FP32 residuals inside the fused residual + restrict, and the
   coarsest level solved in FP64 (plm_vcycle_lp);
4. adds e back into x in FP32.

Steps 4, 1 and 2 of consecutive rounds and the FP64 norms run as one kernel (plm_refine_residual), which
writes the corrected x to the other of two FP32 buffers (the neighbours of a cell still read the old x). Its
power-of-two scale is predicted from the previous rounds' residuals; a prediction more than 2^4 off the
measured RMS stores r again with the measured scale.

The FP32 pressure itself bounds the reachable residual: even the exact solution rounded to FP32 leaves a
relative residual of order 1e-6 on grids of 10^5 cells and more. A round that does not halve the residual
therefore ends the solve (`log.stalled`) instead of spinning at that floor; when the first round itself
reduces less (a weak inner cycle, such as V(1,1) on 16-bit storage), the bar is 1.5 times its ratio, at most
0.95 (stall_threshold).

The level hierarchy is the reference's (Multigrid._build_levels, sim/multigrid.py:106-122), optionally cut
at `max_levels` so that the coarsest level is the one the FP64 solve handles (8 x 12 x 8 at 256 x 384 x 256).
"""

import ctypes
import math
import time
from dataclasses import dataclass, field

import torch

from plume_amd._lib import LIB, GridInfo, call, ptr, stream
from plume_amd.sim.multigrid import MAX_LEVELS, build_levels

# PLM_STORAGE_* in include/plume_hip.h; "f16x2" stores FP16 and sweeps the flat levels (the simulation level
# and the coarser levels of even-sized grids) in packed FP16 arithmetic (two cells per lane)
STORAGE = {"f32": 0, "bf16": 1, "f16": 2, "f16x2": 3}
STALL = 0.5  # a round that does not at least halve the residual ends the solve (V(2,2) and stronger cycles)
DTYPES = {"f32": torch.float32, "bf16": torch.bfloat16, "f16": torch.float16, "f16x2": torch.float16}


def _vptr(t: torch.Tensor) -> int:
    if not t.is_contiguous() or t.device.type != "cuda":
        raise ValueError("expected a contiguous device tensor")
    return t.data_ptr()


@dataclass
class RefinementLog:
    """Relative residual before each refinement round and after the last; `stalled` when the solve stopped
    because a round no longer halved the residual (the FP32 pressure's representation floor)."""

    residuals: list = field(default_factory=list)
    stalled: bool = False

    @property
    def rounds(self) -> int:
        return max(0, len(self.residuals) - 1)


class MixedPrecisionMultigrid:
    """Pressure solve in the reference's PressureSolve role (grid.pressure refined in place from its current
    value, right-hand side grid.divergence), see the module docstring."""

    def __init__(self, grid, storage: str = "bf16", tolerance: float = 1e-6, sweeps: int = 2,
                 inner_cycles: int = 1, max_rounds: int = 60, max_levels: int = MAX_LEVELS):
        if storage not in STORAGE:
            raise ValueError(f"storage must be one of {tuple(STORAGE)}")
        if sweeps < 1 or inner_cycles < 1:
            raise ValueError("sweeps and inner_cycles must be >= 1")
        self.storage = storage
        self.tolerance = float(tolerance)
        self.sweeps = sweeps
        self.inner_cycles = inner_cycles
        self.max_rounds = max_rounds
        levels = build_levels(grid, max_levels)
        if len(levels) < 2:
            raise ValueError("the mixed-precision solve needs at least two levels")
        self.infos = (GridInfo * len(levels))(*(level.info for level in levels))
        self.cells = [level.cells for level in levels]
        dtype = DTYPES[storage]
        device = grid.device
        self.q = [torch.zeros(c, dtype=dtype, device=device) for c in self.cells]
        self.rhs = [torch.zeros(c, dtype=dtype, device=device) for c in self.cells]
        self.scratch = [torch.zeros(c, dtype=dtype, device=device) for c in self.cells[:-1]]
        self.x_alt = torch.zeros(self.cells[0], dtype=torch.float32, device=device)
        self.norm_scratch = torch.zeros(LIB.plm_norm_scratch_doubles(ctypes.byref(self.infos[0])),
                                        dtype=torch.float64, device=device)
        self.norms = torch.zeros(2, dtype=torch.float64, device=device)
        self.log = RefinementLog()

    def _arrays(self):
        n = len(self.cells)
        array = ctypes.c_void_p * n
        return (array(*(_vptr(t) for t in self.q)), array(*(_vptr(t) for t in self.rhs)),
                array(*([_vptr(t) for t in self.scratch] + [None])))

    def relative_residual(self, grid) -> tuple[float, float]:
        """(|b - A x| / |b|, RMS of b - A x) in FP64 (synchronises)."""
        call("plm_residual_norms", ctypes.byref(self.infos[0]), ptr(grid.pressure), ptr(grid.divergence),
             ptr(self.norm_scratch, torch.float64), ptr(self.norms, torch.float64), stream())
        r2, b2 = self.norms.tolist()
        count = grid.pressure.numel()
        rel = math.sqrt(r2 / b2) if b2 > 0 else (0.0 if r2 == 0 else math.inf)
        return rel, math.sqrt(r2 / count)

    def _refine(self, grid, x, e, inv_scale, x_out, r_scale) -> tuple[float, float]:
        """plm_refine_residual on the simulation level: x_out = x + e * inv_scale (e None: no correction, x_out
        None: not written), rhs[0] = storage(r * r_scale) for r = b - A x_out; returns (|r| / |b|, RMS of r)
        from the FP64 norms (synchronises)."""
        call("plm_refine_residual", ctypes.byref(self.infos[0]), ptr(x), None if e is None else _vptr(e),
             inv_scale, STORAGE[self.storage], ptr(grid.divergence), None if x_out is None else ptr(x_out),
             _vptr(self.rhs[0]), r_scale, ptr(self.norm_scratch, torch.float64), ptr(self.norms, torch.float64),
             stream())
        r2, b2 = self.norms.tolist()
        rel = math.sqrt(r2 / b2) if b2 > 0 else (0.0 if r2 == 0 else math.inf)
        return rel, math.sqrt(r2 / grid.pressure.numel())

    def __call__(self, grid) -> None:
        x = grid.pressure
        code = STORAGE[self.storage]
        q, rhs, scratch = self._arrays()
        self.log = RefinementLog()
        buffers = (x, self.x_alt)
        current = 0  # buffers[current] holds the latest x
        stored = None  # the scale rhs[0] was stored with, when it was the latest r's
        bar = STALL  # stall threshold, set from the first round's ratio
        try:
            rel, rms = self.relative_residual(grid)
            while True:
                self.log.residuals.append(rel)
                if rel <= self.tolerance or rms == 0.0 or len(self.log.residuals) > self.max_rounds:
                    return
                if len(self.log.residuals) >= 2:
                    if len(self.log.residuals) == 2:
                        bar = stall_threshold(rel / self.log.residuals[0])
                    if rel > bar * self.log.residuals[-2]:
                        self.log.stalled = True
                        return
                scale = _pow2_scale(rms)
                if stored is None or abs(math.log2(stored / scale)) > 4:
                    self._refine(grid, buffers[current], None, 0.0, None, scale)
                    stored = scale
                call("plm_vcycle_lp", len(self.cells), self.infos, code, q, rhs, scratch, self.inner_cycles,
                     self.sweeps, 1, stream())
                # the next round's r, stored with the scale its predicted RMS asks for
                ratio = rel / self.log.residuals[-2] if len(self.log.residuals) >= 2 else 0.1
                predicted = _pow2_scale(rms * min(max(ratio, 1e-3), 1.0))
                rel, rms = self._refine(grid, buffers[current], self.q[0], 1.0 / stored, buffers[1 - current],
                                        predicted)
                current = 1 - current
                stored = predicted
        finally:
            if current == 1:
                x.copy_(self.x_alt)


def stall_threshold(first_ratio: float) -> float:
    """The residual ratio above which a refinement round counts as stalled, from the first round's ratio:
    STALL for cycles that cut the residual well, up to 0.95 for weak ones (1.5 times the first ratio)."""
    return max(STALL, min(0.95, 1.5 * first_ratio))


def _pow2_scale(rms: float) -> float:
    """The power of two that brings `rms` nearest 1 (exact to apply in any floating-point type)."""
    return 2.0 ** (-round(math.log2(rms)))


def divergent_velocity(grid, seed: int, blobs: int = 64, emitter_radius_cells: float = 12.0,
                       emitter_speed: float = 1.0) -> None:
    """Fills grid.u/v/w with a seeded divergent field: `blobs` Gaussian vortex blobs (random centre, axis,
    strength and radius) plus a radial emitter source at the domain's floor centre, 10 cells up. Evaluated
    in FP64 at each face position, stored as FP32."""
    g = torch.Generator().manual_seed(seed)
    nx, ny, nz = grid.cells
    size = torch.tensor([nx, ny, nz], dtype=torch.float64)
    centers = torch.rand(blobs, 3, generator=g, dtype=torch.float64) * size
    axes = torch.nn.functional.normalize(torch.randn(blobs, 3, generator=g, dtype=torch.float64), dim=1)
    strengths = (torch.rand(blobs, generator=g, dtype=torch.float64) * 2.0 - 1.0) * 0.2
    radii = (0.05 + 0.1 * torch.rand(blobs, generator=g, dtype=torch.float64)) * float(min(nx, ny, nz))
    source = torch.tensor([nx / 2.0, 10.0, nz / 2.0], dtype=torch.float64)
    device = grid.device
    for axis, f in enumerate(grid.velocity_fields()):
        shape = tuple(f.data.shape)
        offset = [0.5, 0.5, 0.5]
        offset[axis] = 0.0
        coords = [torch.arange(shape[a], dtype=torch.float64, device=device) + offset[a] for a in range(3)]
        X, Y, Z = torch.meshgrid(*coords, indexing="ij")
        out = torch.zeros(shape, dtype=torch.float64, device=device)
        for b in range(blobs):
            c = centers[b].tolist()
            dx, dy, dz = X - c[0], Y - c[1], Z - c[2]
            weight = torch.exp(-(dx * dx + dy * dy + dz * dz) / (2.0 * float(radii[b]) ** 2)) * float(strengths[b])
            a = axes[b].tolist()
            # (axis x d)[axis component] of a vortex about `a` through the centre
            d = (dx, dy, dz)
            cross = (a[1] * d[2] - a[2] * d[1], a[2] * d[0] - a[0] * d[2], a[0] * d[1] - a[1] * d[0])
            out += weight * cross[axis] / float(radii[b])
        s = source.tolist()
        dx, dy, dz = X - s[0], Y - s[1], Z - s[2]
        dist = torch.sqrt(dx * dx + dy * dy + dz * dz).clamp_min(1e-9)
        radial = (dx, dy, dz)[axis] / dist
        out += emitter_speed * radial * torch.exp(-(dist * dist) / (2.0 * emitter_radius_cells**2))
        f.data.copy_(out.to(torch.float32))


def fp32_floor(grid, max_cycles: int = 60) -> float:
    """The lowest FP64 relative residual the plain FP32 multigrid reaches on grid.divergence from q = 0: V-cycles
    until one no longer cuts the residual by 10%. That stall is set by FP32 itself, the pressure's rounding
    and the rounding of A q in the FP32 residual (terms of size q / h^2 that cancel), so no FP32 pressure
    reports much less. Leaves grid.pressure at the stalled solution."""
    from plume_amd.sim.multigrid import Multigrid, relative_residual

    solve = Multigrid(grid, 1, 2)
    grid.pressure.zero_()
    best = previous = math.inf
    for _ in range(max_cycles):
        solve(grid)
        residual = relative_residual(grid.info, grid.pressure, grid.divergence)
        best = min(best, residual)
        if residual > 0.9 * previous:
            break
        previous = residual
    return best


@dataclass
class Problem:
    """A projection problem: the grid (its divergence set), the divergent velocity and its divergence."""

    grid: object
    velocity: list
    divergence: torch.Tensor


def divergent_problem(cells, device: str, seed: int = 20261208) -> Problem:
    """The mixed-precision projection problem on a `cells` puff grid (1 cm cells, floor closed, the rest
    open): divergent_velocity with `seed`, closed faces zeroed, its divergence as the right-hand side."""
    from plume_amd.recipe.derived import derive
    from plume_amd.recipe.load import fill_dependent_defaults, parse_recipe
    from plume_amd.sim.grid import make_grid
    from plume_amd.sim.pressure import compute_divergence, zero_closed_faces
    from plume_amd.sim.scenes import puff_table

    recipe = fill_dependent_defaults(parse_recipe(puff_table(tuple(cells))), "mixed")
    derived = derive(recipe)
    grid = make_grid(derived, recipe.domain.cell_m, recipe.domain.faces, recipe.domain.ambient_k, device)
    divergent_velocity(grid, seed=seed)
    zero_closed_faces(grid)
    compute_divergence(grid)
    return Problem(grid, [f.data.clone() for f in grid.velocity_fields()], grid.divergence.clone())


def solve_forms(problem: Problem, target: float, forms=("f32", "bf16", "f16", "f16x2"), max_levels: int = 6,
                repeats: int = 2, sweeps: int = 2) -> dict:
    """Solves the problem from q = 0 to `target` with each form: "f32" the plain FP32 multigrid in tolerance
    mode (autotuned tiles), "bf16" / "f16" / "f16x2" iterative refinement around that storage, every form with
    V(sweeps, sweeps) cycles. The last of `repeats`
    solves is timed (the first warms the code paths). Returns {form: (entry, pressure)} where entry holds the
    seconds, cells per second, FP64 relative residual and post-projection divergence |div u| / |div u0|."""
    from plume_amd.sim.multigrid import Multigrid, relative_residual
    from plume_amd.sim.pressure import compute_divergence, subtract_gradient

    grid = problem.grid
    cells = grid.pressure.numel()
    div0_norm = float(torch.linalg.vector_norm(problem.divergence.double()))
    out = {}
    for name in forms:
        if name == "f32":
            solve = Multigrid(grid, 2, sweeps, tolerance=target, max_cycles=80)
            solve.autotune()
        else:
            solve = MixedPrecisionMultigrid(grid, name, tolerance=target, sweeps=sweeps, max_levels=max_levels)
        for _ in range(repeats):
            grid.pressure.zero_()
            torch.cuda.synchronize()
            start = time.perf_counter()
            solve(grid)
            torch.cuda.synchronize()
            seconds = time.perf_counter() - start
        residual = relative_residual(grid.info, grid.pressure, grid.divergence)
        q = grid.pressure.clone()
        for f, v0 in zip(grid.velocity_fields(), problem.velocity):
            f.data.copy_(v0)
        subtract_gradient(grid)
        compute_divergence(grid)
        post = float(torch.linalg.vector_norm(grid.divergence.double())) / div0_norm
        grid.divergence.copy_(problem.divergence)
        entry = {"seconds": seconds, "cell_solves_per_s": cells / seconds, "relative_residual": residual,
                 "post_projection_divergence": post}
        if name == "f32":
            entry["v_cycles"] = solve.last_cycles()
        else:
            entry.update(refinement_rounds=solve.log.rounds, inner_v_cycles=solve.log.rounds * solve.inner_cycles,
                         stalled=solve.log.stalled)
        out[name] = (entry, q)
    return out
