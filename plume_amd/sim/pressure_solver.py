"""Choosing the pressure solve a recipe asks for (sim/pressure_solver.py of the reference)."""

from plume_amd.sim.multigrid import Multigrid
from plume_amd.sim.pressure import GaussSeidel


def make_pressure_solve(grid, solver):
    """The solve for `grid`; multigrid allocates its coarse levels here, once per run."""
    if solver.pressure_solver == "gauss_seidel":
        return GaussSeidel(solver.pressure_iterations)
    tolerance = getattr(solver, "multigrid_tolerance", 0.0)
    return Multigrid(grid, solver.multigrid_cycles, solver.multigrid_sweeps, tolerance=tolerance,
                     max_cycles=getattr(solver, "multigrid_max_cycles", None) if tolerance > 0 else None)
