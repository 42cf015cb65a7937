"""One simulation step in the reference's order (sim/stepper.py): emit, buoyancy, confinement, turbulence,
drag, advect velocity (all three from the old velocity, then swap), project, advect smoke and hold its
budget, advect temperature, dissipate."""

import numpy as np

from plume_amd.sim.advect import advect_scalars, advect_velocity
from plume_amd.sim.conserve import SmokeBudget
from plume_amd.sim.dissipate import dissipate
from plume_amd.sim.emitters import emit, emitter_params
from plume_amd.sim.forces import apply_buoyancy, apply_drag
from plume_amd.sim.pressure import project
from plume_amd.sim.pressure_solver import make_pressure_solve
from plume_amd.sim.turbulence import apply_turbulence
from plume_amd.sim.vorticity import VorticityConfinement


class Stepper:
    """Advances a grid one step at a time under a recipe's emitters, forces and dissipation
    (sim/stepper.py:17-26)."""

    def __init__(self, grid, recipe):
        self.grid = grid
        self.recipe = recipe
        self.emitters = emitter_params(recipe.emitter, grid.device)
        self.pressure_solve = make_pressure_solve(grid, recipe.solver)
        self.smoke_budget = SmokeBudget(grid)
        self.confinement = VorticityConfinement(grid) if recipe.forces.vorticity_confinement > 0 else None

    def step(self, time_s: float, dt: float) -> None:
        """Advances from `time_s` to `time_s + dt` (sim/stepper.py:28-44)."""
        self.before_projection(time_s, dt)
        project(self.grid, self.pressure_solve)
        self.after_projection(dt)

    def before_projection(self, time_s: float, dt: float) -> None:
        """The step up to the projection: emit, forces, velocity advection (sim/stepper.py:30-41)."""
        grid, recipe = self.grid, self.recipe
        ambient = recipe.domain.ambient_k
        if self._any_emitter_active(time_s):
            emit(grid, self.emitters, recipe.seed, time_s, dt)
        forces = recipe.forces
        apply_buoyancy(grid, forces.gravity_mps2, forces.heat_buoyancy, forces.smoke_weight, ambient, dt)
        if self.confinement is not None:
            self.confinement.apply(grid, forces.vorticity_confinement, dt)
        turbulence = forces.turbulence
        apply_turbulence(grid, recipe.seed, turbulence.scale_m, turbulence.strength_mps2, turbulence.speed_mps,
                         time_s, dt)
        apply_drag(grid, forces.wind_mps, forces.drag_per_s, dt)
        self._advect_velocity(dt)

    def after_projection(self, dt: float) -> None:
        """The step after the projection: scalar advection with the smoke budget, dissipation (sim/stepper.py:
        42-44)."""
        recipe = self.recipe
        self._advect_scalars(dt)
        dissipate(self.grid, recipe.dissipation.smoke_per_s, recipe.dissipation.cooling_per_s,
                  recipe.domain.ambient_k, dt)

    def _any_emitter_active(self, time_s: float) -> bool:
        """is_active (sim/emitters.py:80-82) of any emitter, compared in float32 as the kernels compare; when
        none is active the emit kernels would change nothing, so they are not launched."""
        t = np.float32(time_s)
        return any(np.float32(p.start_s) <= t < np.float32(p.end_s) for p in self.emitters.params)

    def _advect_velocity(self, dt: float) -> None:
        """sim/stepper.py:46-52, as one fused launch pair over u, v and w."""
        advect_velocity(self.grid, dt)
        for field in self.grid.velocity_fields():
            field.swap()

    def _advect_scalars(self, dt: float) -> None:
        """sim/stepper.py:54-60. Smoke and temperature share one trace (temperature does not read smoke, so
        advecting it before the smoke budget changes nothing)."""
        velocity = tuple(field.data for field in self.grid.velocity_fields())
        advect_scalars(self.grid, velocity, dt)
        self.smoke_budget.hold(self.grid)
        self.grid.smoke.swap()
        self.grid.temperature.swap()
