"""The differentiable rollout restated on the FP64 oracle (the checker of plume_amd.sim.adjoint).

Per substep, in the order of DiffSim.substep: per-cell emission (smoke += dt S_s, temperature += dt S_T,
v += dt S_v), buoyancy (sim/forces.py:13-31) with smoke_weight = alpha / g, vorticity confinement
(sim/vorticity.py:95-106), MacCormack of u, v, w by the old velocity and the swaps, projection with multigrid
to an FP64 relative residual of 1e-13 (sim/pressure.py:144-154), MacCormack of smoke and temperature by the
projected velocity. The heat and thickness images after every `checkpoint` substeps come from the oracle's
ray march with a fixed sample count (render/raymarch.py:86-140).
"""

from dataclasses import dataclass

import numpy as np


@dataclass
class Camera:
    """The view and look of the rollout's render as plain numbers (ViewParams / LookParams)."""

    right: tuple
    up: tuple
    toward: tuple
    left_m: float
    top_m: float
    pixel_m: float
    perspective: int
    opacity_per_density: float
    glow_start_k: float
    smoke_hides_heat: float
    normal_fade_thickness: float
    shade_fill: float

    @classmethod
    def of(cls, sim):
        v, lk = sim.view, sim.look
        return cls(tuple(v.right), tuple(v.up), tuple(v.toward), v.left_m, v.top_m, v.pixel_m, v.perspective,
                   lk.opacity_per_density, lk.glow_start_k, lk.smoke_hides_heat, lk.normal_fade_thickness,
                   lk.shade_fill)

    def as_array(self) -> np.ndarray:
        return np.array([*self.right, *self.up, *self.toward, self.left_m, self.top_m, self.pixel_m,
                         self.perspective, self.opacity_per_density, self.glow_start_k, self.smoke_hides_heat,
                         self.normal_fade_thickness, self.shade_fill], dtype=np.float64)

    @classmethod
    def from_array(cls, a):
        a = [float(x) for x in a]
        return cls(tuple(a[0:3]), tuple(a[3:6]), tuple(a[6:9]), a[9], a[10], a[11], int(a[12]), a[13], a[14],
                   a[15], a[16], a[17])


def rollout_images(o, recipe, derived, dt: float, camera: Camera, sources64, alpha: float, substeps: int,
                   checkpoint: int, pixels: int, march_steps: int, initial=None) -> np.ndarray:
    """(substeps / checkpoint, pixels, pixels, 2) heat and thickness images of the FP64 rollout from rest
    (zero velocity and smoke, ambient temperature) or from the fields of `initial`."""
    from oracle.sim import Multigrid, VorticityConfinement, advect_maccormack, apply_buoyancy, make_grid
    from oracle.sim import poisson_residual, project

    grid = make_grid(o, derived, recipe.domain.cell_m, recipe.domain.faces, recipe.domain.ambient_k)
    conf = VorticityConfinement(grid)
    for name, values in (initial or {}).items():
        getattr(grid, name).data[...] = np.asarray(values, dtype=np.float64)
    amb, g = recipe.domain.ambient_k, recipe.forces.gravity_mps2

    def solve(gr):
        s = Multigrid(gr, 1, 2)
        for _ in range(60):
            s(gr)
            if poisson_residual(gr) < 1e-13:
                break

    view = o.view(camera.right, camera.up, camera.toward, camera.left_m, camera.top_m, camera.pixel_m,
                  camera.perspective, (0.0, 0.0, 0.0))
    look = o.look(camera.opacity_per_density, camera.glow_start_k, camera.smoke_hides_heat,
                  camera.normal_fade_thickness, 1, camera.shade_fill)
    images = []
    for _ in range(substeps // checkpoint):
        for _ in range(checkpoint):
            grid.smoke.data[...] += dt * sources64[0]
            grid.temperature.data[...] += dt * sources64[1]
            grid.v.data[...] += dt * sources64[2]
            apply_buoyancy(grid, g, recipe.forces.heat_buoyancy, alpha / g, amb, dt)
            conf.apply(grid, recipe.forces.vorticity_confinement, dt)
            fields = grid.velocity_fields()
            old = tuple(f.data for f in fields)
            for f in fields:
                advect_maccormack(f, old, grid, dt)
            for f in fields:
                f.swap()
            project(grid, solve)
            vel = tuple(f.data for f in grid.velocity_fields())
            advect_maccormack(grid.smoke, vel, grid, dt)
            advect_maccormack(grid.temperature, vel, grid, dt)
            grid.smoke.swap()
            grid.temperature.swap()
        out = np.zeros((pixels, pixels, 5))
        o.raymarch(grid.info, grid.smoke.data, grid.temperature.data, np.zeros((1, 1, 1, 3)), np.ones((1, 1, 1)),
                   view, look, out, march_steps)
        images.append(out[..., :2].copy())
    return np.stack(images)
