"""Recipe tables of the benchmark scenes: the buoyant smoke puff at any grid size.

The puff maps the benchmark's normalised description onto the reference's recipe keys (recipe/schema.py):
1 cm cells, floor closed and the other faces open (the schema's defaults), ambient temperature 1 and emitter
temperature 2 so that buoyancy is beta (T - 1) - alpha s exactly (heat_buoyancy = beta / g, smoke_weight =
alpha / g, sim/forces.py:20-23), a spherical emitter 10 cm above the floor centre for the first 0.5 s
(12 frames at 24 fps), 4 substeps per frame, multigrid V(2, 2) to a relative residual.
"""

GRAVITY = 9.81
CELL_M = 0.01


def puff_table(cells, frames: int = 96, radius_cells: float = 12.0, alpha: float = 0.1, beta: float = 1.0,
               confinement: float = 0.3, tolerance: float = 1e-5, cycles: int = 2, max_cycles: int = 24,
               seed: int = 7) -> dict:
    """The recipe table (as `plume run` reads it from TOML) of the puff on a `cells` grid."""
    size = [n * CELL_M for n in cells]
    return {
        "seed": seed,
        "domain": {"size_m": size, "cell_m": CELL_M, "ambient_k": 1.0},
        "time": {"duration_s": frames / 24.0, "frames": frames, "substeps": 4},
        "emitter": [{"name": "puff", "center_m": [0.0, 0.10, 0.0], "radius_m": radius_cells * CELL_M,
                     "edge_softness": 0.6, "end_s": 0.5, "smoke_per_s": 2.0, "temperature_k": 2.0}],
        "forces": {"gravity_mps2": GRAVITY, "heat_buoyancy": beta / GRAVITY, "smoke_weight": alpha / GRAVITY,
                   "vorticity_confinement": confinement},
        "solver": {"multigrid_cycles": cycles, "multigrid_sweeps": 2, "multigrid_tolerance": tolerance,
                   "multigrid_max_cycles": max_cycles},
        "render": {"glow_start_k": 1.0, "opacity_per_density": 20.0, "light_dir": [0.5, 0.7, 0.5]},
    }


def differentiable_recipe(cells=(64, 96, 64), radius_cells: float = 6.0, tolerance: float = 1e-5):
    """(recipe, derived) of the differentiable-rollout scene: the puff (alpha 0.1, beta 1, confinement 0.3) on
    `cells` with the glow starting below the ambient 1, so that the heat channel's max(T - glow, 0) has no kink
    in still air."""
    from plume_amd.recipe.derived import derive
    from plume_amd.recipe.load import fill_dependent_defaults, parse_recipe

    table = puff_table(cells, frames=2, radius_cells=radius_cells, tolerance=tolerance)
    table["render"]["glow_start_k"] = 0.5
    recipe = fill_dependent_defaults(parse_recipe(table), "differentiable")
    return recipe, derive(recipe)


def differentiable_puff(device: str, cells=(64, 96, 64), pixels: int = 128, march_steps: int = 192,
                        tolerance: float = 1e-5, radius_cells: float = 6.0):
    """The differentiable-rollout scene (differentiable_recipe) with per-cell sources of a Gaussian-falloff
    sphere of `radius_cells` centred 10 cells above the floor (smoke 2 /s, heat 4 K/s on the normalised scale,
    lift 1 m/s^2 on v faces), and a DiffSim rendering `pixels`^2 heat and thickness images with `march_steps`
    samples per ray. Returns (recipe, derived, sim, sources)."""
    from plume_amd.sim.adjoint import DiffSim, sphere_sources

    recipe, derived = differentiable_recipe(cells, radius_cells, tolerance)
    sim = DiffSim(recipe, derived, device, pixels=pixels, march_steps=march_steps, tolerance=tolerance)
    center = (cells[0] / 2.0, 10.0, cells[2] / 2.0)
    sources = sphere_sources(sim.grid, center, radius_cells, falloff="gaussian")
    return recipe, derived, sim, sources
