"""Short rollouts and rendered frames of whole substeps: the HIP path against the FP32 oracle, determinism,
and the reference's frame-loop behaviour (tests/sim/test_frame_loop.py, test_step_cap.py of chadmv/plume).

Tolerance for rollouts: transcendental functions (the emitter's pow, the march's exp) differ by an ulp
between the GPU and glibc, and the MacCormack clamp can turn such a difference into a local jump, so the
bar is the relative L2 error of whole fields (1e-5 over 6 substeps) and of rendered images (1e-4).
"""

import numpy as np
import pytest
import torch

from support import TINY, tiny_recipe

pytestmark = pytest.mark.gpu


def rel_l2(actual, expected):
    actual = np.asarray(actual, dtype=np.float64)
    expected = np.asarray(expected, dtype=np.float64)
    return float(np.linalg.norm(actual - expected) / max(np.linalg.norm(expected), 1e-30))


def puff_recipe(**solver):
    return tiny_recipe(emitter=[{"name": "puff", "center_m": [0.0, 0.25, 0.0], "radius_m": 0.15, "end_s": 0.4,
                                 "smoke_per_s": 5.0, "temperature_k": 600.0, "velocity_mps": [0.0, 0.4, 0.0],
                                 "noise": {"scale_m": 0.1, "strength": 0.5}}],
                       forces={"vorticity_confinement": 2.0, "drag_per_s": 0.5,
                               "turbulence": {"scale_m": 0.2, "strength_mps2": 0.12, "speed_mps": 0.2}},
                       dissipation={"smoke_per_s": 1.6, "cooling_per_s": 1.0}, solver=solver)


def stepped(recipe, device, o, steps):
    from oracle.sim import Stepper as OracleStepper, make_grid as oracle_make_grid
    from plume_amd.recipe.derived import derive
    from plume_amd.sim.grid import make_grid
    from plume_amd.sim.stepper import Stepper

    derived = derive(recipe)
    grid = make_grid(derived, recipe.domain.cell_m, recipe.domain.faces, recipe.domain.ambient_k, device)
    reference = oracle_make_grid(o, derived, recipe.domain.cell_m, recipe.domain.faces, recipe.domain.ambient_k)
    gpu_stepper, oracle_stepper = Stepper(grid, recipe), OracleStepper(reference, recipe)
    for index in range(steps):
        gpu_stepper.step(index * derived.dt_s, derived.dt_s)
        oracle_stepper.step(index * derived.dt_s, derived.dt_s)
    torch.cuda.synchronize()
    return grid, reference, gpu_stepper, oracle_stepper


def test_six_substeps_with_every_force_match_the_oracle(device, oracle32):
    grid, reference, _, _ = stepped(puff_recipe(), device, oracle32, 6)
    for name in ("u", "v", "w", "smoke", "temperature"):
        error = rel_l2(getattr(grid, name).data.cpu().numpy(), getattr(reference, name).data)
        assert error < 1e-5, (name, error)
    assert rel_l2(grid.pressure.cpu().numpy(), reference.pressure) < 1e-5
    assert float(grid.smoke.data.sum()) > 0


def test_gauss_seidel_rollout_matches_the_oracle(device, oracle32):
    grid, reference, _, _ = stepped(puff_recipe(pressure_solver="gauss_seidel", pressure_iterations=30), device,
                                    oracle32, 4)
    for name in ("u", "v", "w", "smoke"):
        assert rel_l2(getattr(grid, name).data.cpu().numpy(), getattr(reference, name).data) < 1e-5, name


def test_identical_runs_give_identical_bytes(device, oracle32):
    from plume_amd.recipe.derived import derive
    from plume_amd.sim.run import simulate

    recipe = puff_recipe()
    frames = []
    for _ in range(2):
        captured = []
        simulate(recipe, derive(recipe), device, lambda frame, grid: captured.append(
            np.concatenate([grid.smoke.data.cpu().numpy().ravel(), grid.v.data.cpu().numpy().ravel()])))
        frames.append(captured)
    assert all(a.tobytes() == b.tobytes() for a, b in zip(*frames))


def run_frames(recipe, device, collect):
    from plume_amd.recipe.derived import derive
    from plume_amd.sim.run import simulate_frames

    results = []
    frames = simulate_frames(recipe, derive(recipe), device)
    while True:
        try:
            _, grid = next(frames)
        except StopIteration as done:
            return results, done.value
        results.append(collect(grid))


def smoke_height(grid):
    smoke = grid.smoke.data.cpu().numpy()
    heights = np.arange(smoke.shape[1]) + 0.5
    return float((smoke.sum(axis=(0, 2)) * heights).sum() / smoke.sum())


def test_hot_puff_rises(device):
    recipe = tiny_recipe(emitter=[{"name": "puff", "center_m": [0.0, 0.25, 0.0], "radius_m": 0.15, "end_s": 0.2,
                                   "smoke_per_s": 5.0, "temperature_k": 600.0}])
    heights, _ = run_frames(recipe, device, smoke_height)
    assert heights == sorted(heights)
    assert heights[-1] - heights[0] > 2.0


def test_smoke_is_nearly_conserved_in_a_closed_box(device):
    closed = {"bottom": "closed", "top": "closed", "left": "closed", "right": "closed", "back": "closed",
              "front": "closed"}
    recipe = tiny_recipe(domain={"size_m": [1.2, 1.2, 1.2], "cell_m": 0.05, "faces": closed},
                         emitter=[{"name": "puff", "center_m": [0.0, 0.4, 0.0], "radius_m": 0.2, "end_s": 0.05,
                                   "smoke_per_s": 5.0, "temperature_k": 320.0}],
                         time={"duration_s": 2.0, "frames": 8, "substeps": 5})
    totals, _ = run_frames(recipe, device, lambda grid: float(grid.smoke.data.double().sum()))
    assert totals[-1] == pytest.approx(totals[0], rel=0.15)
    assert totals[-1] <= totals[0] * (1 + 1e-5)


def test_step_cap_counts_match_the_reference(device):
    _, log = run_frames(tiny_recipe(), device, lambda grid: None)
    assert log.steps_per_frame == (3, 3, 3, 3) and log.added_steps == 0
    recipe = tiny_recipe(emitter=[{"name": "jet", "center_m": [0.0, 0.3, 0.0], "radius_m": 0.15, "end_s": 0.5,
                                   "smoke_per_s": 5.0, "velocity_mps": [0.0, 4.0, 0.0]}])
    _, log = run_frames(recipe, device, lambda grid: None)
    assert log.steps_per_frame[0] == 20
    assert log.steps_per_frame[1] >= 20


def test_more_than_max_substeps_is_a_recipe_error(device):
    from plume_amd.recipe.errors import RecipeError

    recipe = tiny_recipe(emitter=[{"name": "jet", "center_m": [0.0, 0.3, 0.0], "radius_m": 0.15, "end_s": 0.5,
                                   "velocity_mps": [0.0, 4.0, 0.0]}], solver={"max_substeps": 10})
    with pytest.raises(RecipeError, match="max_substeps"):
        run_frames(recipe, device, lambda grid: None)


def test_step_cap_matches_the_oracle_loop(device, oracle32):
    from oracle.sim import simulate_frames as oracle_frames
    from plume_amd.recipe.derived import derive

    recipe = tiny_recipe(emitter=[{"name": "jet", "center_m": [0.0, 0.3, 0.0], "radius_m": 0.15, "end_s": 0.5,
                                   "smoke_per_s": 5.0, "velocity_mps": [0.0, 2.0, 0.0]}])
    _, log = run_frames(recipe, device, lambda grid: None)
    frames = oracle_frames(oracle32, recipe, derive(recipe))
    while True:
        try:
            next(frames)
        except StopIteration as done:
            oracle_log = done.value
            break
    assert log.steps_per_frame == oracle_log.steps_per_frame


@pytest.mark.parametrize("channels", ["heat,thickness,shade,unused", "heat,thickness,normal_x,normal_y"])
def test_rendered_frames_of_a_rollout_match_the_oracle(device, oracle32, channels):
    from oracle.render import render_oracle
    from plume_amd.recipe.derived import derive
    from plume_amd.render.camera import make_camera, make_domain_view
    from plume_amd.render.frame_renderer import FrameRenderer

    recipe = puff_recipe()
    grid, reference, _, _ = stepped(recipe, device, oracle32, 6)
    derived = derive(recipe)
    view = make_domain_view(make_camera(20.0, 30.0), derived.domain_min_m, derived.size_m, recipe.domain.cell_m,
                            recipe.render.px_per_cell)
    # render the oracle's state on both sides, so the comparison isolates the render kernels
    grid.smoke.data.copy_(torch.from_numpy(reference.smoke.data))
    grid.temperature.data.copy_(torch.from_numpy(reference.temperature.data))
    settings = recipe.render
    image = FrameRenderer(view, settings, derived.cells, channels, device).render(grid)
    expected = render_oracle(oracle32, reference, view, settings, channels)
    assert image.shape == expected.shape == (view.height_px, view.width_px, 5)
    for channel in range(5):
        error = rel_l2(image[..., channel], expected[..., channel])
        assert error < 1e-4, (channel, error)


def test_tiny_recipe_file_loads_and_runs(device, tmp_path):
    from plume_amd.recipe.derived import derive
    from plume_amd.recipe.load import load_recipe
    from plume_amd.sim.run import simulate

    path = tmp_path / "tiny.toml"
    path.write_text(TINY)
    recipe = load_recipe(path)
    log = simulate(recipe, derive(recipe), device, lambda frame, grid: None)
    assert len(log.steps_per_frame) == 4 and min(log.steps_per_frame) >= 3
