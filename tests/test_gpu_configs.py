"""Parity at the benchmark workloads' own sizes: the HIP path against the oracle on every configuration.

- 128 x 192 x 128 (the headline substep): every operator of the substep on seeded fields (velocities U(-1, 1)
  m/s, smoke and temperature U(0, 1), seeds 20261203 + k) with the default faces and with all faces closed:
  bit-identical to the FP32 oracle where no transcendental is involved (advection, buoyancy, curl, divergence,
  sweeps, residual, restriction, prolongation, projection, the tolerance solve), within 1e-5 of the field's
  scale otherwise; four whole substeps of the puff within 1e-5 relative L2; one 256 x 256 px, 192-sample frame
  with a 64-step light volume within 2e-5 of the image's scale.
- 64 x 96 x 64 (the puff run on the oracle's path): frames of the recipe run with the same step counts and
  fields within 1e-5 relative L2, and 128 x 128 heat / thickness / shade images.
- 64 x 96 x 64, 32 substeps (the differentiable rollout): directional derivatives of the adjoint against FP64
  central differences of the oracle (tests/golden/adjoint_fd_64x96x64.json), within 1e-3 relative.
- 32 members of 48³ to 128 x 192 x 128 (the batched ensemble): replayed frames bit-equal to each member
  stepped alone, every solve's residual gate met, one member against the oracle.
- 256 x 384 x 256 (the mixed-precision projection): every storage form reaches the common target and its
  post-projection divergence is no worse than the plain FP32 solve's.

The reference's bounds: tests/sim/test_multigrid.py:52-75 (convergence), tests/cli/test_cpu_gpu.py:21-33
(CPU-against-GPU images).
"""

import json
from pathlib import Path

import numpy as np
import pytest
import torch

from support import ALL_CLOSED, DEFAULT_FACES, bit_equal, compare, gpu_grid, load_state, oracle_grid, random_state, \
    rng, to_np

pytestmark = [pytest.mark.gpu, pytest.mark.config]

C2 = (128, 192, 128)
C1 = (64, 96, 64)
CELL_M = 0.01
DT = 0.005  # max |u| dt / cell <= 0.5 cells for U(-1, 1) m/s at 1 cm cells
FACES = [pytest.param(DEFAULT_FACES, id="default-faces"), pytest.param(ALL_CLOSED, id="closed")]
GOLDEN = Path(__file__).resolve().parent / "golden"


def rel_l2(actual, expected):
    actual = np.asarray(actual, dtype=np.float64)
    expected = np.asarray(expected, dtype=np.float64)
    return float(np.linalg.norm(actual - expected) / max(np.linalg.norm(expected), 1e-30))


def c2_pair(device, oracle32, faces, k, ambient=293.0):
    print(f"[config] 128x192x128 faces={faces} seed={20261203 + k}", flush=True)
    gpu = gpu_grid(C2, CELL_M, faces, ambient, device)
    orc = oracle_grid(oracle32, C2, CELL_M, faces, ambient)
    load_state(gpu, orc, random_state(C2, k))
    return gpu, orc


# ---------------------------------------------------------------------- 128 x 192 x 128, per operator

@pytest.mark.parametrize("faces", FACES)
def test_c2_advection_is_bit_exact(device, oracle32, faces):
    from oracle.sim import advect_maccormack as oracle_advect
    from plume_amd.sim.advect import advect_scalars, advect_velocity

    gpu, orc = c2_pair(device, oracle32, faces, 0)
    advect_velocity(gpu, DT)
    velocity = tuple(f.data for f in orc.velocity_fields())
    for name in ("u", "v", "w"):
        oracle_advect(getattr(orc, name), velocity, orc, DT)
        assert bit_equal(to_np(getattr(gpu, name).next), getattr(orc, name).next), name
    advect_scalars(gpu, tuple(f.data for f in gpu.velocity_fields()), DT)
    for name in ("smoke", "temperature"):
        oracle_advect(getattr(orc, name), velocity, orc, DT)
        assert bit_equal(to_np(getattr(gpu, name).next), getattr(orc, name).next), name


@pytest.mark.parametrize("faces", FACES)
def test_c2_forces_match(device, oracle32, faces):
    from oracle.sim import VorticityConfinement as OracleConfinement, apply_buoyancy as oracle_buoyancy
    from plume_amd.sim.forces import apply_buoyancy
    from plume_amd.sim.vorticity import VorticityConfinement

    gpu, orc = c2_pair(device, oracle32, faces, 1)
    # the benchmark's normalised buoyancy: ambient 1, heat_buoyancy beta / g, smoke_weight alpha / g
    apply_buoyancy(gpu, 9.81, 1.0 / 9.81, 0.1 / 9.81, 1.0, DT)
    oracle_buoyancy(orc, 9.81, 1.0 / 9.81, 0.1 / 9.81, 1.0, DT)
    assert bit_equal(to_np(gpu.v.data), orc.v.data)
    confinement, oconf = VorticityConfinement(gpu), OracleConfinement(orc)
    confinement.apply(gpu, 0.3, DT)
    oconf.apply(orc, 0.3, DT)
    assert bit_equal(to_np(confinement.curl), oconf.curl)
    compare(to_np(confinement.push), oconf.push, what="push")
    for name in ("u", "v", "w"):
        compare(to_np(getattr(gpu, name).data), getattr(orc, name).data, what=name)


@pytest.mark.parametrize("faces", FACES)
def test_c2_pressure_operators_are_bit_exact(device, oracle32, faces):
    from plume_amd._lib import call, int3, ptr, stream
    from plume_amd.sim.multigrid import build_levels, residual
    from plume_amd.sim.pressure import compute_divergence, red_black_sweeps, subtract_gradient, zero_closed_faces

    gpu, orc = c2_pair(device, oracle32, faces, 2)
    zero_closed_faces(gpu)
    for axis, f in enumerate(orc.velocity_fields()):
        oracle32.zero_closed_faces(orc.info, f.data, axis)
    compute_divergence(gpu)
    oracle32.divergence(orc.info, *(f.data for f in orc.velocity_fields()), orc.divergence)
    assert bit_equal(to_np(gpu.divergence), orc.divergence)
    q0 = rng(11).uniform(-1e-3, 1e-3, C2).astype(np.float32)
    gpu.pressure.copy_(torch.from_numpy(q0))
    orc.pressure[...] = q0
    scratch = torch.empty_like(gpu.pressure)
    red_black_sweeps(gpu.pressure, gpu.divergence, gpu.info, gpu.cells, device, 2, 1.0, scratch)
    oracle32.rb_sweeps(orc.info, orc.pressure, orc.divergence, 2, 1.0)
    assert bit_equal(to_np(gpu.pressure), orc.pressure), "fused sweeps"
    levels = build_levels(gpu)
    r = to_np(residual(levels[0]))
    expected = np.zeros_like(orc.pressure)
    oracle32.residual(orc.info, orc.pressure, orc.divergence, expected)
    assert bit_equal(r, expected)
    coarse = levels[1]
    call("plm_restrict", int3(C2), ptr(levels[0].residual), int3(coarse.cells), ptr(coarse.rhs), stream())
    expected_coarse = np.zeros(coarse.cells, dtype=np.float32)
    oracle32.restrict(expected, expected_coarse)
    assert bit_equal(to_np(coarse.rhs), expected_coarse)
    call("plm_prolong_add", int3(coarse.cells), ptr(coarse.rhs), int3(C2), ptr(gpu.pressure), stream())
    oracle32.prolong_add(expected_coarse, orc.pressure)
    assert bit_equal(to_np(gpu.pressure), orc.pressure)
    subtract_gradient(gpu)
    for axis, f in enumerate(orc.velocity_fields()):
        oracle32.subtract_gradient(orc.info, f.data, orc.pressure, axis)
    for name in ("u", "v", "w"):
        assert bit_equal(to_np(getattr(gpu, name).data), getattr(orc, name).data), name


@pytest.mark.parametrize("faces", FACES)
def test_c2_projection_and_tolerance_solve_are_bit_exact(device, oracle32, faces):
    """Two fixed V(2,2) cycles, then the converge-to-1e-5 mode from that warm start: the same cycle count as
    the oracle's, the same pressure bit for bit, a relative residual at or below 1e-5 and walls exactly 0."""
    from oracle.sim import Multigrid as OracleMultigrid, project as oracle_project, relative_residual as orc_ratio
    from plume_amd.sim.multigrid import Multigrid, relative_residual
    from plume_amd.sim.pressure import project

    gpu, orc = c2_pair(device, oracle32, faces, 3)
    project(gpu, Multigrid(gpu, 2, 2))
    oracle_project(orc, OracleMultigrid(orc, 2, 2))
    assert bit_equal(to_np(gpu.pressure), orc.pressure)
    for name in ("u", "v", "w"):
        assert bit_equal(to_np(getattr(gpu, name).data), getattr(orc, name).data), name
    # a fresh divergent field, solved to the target from the warm start above
    load_state(gpu, orc, {k: v for k, v in random_state(C2, 4).items() if k in ("u", "v", "w")} | {
        "smoke": to_np(gpu.smoke.data), "temperature": to_np(gpu.temperature.data)})
    solve, osolve = Multigrid(gpu, 2, 2, tolerance=1e-5, max_cycles=24), OracleMultigrid(orc, 2, 2, 1e-5, 24)
    project(gpu, solve)
    oracle_project(orc, osolve)
    assert solve.last_cycles() == osolve.cycles_log[-1] >= 2
    assert bit_equal(to_np(gpu.pressure), orc.pressure)
    assert relative_residual(gpu.info, gpu.pressure, gpu.divergence) <= 1e-5
    assert orc_ratio(orc) <= 1e-5
    assert solve.last_relative_residual() == pytest.approx(orc_ratio(orc), rel=1e-9)
    if faces == ALL_CLOSED:
        assert not gpu.u.data[0].any() and not gpu.u.data[-1].any() and not gpu.v.data[:, 0].any()


def puff_recipe(cells, radius_cells):
    from plume_amd.recipe.load import fill_dependent_defaults, parse_recipe
    from plume_amd.sim.scenes import puff_table

    return fill_dependent_defaults(parse_recipe(puff_table(cells, radius_cells=radius_cells)), "puff")


def test_c2_four_substeps_match_the_oracle(device, oracle32):
    """The headline scene (radius-12 emitter active, buoyancy alpha 0.1 / beta 1, confinement 0.3, multigrid to
    1e-5) for 4 substeps from a seeded moving state (velocities U(-0.5, 0.5) m/s, smoke U(0, 1), temperature
    1 + U(0, 1) over the ambient 1)."""
    from oracle.sim import Stepper as OracleStepper, make_grid as oracle_make_grid
    from plume_amd.recipe.derived import derive
    from plume_amd.sim.grid import make_grid
    from plume_amd.sim.stepper import Stepper

    recipe = puff_recipe(C2, 12)
    derived = derive(recipe)
    grid = make_grid(derived, recipe.domain.cell_m, recipe.domain.faces, recipe.domain.ambient_k, device)
    reference = oracle_make_grid(oracle32, derived, recipe.domain.cell_m, recipe.domain.faces, recipe.domain.ambient_k)
    state = random_state(C2, 5)
    for name in ("u", "v", "w"):
        state[name] = (0.5 * state[name]).astype(np.float32)
    state["temperature"] = (1.0 + state["temperature"]).astype(np.float32)
    load_state(grid, reference, state)
    stepper, oracle_stepper = Stepper(grid, recipe), OracleStepper(reference, recipe)
    for index in range(4):
        print(f"[config] 128x192x128 substep {index}", flush=True)
        stepper.step(index * DT, DT)
        oracle_stepper.step(index * DT, DT)
        assert stepper.pressure_solve.last_cycles() == oracle_stepper.pressure_solve.cycles_log[-1]
    torch.cuda.synchronize()
    for name in ("u", "v", "w", "smoke", "temperature"):
        error = rel_l2(to_np(getattr(grid, name).data), getattr(reference, name).data)
        assert error < 1e-5, (name, error)
    assert rel_l2(to_np(grid.pressure), reference.pressure) < 1e-5


def test_c2_frame_matches_the_oracle(device, oracle32):
    """256 x 256 px, 192 samples per ray, the light volume with 64 steps per cell (the benchmark's frame), shade
    layout, on seeded smoke and temperature."""
    from oracle.render import render_oracle
    from plume_amd.recipe.schema import SHADE_LAYOUT
    from plume_amd.render.camera import make_camera, square_domain_view
    from plume_amd.render.frame_renderer import FrameRenderer

    recipe = puff_recipe(C2, 12)
    gpu, orc = c2_pair(device, oracle32, DEFAULT_FACES, 6, ambient=1.0)
    # smoke U(0, 0.05): an optical depth of about 1 across the box, so light and shade vary over the image
    smoke = (0.05 * random_state(C2, 6)["smoke"]).astype(np.float32)
    temperature = (1.0 + random_state(C2, 7)["temperature"]).astype(np.float32)
    for name, values in (("smoke", smoke), ("temperature", temperature)):
        getattr(gpu, name).data.copy_(torch.from_numpy(values))
        getattr(orc, name).data[...] = values
    settings = recipe.render
    camera = make_camera(recipe.camera.elevation_deg, recipe.camera.azimuth_deg)
    view = square_domain_view(camera, orc.origin_m, tuple(n * CELL_M for n in C2), 256)
    renderer = FrameRenderer(view, settings, C2, SHADE_LAYOUT, device, max_steps=192, light_steps=64)
    image = renderer.render(gpu)
    expected = render_oracle(oracle32, orc, view, settings, SHADE_LAYOUT, max_steps=192, light_steps=64)
    for channel in range(5):
        compare(image[..., channel], expected[..., channel], rel=2e-5, atol=1e-6, what=f"channel {channel}")
    assert float(np.abs(expected[..., 1]).max()) > 0.1


# ---------------------------------------------------------------------- 64 x 96 x 64, the puff run

def test_c1_puff_frames_and_images_match_the_oracle(device, oracle32):
    """Six frames of the puff recipe at 64 x 96 x 64 (radius-6 emitter), run by each side's frame loop: the
    same substeps per frame and fields within 1e-5 relative L2 after each; then 128 x 128 heat, thickness and
    shade images of the last frame: rendered from the same fields within 2e-5 of each channel's scale, and
    each side's image of its own rollout within 1e-4 relative L2."""
    from oracle.render import render_oracle
    from oracle.sim import simulate_frames as oracle_frames
    from plume_amd.recipe.derived import derive
    from plume_amd.recipe.schema import SHADE_LAYOUT
    from plume_amd.render.camera import make_camera, square_domain_view
    from plume_amd.render.frame_renderer import FrameRenderer
    from plume_amd.sim.run import simulate_frames

    recipe = puff_recipe(C1, 6)
    derived = derive(recipe)
    frames, oframes = simulate_frames(recipe, derived, device), oracle_frames(oracle32, recipe, derived)
    for frame in range(6):
        index, grid = next(frames)
        oindex, reference = next(oframes)
        assert index == oindex == frame
        print(f"[config] 64x96x64 frame {frame}", flush=True)
        for name in ("u", "v", "w", "smoke", "temperature"):
            error = rel_l2(to_np(getattr(grid, name).data), getattr(reference, name).data)
            assert error < 1e-5, (frame, name, error)
    assert float(grid.smoke.data.sum()) > 0.0
    settings = recipe.render
    camera = make_camera(recipe.camera.elevation_deg, recipe.camera.azimuth_deg)
    view = square_domain_view(camera, derived.domain_min_m, derived.size_m, 128)
    own = FrameRenderer(view, settings, C1, SHADE_LAYOUT, device).render(grid)
    theirs = render_oracle(oracle32, reference, view, settings, SHADE_LAYOUT)
    for channel in range(3):
        assert rel_l2(own[..., channel], theirs[..., channel]) < 1e-4, channel
    # the same fields on both sides
    for name in ("smoke", "temperature"):
        getattr(reference, name).data[...] = to_np(getattr(grid, name).data)
    theirs = render_oracle(oracle32, reference, view, settings, SHADE_LAYOUT)
    for channel in range(3):
        compare(own[..., channel], theirs[..., channel], rel=2e-5, atol=1e-6, what=f"channel {channel}")
    assert float(np.abs(theirs[..., 0]).max()) > 0.0 and float(np.abs(theirs[..., 1]).max()) > 0.05


# ---------------------------------------------------------------------- 64 x 96 x 64, the differentiable rollout

def test_c3_gradient_matches_fp64_central_differences(device):
    """Config 3 at its size: the adjoint's directional derivatives of <W, images> (each source scaled by itself,
    alpha) against the FP64 oracle's central differences at h = 1e-3 (within 1e-3 relative, the benchmark's
    bar) and h = 1e-5 (reported)."""
    from golden.make_adjoint_fd import ALPHA, CHECKPOINT, SUBSTEPS, weights
    from plume_amd.sim.adjoint import Sources, rollout
    from plume_amd.sim.scenes import differentiable_puff

    golden = json.loads((GOLDEN / "adjoint_fd_64x96x64.json").read_text())
    recipe, derived, sim, sources = differentiable_puff(device, C1, 128, 192)
    params = [t.clone().requires_grad_(True) for t in sources.tensors()]
    alpha = torch.tensor(ALPHA, dtype=torch.float64, requires_grad=True)
    images = rollout(Sources(*params), alpha, sim, SUBSTEPS, CHECKPOINT)
    w = torch.from_numpy(weights()).to(device)
    assert tuple(images.shape) == tuple(w.shape)
    sums = images.detach().sum(dim=(1, 2)).double().cpu().numpy()
    np.testing.assert_allclose(sums, np.array(golden["image_sums"]), rtol=1e-4)
    images.backward(w)
    derivatives = {}
    for k, name in enumerate(("smoke", "temperature", "lift")):
        derivatives[name] = float((params[k].grad.double() * params[k].detach().double()).sum())
    derivatives["alpha"] = float(alpha.grad) * ALPHA
    print("[config] adjoint vs FP64 central differences:", {
        name: (value, golden["central_differences"][name]) for name, value in derivatives.items()}, flush=True)
    for name, value in derivatives.items():
        fd = golden["central_differences"][name]
        assert abs(value - fd["0.001"]) <= 1e-3 * abs(fd["0.001"]), (name, value, fd)


# ---------------------------------------------------------------------- the 32-member ensemble

def test_c4_ensemble_replays_equal_eager_members_and_meet_their_gates(device, oracle32):
    """32 seeded members over 48³, 64 x 96 x 64, 96 x 128 x 96 and 128 x 192 x 128 (emitters off after 0.1 s so
    that frames 3 on replay): each replayed frame equals every member stepped alone and eagerly, bit for bit,
    no solve in a replayed graph ended above its tolerance, and the first 48³ member follows the FP32 oracle
    stepped with the ensemble's substep counts within 1e-5 relative L2."""
    from oracle.sim import Stepper as OracleStepper, make_grid as oracle_make_grid
    from plume_amd.recipe.derived import derive
    from plume_amd.recipe.load import fill_dependent_defaults, parse_recipe
    from plume_amd.sim.ensemble import Ensemble, ensemble_tables
    from plume_amd.sim.grid import make_grid
    from plume_amd.sim.multigrid import relative_residual
    from plume_amd.sim.stepper import Stepper

    frames = 6
    tables = ensemble_tables(32, seed=20261204, frames=frames)
    for table in tables:
        table["emitter"][0]["end_s"] = 0.1
    ens = Ensemble(tables, device)
    ens.autotune()
    sizes = {m.grid.cells for m in ens.members}
    assert len(ens.members) == 32 and len(sizes) >= 3, sizes
    alone = []
    for table in tables:
        recipe = fill_dependent_defaults(parse_recipe(table), "ensemble")
        derived = derive(recipe)
        grid = make_grid(derived, recipe.domain.cell_m, recipe.domain.faces, recipe.domain.ambient_k, device)
        alone.append((grid, Stepper(grid, recipe)))
    probe = next(i for i, m in enumerate(ens.members) if m.grid.cells == (48, 48, 48))
    orecipe = ens.members[probe].recipe
    reference = oracle_make_grid(oracle32, ens.members[probe].derived, orecipe.domain.cell_m, orecipe.domain.faces,
                                 orecipe.domain.ambient_k)
    ostepper = OracleStepper(reference, orecipe)
    replayed = 0
    for frame in range(frames):
        steps = ens.run_frame(replay=True)
        replayed += ens.log[-1][2]
        dt = ens.frame_s / steps
        print(f"[config] ensemble frame {frame}: {steps} substeps, replayed {ens.log[-1][2]}", flush=True)
        for grid, stepper in alone:
            for index in range(steps):
                stepper.step(frame * ens.frame_s + index * dt, dt)
        for index in range(steps):
            ostepper.step(frame * ens.frame_s + index * dt, dt)
        for m, (grid, stepper) in zip(ens.members, alone):
            for name in ("u", "v", "w", "smoke", "temperature"):
                a, b = getattr(m.grid, name).data, getattr(grid, name).data
                assert torch.equal(a.view(torch.int32), b.view(torch.int32)), (frame, name, m.grid.cells)
            assert torch.equal(m.grid.pressure.view(torch.int32), grid.pressure.view(torch.int32))
            assert relative_residual(m.grid.info, m.grid.pressure, m.grid.divergence) <= m.recipe.solver.multigrid_tolerance
    assert replayed >= 2 and not ens.redone, (ens.log, ens.redone)
    member = ens.members[probe].grid
    for name in ("u", "v", "w", "smoke", "temperature"):
        error = rel_l2(to_np(getattr(member, name).data), getattr(reference, name).data)
        assert error < 1e-5, (name, error)


# ---------------------------------------------------------------------- 256 x 384 x 256, mixed precision

def test_c5_every_storage_form_reaches_the_target_at_full_size(device):
    """The seeded divergent field at 256 x 384 x 256 (64 Gaussian vortex blobs plus the emitter source), solved
    from q = 0 by plain FP32 V(2,2) cycles and by FP32 refinement around BF16 and FP16 multigrid with the FP64
    8 x 12 x 8 coarsest level: each reaches the common target max(1e-6, 2 x the FP32 floor) in FP64 relative
    residual, the mixed forms' post-projection divergence is within 1.25 x the plain solve's, and their
    pressures agree with it to 1e-3 relative L2."""
    from plume_amd.sim.mixed import divergent_problem, fp32_floor, solve_forms

    problem = divergent_problem((256, 384, 256), device)
    floor = fp32_floor(problem.grid)
    target = max(1e-6, 2.0 * floor)
    print(f"[config] 256x384x256 FP32 floor {floor:.3g}, target {target:.3g}", flush=True)
    forms = solve_forms(problem, target)
    for name, (entry, _) in forms.items():
        print(f"[config] {name}: {entry}", flush=True)
        assert entry["relative_residual"] <= target, (name, entry)
    plain_div = forms["f32"][0]["post_projection_divergence"]
    plain_q = forms["f32"][1]
    for name in ("bf16", "f16"):
        entry, q = forms[name]
        assert entry["post_projection_divergence"] <= 1.25 * plain_div, (name, entry, plain_div)
        assert rel_l2(to_np(q), to_np(plain_q)) < 1e-3, name
