"""Per-operator parity: each HIP kernel against the FP32 oracle on identical seeded inputs.

Bar: max |gpu - oracle| <= 1e-5 * max |oracle| (FP32 operators); the integer flags and the noise hash must
match bit for bit. Operators without transcendentals are also expected to be bit-identical (both sides
round every float operation once, in the reference's order), which the tests report and, for the
projection and multigrid, assert.
"""

import ctypes

import numpy as np
import pytest
import torch

from support import (ALL_CLOSED, ALL_OPEN, DEFAULT_FACES, bit_equal, compare, gpu_grid, load_state, oracle_grid,
                     random_state, rng, to_np)

pytestmark = pytest.mark.gpu

GRIDS = [
    pytest.param((16, 16, 16), ALL_CLOSED, id="16cubed-closed"),
    pytest.param((31, 33, 29), DEFAULT_FACES, id="odd-default"),
    pytest.param((31, 33, 29), ALL_CLOSED, id="odd-closed"),
    pytest.param((17, 12, 9), ALL_OPEN, id="small-open"),
]
CELL_M = 0.01
DT = 0.005  # max |u| dt / cell <= 0.5 cells for U(-1, 1) m/s at 1 cm cells


def pair(device, oracle32, cells, faces, k=0, ambient=293.0):
    gpu = gpu_grid(cells, CELL_M, faces, ambient, device)
    orc = oracle_grid(oracle32, cells, CELL_M, faces, ambient)
    state = random_state(cells, k)
    load_state(gpu, orc, state)
    return gpu, orc


# ---------------------------------------------------------------------- boundary flags, noise hash

@pytest.mark.parametrize("cells, faces", GRIDS)
def test_boundary_flags_are_bit_exact(device, oracle32, cells, faces):
    from oracle.sim import level_infos
    from plume_amd.sim.multigrid import build_levels
    from plume_amd.sim.pressure import boundary_flags

    gpu, orc = pair(device, oracle32, cells, faces)
    levels = build_levels(gpu)
    oracle_levels = level_infos(oracle32, cells, CELL_M, orc.origin_m, orc.open_lo, orc.open_hi)
    assert [lv.cells for lv in levels] == [c for c, _ in oracle_levels]
    for level, (level_cells, info) in zip(levels, oracle_levels):
        expected = np.zeros(level_cells, dtype=np.uint16)
        oracle32.build_flags(info, expected)
        actual = to_np(boundary_flags(level.info, device)).view(np.uint16)
        assert bit_equal(actual, expected), level_cells
        assert tuple(level.info.fill_hi) == tuple(info.fill_hi)
        assert level.info.ghost_gap == info.ghost_gap
        assert level.info.cell_m == info.cell_m


def test_noise_hash_and_lattice_are_bit_exact(device, oracle32):
    from plume_amd.sim.noise import lattice_values, value_noise

    r = rng(7)
    cells = r.integers(-5000, 5000, size=(4096, 3)).astype(np.int32)
    for seed, stream in ((0, 0), (7, 1), (2**31 - 1, (1 << 24) + 5)):
        actual = to_np(lattice_values(seed, stream, torch.from_numpy(cells).to(device)))
        expected = np.array([oracle32.lattice_value(seed, stream, *c) for c in cells], dtype=np.float32)
        assert bit_equal(actual, expected)
    points = r.uniform(-40, 40, size=(4096, 3)).astype(np.float32)
    actual = to_np(value_noise(3, 2, torch.from_numpy(points).to(device)))
    expected = np.array([oracle32.value_noise(3, 2, p) for p in points], dtype=np.float32)
    assert bit_equal(actual, expected)


# ---------------------------------------------------------------------- advection

@pytest.mark.parametrize("cells, faces", GRIDS)
@pytest.mark.parametrize("fuse_back", [False, True], ids=["three-pass", "fused-back"])
def test_maccormack_matches_oracle(device, oracle32, cells, faces, fuse_back):
    from oracle.sim import advect_maccormack as oracle_advect
    from plume_amd.sim.advect import advect_maccormack

    gpu, orc = pair(device, oracle32, cells, faces)
    for name in ("smoke", "temperature", "u", "v", "w"):
        field, ofield = getattr(gpu, name), getattr(orc, name)
        velocity = tuple(f.data for f in gpu.velocity_fields())
        advect_maccormack(field, velocity, gpu, DT, fuse_back=fuse_back)
        oracle_advect(ofield, tuple(f.data for f in orc.velocity_fields()), orc, DT)
        compare(to_np(field.hat), ofield.hat, what=f"{name} hat")
        compare(to_np(field.next), ofield.next, what=f"{name} next")
        assert bit_equal(to_np(field.next), ofield.next), f"{name}: MacCormack not bit-identical"


@pytest.mark.parametrize("cells, faces", GRIDS)
def test_semi_lagrangian_matches_oracle(device, oracle32, cells, faces):
    from oracle.sim import advect_semi_lagrangian as oracle_sl
    from plume_amd.sim.advect import advect_semi_lagrangian

    gpu, orc = pair(device, oracle32, cells, faces)
    advect_semi_lagrangian(gpu.smoke, tuple(f.data for f in gpu.velocity_fields()), gpu, 3 * DT)
    oracle_sl(orc.smoke, tuple(f.data for f in orc.velocity_fields()), orc, 3 * DT)
    assert bit_equal(to_np(gpu.smoke.next), orc.smoke.next)


@pytest.mark.parametrize("cells, faces", GRIDS)
def test_fused_velocity_advection_matches_per_field_oracle(device, oracle32, cells, faces):
    from oracle.sim import advect_maccormack as oracle_advect
    from plume_amd.sim.advect import advect_velocity

    gpu, orc = pair(device, oracle32, cells, faces)
    advect_velocity(gpu, DT)
    velocity = tuple(f.data for f in orc.velocity_fields())
    for field, ofield in zip(gpu.velocity_fields(), orc.velocity_fields()):
        oracle_advect(ofield, velocity, orc, DT)
        compare(to_np(field.next), ofield.next, what="velocity next")
        assert bit_equal(to_np(field.next), ofield.next)


@pytest.mark.parametrize("cells, faces", GRIDS)
def test_velocity_advection_with_carried_bounds_is_bit_exact(device, oracle32, cells, faces):
    from oracle.sim import advect_maccormack as oracle_advect
    from plume_amd.sim.advect import advect_velocity

    gpu, orc = pair(device, oracle32, cells, faces, k=5)
    gpu.u.data[2, 3, 4] = -0.0  # a -0.0 operand on an aligned axis takes its partner's sign
    orc.u.data[2, 3, 4] = -0.0
    advect_velocity(gpu, 2 * DT, carry_bounds=True)
    velocity = tuple(f.data for f in orc.velocity_fields())
    for field, ofield in zip(gpu.velocity_fields(), orc.velocity_fields()):
        oracle_advect(ofield, velocity, orc, 2 * DT)
        assert bit_equal(to_np(field.hat), ofield.hat)
        assert bit_equal(to_np(field.next), ofield.next)


@pytest.mark.parametrize("cells, faces", GRIDS)
@pytest.mark.parametrize("fields", [("smoke", "temperature"), ("smoke",)], ids=["pair", "single"])
def test_scalar_advection_sharing_one_trace_is_bit_exact(device, oracle32, cells, faces, fields):
    from oracle.sim import advect_maccormack as oracle_advect
    from plume_amd.sim.advect import advect_scalars

    gpu, orc = pair(device, oracle32, cells, faces, k=6, ambient=293.0)
    velocity = tuple(f.data for f in gpu.velocity_fields())
    advect_scalars(gpu, velocity, 3 * DT, [getattr(gpu, name) for name in fields])
    for name in fields:
        ofield = getattr(orc, name)
        oracle_advect(ofield, tuple(f.data for f in orc.velocity_fields()), orc, 3 * DT)
        assert bit_equal(to_np(getattr(gpu, name).hat), ofield.hat), name
        assert bit_equal(to_np(getattr(gpu, name).next), ofield.next), name


# ---------------------------------------------------------------------- forces

@pytest.mark.parametrize("cells, faces", GRIDS)
def test_buoyancy_matches_oracle(device, oracle32, cells, faces):
    from oracle.sim import apply_buoyancy as oracle_buoyancy
    from plume_amd.sim.forces import apply_buoyancy

    gpu, orc = pair(device, oracle32, cells, faces)
    apply_buoyancy(gpu, 9.81, 1.0 / 9.81, 0.1 / 9.81, 1.0, DT)
    oracle_buoyancy(orc, 9.81, 1.0 / 9.81, 0.1 / 9.81, 1.0, DT)
    assert bit_equal(to_np(gpu.v.data), orc.v.data)


@pytest.mark.parametrize("cells, faces", GRIDS[:2])
def test_drag_and_dissipation_match_oracle(device, oracle32, cells, faces):
    from oracle.sim import apply_drag as oracle_drag, dissipate as oracle_dissipate
    from plume_amd.sim.dissipate import dissipate
    from plume_amd.sim.forces import apply_drag

    gpu, orc = pair(device, oracle32, cells, faces)
    apply_drag(gpu, (0.5, 0.0, -0.25), 0.5, DT)
    oracle_drag(orc, (0.5, 0.0, -0.25), 0.5, DT)
    dissipate(gpu, 1.6, 1.0, 293.0, DT)
    oracle_dissipate(orc, 1.6, 1.0, 293.0, DT)
    for name in ("u", "v", "w", "smoke", "temperature"):
        assert bit_equal(to_np(getattr(gpu, name).data), getattr(orc, name).data), name


@pytest.mark.parametrize("cells, faces", GRIDS)
@pytest.mark.parametrize("fuse", [False, True], ids=["four-pass", "fused"])
def test_confinement_matches_oracle(device, oracle32, cells, faces, fuse):
    from oracle.sim import VorticityConfinement as OracleConfinement
    from plume_amd.sim.vorticity import VorticityConfinement

    gpu, orc = pair(device, oracle32, cells, faces)
    confinement, oconf = VorticityConfinement(gpu, fuse=fuse), OracleConfinement(orc)
    confinement.apply(gpu, 0.3, DT)
    oconf.apply(orc, 0.3, DT)
    compare(to_np(confinement.curl), oconf.curl, what="curl")
    assert bit_equal(to_np(confinement.curl), oconf.curl)
    compare(to_np(confinement.push), oconf.push, what="push")
    for name in ("u", "v", "w"):
        compare(to_np(getattr(gpu, name).data), getattr(orc, name).data, what=name)


def emitter_settings(**changes):
    from support import tiny_recipe

    emitter = {"name": "e", "center_m": [0.0, 0.12, 0.0], "radius_m": 0.06, "end_s": 1.0, "smoke_per_s": 2.0,
               "temperature_k": 900.0, "velocity_mps": [0.1, 0.5, -0.2], "edge_softness": 0.6,
               "noise": {"scale_m": 0.03, "strength": 0.7}}
    emitter.update(changes)
    return tiny_recipe(emitter=[emitter]).emitter


@pytest.mark.parametrize("cells, faces", GRIDS[:2])
@pytest.mark.parametrize("variant", ["sphere-noise", "box-outward"])
def test_emit_matches_oracle(device, oracle32, cells, faces, variant):
    from oracle.sim import emit as oracle_emit, emitter_params as oracle_params
    from plume_amd.sim.emitters import emit, emitter_params

    changes = {} if variant == "sphere-noise" else {"shape": "box", "size_m": [0.1, 0.08, 0.12],
                                                    "velocity_mode": "outward"}
    emitters = emitter_settings(**changes)
    gpu, orc = pair(device, oracle32, cells, faces, ambient=293.0)
    emit(gpu, emitter_params(emitters, device), 7, 0.1, DT)
    oracle_emit(orc, oracle_params(emitters), 7, 0.1, DT)
    for name in ("smoke", "temperature", "u", "v", "w"):
        compare(to_np(getattr(gpu, name).data), getattr(orc, name).data, what=name)


@pytest.mark.parametrize("case", ["sphere-noise", "box-outward", "corner-and-inactive", "outside-the-grid"])
def test_emit_over_the_emitters_box_equals_the_whole_grid_launch(device, oracle32, case):
    """plm_emit_box over the active emitters' box (plume_amd.sim.emitters.emit) leaves every field bit for bit as
    plm_emit over the whole grid: an emitter in the middle, a box emitter with outward velocity, an emitter
    overhanging a corner beside one not yet active, and an emitter wholly outside the grid."""
    from support import tiny_recipe

    from plume_amd.sim.emitters import emit, emit_box, emit_whole_grid, emitter_params

    base = {"name": "e", "center_m": [0.0, 0.12, 0.0], "radius_m": 0.06, "end_s": 1.0, "smoke_per_s": 2.0,
            "temperature_k": 900.0, "velocity_mps": [0.1, 0.5, -0.2], "edge_softness": 0.6,
            "noise": {"scale_m": 0.03, "strength": 0.7}}
    if case == "sphere-noise":
        emitters = [base]
    elif case == "box-outward":
        emitters = [dict(base, shape="box", size_m=[0.1, 0.08, 0.12], velocity_mode="outward")]
    elif case == "corner-and-inactive":
        emitters = [dict(base, center_m=[0.15, 0.02, -0.14], radius_m=0.05),
                    dict(base, name="later", center_m=[-0.1, 0.2, 0.1], start_s=0.5)]
    else:
        emitters = [dict(base, center_m=[1.0, 0.1, 0.0], radius_m=0.05)]
    settings = tiny_recipe(emitter=emitters).emitter
    params = emitter_params(settings, device)
    gpu_box, _ = pair(device, oracle32, (31, 33, 29), DEFAULT_FACES, k=3)
    gpu_all, _ = pair(device, oracle32, (31, 33, 29), DEFAULT_FACES, k=3)
    box = emit_box(gpu_box.info, params.params, 0.1)
    if case == "outside-the-grid":
        assert box is None
    else:
        lo, hi = box
        assert all(0 <= lo[a] < hi[a] <= gpu_box.info.dims[a] for a in range(3))
        assert any(hi[a] - lo[a] < gpu_box.info.dims[a] for a in range(3)), "the box should not be the whole grid"
    names = ("smoke", "temperature", "u", "v", "w")
    before = {name: to_np(getattr(gpu_all, name).data) for name in names}
    emit(gpu_box, params, 7, 0.1, DT)
    emit_whole_grid(gpu_all, params, 7, 0.1, DT)
    changed = False
    for name in names:
        a, b = to_np(getattr(gpu_box, name).data), to_np(getattr(gpu_all, name).data)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), name
        changed = changed or not np.array_equal(b.view(np.uint32), before[name].view(np.uint32))
    assert changed == (case != "outside-the-grid")


@pytest.mark.parametrize("cells, faces", GRIDS[:2])
def test_turbulence_matches_oracle(device, oracle32, cells, faces):
    from oracle.sim import apply_turbulence as oracle_turbulence
    from plume_amd.sim.turbulence import apply_turbulence

    gpu, orc = pair(device, oracle32, cells, faces)
    apply_turbulence(gpu, 7, 0.05, 0.12, 0.2, 0.37, DT)
    oracle_turbulence(orc, 7, 0.05, 0.12, 0.2, 0.37, DT)
    for name in ("u", "v", "w"):
        compare(to_np(getattr(gpu, name).data), getattr(orc, name).data, what=name)
        assert bit_equal(to_np(getattr(gpu, name).data), getattr(orc, name).data), name


@pytest.mark.parametrize("scale_m", [0.011, 0.0102, 0.5, 0.009])
def test_turbulence_is_bit_exact_across_noise_scales(device, oracle32, scale_m):
    """On 1 cm cells: 0.011 m puts half a cell at 0.45 lattice units (the two points of each curl difference often
    lie in neighbouring lattice cells), 0.0102 m at 0.49 (the largest the shared-lattice form takes), 0.5 m at
    0.01; at 0.009 m a cell is more than one lattice unit and every point's noise is evaluated on its own."""
    from oracle.sim import apply_turbulence as oracle_turbulence
    from plume_amd.sim.turbulence import apply_turbulence

    gpu, orc = pair(device, oracle32, (31, 33, 29), DEFAULT_FACES)
    apply_turbulence(gpu, 11, scale_m, 0.12, 0.2, 1.63, DT)
    oracle_turbulence(orc, 11, scale_m, 0.12, 0.2, 1.63, DT)
    for name in ("u", "v", "w"):
        assert bit_equal(to_np(getattr(gpu, name).data), getattr(orc, name).data), name


# ---------------------------------------------------------------------- pressure and multigrid

@pytest.mark.parametrize("cells, faces", GRIDS)
def test_divergence_sweeps_and_gradient_are_bit_exact(device, oracle32, cells, faces):
    from plume_amd.sim.pressure import compute_divergence, red_black_sweep, red_black_sweeps, subtract_gradient, \
        zero_closed_faces

    gpu, orc = pair(device, oracle32, cells, faces)
    zero_closed_faces(gpu)
    for axis, f in enumerate(orc.velocity_fields()):
        oracle32.zero_closed_faces(orc.info, f.data, axis)
    compute_divergence(gpu)
    oracle32.divergence(orc.info, *(f.data for f in orc.velocity_fields()), orc.divergence)
    assert bit_equal(to_np(gpu.divergence), orc.divergence)
    # one colour at a time, then fused two-colour sweeps (ping-pong), from a random warm start
    q0 = rng(11).uniform(-1e-3, 1e-3, gpu.cells).astype(np.float32)
    gpu.pressure.copy_(torch.from_numpy(q0))
    orc.pressure[...] = q0
    red_black_sweep(gpu.pressure, gpu.divergence, gpu.info, 0, 1.5)
    red_black_sweep(gpu.pressure, gpu.divergence, gpu.info, 1, 1.5)
    oracle32.rb_sweeps(orc.info, orc.pressure, orc.divergence, 1, 1.5)
    assert bit_equal(to_np(gpu.pressure), orc.pressure)
    scratch = torch.empty_like(gpu.pressure)
    for sweeps in (1, 2, 3):
        red_black_sweeps(gpu.pressure, gpu.divergence, gpu.info, gpu.cells, device, sweeps, 1.0, scratch)
        oracle32.rb_sweeps(orc.info, orc.pressure, orc.divergence, sweeps, 1.0)
        assert bit_equal(to_np(gpu.pressure), orc.pressure), f"fused sweeps x{sweeps}"
    subtract_gradient(gpu)
    for axis, f in enumerate(orc.velocity_fields()):
        oracle32.subtract_gradient(orc.info, f.data, orc.pressure, axis)
    for name in ("u", "v", "w"):
        assert bit_equal(to_np(getattr(gpu, name).data), getattr(orc, name).data), name


@pytest.mark.parametrize("cells, faces", GRIDS)
def test_closing_the_faces_with_the_divergence_is_bit_exact(device, oracle32, cells, faces):
    """plm_close_divergence (project's zeroing of the closed faces and its divergence in one launch) leaves the
    faces and the divergence as the oracle's two steps do."""
    from plume_amd.sim.pressure import close_faces_and_divergence

    gpu, orc = pair(device, oracle32, cells, faces)
    close_faces_and_divergence(gpu)
    for axis, f in enumerate(orc.velocity_fields()):
        oracle32.zero_closed_faces(orc.info, f.data, axis)
    oracle32.divergence(orc.info, *(f.data for f in orc.velocity_fields()), orc.divergence)
    assert bit_equal(to_np(gpu.divergence), orc.divergence)
    for name in ("u", "v", "w"):
        assert bit_equal(to_np(getattr(gpu, name).data), getattr(orc, name).data), name


@pytest.mark.parametrize("cells", [(64, 96, 64), (65, 47, 33)], ids=["64x96x64", "odd-large"])
def test_fused_sweep_on_large_grids_is_bit_exact(device, oracle32, cells):
    from plume_amd.sim.pressure import red_black_sweeps

    for faces in (DEFAULT_FACES, ALL_CLOSED):
        gpu, orc = pair(device, oracle32, cells, faces)
        oracle32.divergence(orc.info, *(f.data for f in orc.velocity_fields()), orc.divergence)
        gpu.divergence.copy_(torch.from_numpy(orc.divergence))
        scratch = torch.empty_like(gpu.pressure)
        red_black_sweeps(gpu.pressure, gpu.divergence, gpu.info, gpu.cells, device, 4, 1.0, scratch)
        oracle32.rb_sweeps(orc.info, orc.pressure, orc.divergence, 4, 1.0)
        assert bit_equal(to_np(gpu.pressure), orc.pressure)


@pytest.mark.parametrize("level, cells, faces", [(0, (45, 38, 70), DEFAULT_FACES), (1, (45, 38, 70), DEFAULT_FACES),
                                                  (0, (40, 36, 128), ALL_CLOSED), (1, (90, 90, 127), DEFAULT_FACES),
                                                  (1, (90, 90, 127), ALL_CLOSED), (0, (70, 66, 96), ALL_OPEN),
                                                  (1, (64, 96, 64), DEFAULT_FACES), (1, (96, 64, 160), ALL_OPEN),
                                                  (2, (128, 64, 128), ALL_CLOSED), (1, (96, 64, 160), ALL_CLOSED),
                                                  (0, (37, 30, 128), DEFAULT_FACES), (1, (70, 60, 128), DEFAULT_FACES),
                                                  (1, (40, 34, 128), ALL_CLOSED), (0, (19, 13, 256), ALL_OPEN),
                                                  (0, (21, 18, 64), ALL_CLOSED)],
                         ids=["top", "coarse", "top-closed-k-full", "coarse-k-full-open", "coarse-k-full-closed",
                              "top-open-k-partial", "flat-k32-default", "flat-k80-open", "flat-corners-closed",
                              "flat-k80-closed", "k128-top", "k64-flat-ghost", "k64-flat-closed", "k256-open",
                              "k64-top-closed"])
def test_every_sweep_tile_shape_is_bit_exact(device, oracle32, level, cells, faces):
    """The autotuner may pick any tile shape; each gives the reference's sweep, bit for bit (a coarse level
    exercises the overhang couplings; 128 cells along k with closed faces give tiles whose 64 k cells all
    exist, the unit-coupling fast path, with Σc = 5 at both k ends; level 1 of 90 x 90 x 127 has 64 cells
    along k with fill 0.5, so the general-coupling path meets the last cell, its neighbour, and a ghost or
    closed face). Levels with fill (1, 1, 1) below the top take the flat path with ghost couplings other than
    1 on open faces (Σc 4 + 1 / 0.75 and the like: the IEEE division), fewer than 64 cells along k (32) or a
    last k tile of 16 cells (80), and corner cells whose three closed faces leave Σc = 3; levels of whole 64-cell
    k tiles (64, 128, 256) with tiles that overhang the level's end in i and j, under every face kind."""
    from oracle.sim import level_infos
    from plume_amd._lib import LIB, call, int3, ptr, stream
    from plume_amd.sim.multigrid import build_levels
    from plume_amd.sim.pressure import red_black_sweeps

    gpu, orc = pair(device, oracle32, cells, faces)
    lv = build_levels(gpu)[level]
    info = level_infos(oracle32, cells, CELL_M, orc.origin_m, orc.open_lo, orc.open_hi)[level][1]
    r = rng(21)
    rhs = r.uniform(-1, 1, lv.cells).astype(np.float32)
    q0 = r.uniform(-1e-3, 1e-3, lv.cells).astype(np.float32)
    expected = q0.copy()
    oracle32.rb_sweeps(info, expected, rhs.copy(), 2, 1.0)
    shapes = int(LIB.plm_sweep_tile_shapes(None, None, None, 0))
    rhs_t = torch.from_numpy(rhs).to(device)
    scratch = torch.empty_like(rhs_t)
    try:
        for shape in range(shapes):
            call("plm_set_sweep_tile", int3(lv.cells), shape)
            q = torch.from_numpy(q0).to(device)
            red_black_sweeps(q, rhs_t, lv.info, lv.cells, device, 2, 1.0, scratch)
            assert bit_equal(to_np(q), expected), shape
    finally:
        call("plm_set_sweep_tile", int3(lv.cells), -1)


def test_autotune_pins_a_shape_and_keeps_results(device, oracle32):
    from oracle.sim import Multigrid as OracleMultigrid, project as oracle_project
    from plume_amd.sim.multigrid import Multigrid
    from plume_amd.sim.pressure import project

    gpu, orc = pair(device, oracle32, (96, 128, 96), DEFAULT_FACES)
    solve = Multigrid(gpu, 2, 2)
    tuned = solve.autotune(repeats=2)
    assert (96, 128, 96) in tuned and 0 <= tuned[(96, 128, 96)][0] < 6
    project(gpu, solve)
    oracle_project(orc, OracleMultigrid(orc, 2, 2))
    assert bit_equal(to_np(gpu.pressure), orc.pressure)


@pytest.mark.parametrize("cells, faces", GRIDS)
def test_residual_restrict_prolong_are_bit_exact(device, oracle32, cells, faces):
    from plume_amd._lib import call, int3, ptr, stream
    from plume_amd.sim.multigrid import build_levels, residual

    gpu, orc = pair(device, oracle32, cells, faces)
    oracle32.divergence(orc.info, *(f.data for f in orc.velocity_fields()), orc.divergence)
    gpu.divergence.copy_(torch.from_numpy(orc.divergence))
    q0 = rng(12).uniform(-1e-3, 1e-3, gpu.cells).astype(np.float32)
    gpu.pressure.copy_(torch.from_numpy(q0))
    orc.pressure[...] = q0
    levels = build_levels(gpu)
    r = to_np(residual(levels[0]))
    expected = np.zeros_like(orc.pressure)
    oracle32.residual(orc.info, orc.pressure, orc.divergence, expected)
    assert bit_equal(r, expected)
    if len(levels) > 1:
        coarse = levels[1]
        call("plm_restrict", int3(cells), ptr(levels[0].residual), int3(coarse.cells), ptr(coarse.rhs), stream())
        expected_coarse = np.zeros(coarse.cells, dtype=np.float32)
        oracle32.restrict(expected, expected_coarse)
        assert bit_equal(to_np(coarse.rhs), expected_coarse)
        call("plm_prolong_add", int3(coarse.cells), ptr(coarse.rhs), int3(cells), ptr(gpu.pressure), stream())
        oracle32.prolong_add(expected_coarse, orc.pressure)
        assert bit_equal(to_np(gpu.pressure), orc.pressure)


def jet(cells, faces, device, oracle32):
    """A smooth upward jet (a Gaussian blob of v), the reference's hardest case for Gauss-Seidel
    (tests/sim/test_multigrid.py:32-36 of the reference)."""
    gpu = gpu_grid(cells, 0.1, faces, device=device)
    orc = oracle_grid(oracle32, cells, 0.1, faces)
    shape = orc.v.data.shape
    idx = np.indices(shape).astype(np.float64) + 0.5
    center = [n / 2 for n in shape]
    radius = min(cells) / 5
    blob = np.exp(-sum((idx[a] - center[a]) ** 2 for a in range(3)) / (2 * radius**2)).astype(np.float32)
    gpu.v.data.copy_(torch.from_numpy(blob))
    orc.v.data[...] = blob
    return gpu, orc


@pytest.mark.parametrize("cells", [(15, 17, 13), (32, 32, 32), (31, 33, 29), (6, 96, 6)])
@pytest.mark.parametrize("faces", [ALL_CLOSED, DEFAULT_FACES], ids=["closed", "open"])
def test_projection_with_multigrid_is_bit_exact(device, oracle32, cells, faces):
    from oracle.sim import Multigrid as OracleMultigrid, project as oracle_project
    from plume_amd.sim.multigrid import Multigrid
    from plume_amd.sim.pressure import project

    gpu, orc = jet(cells, faces, device, oracle32)
    project(gpu, Multigrid(gpu, 2, 2))
    oracle_project(orc, OracleMultigrid(orc, 2, 2))
    assert bit_equal(to_np(gpu.pressure), orc.pressure)
    for name in ("u", "v", "w"):
        assert bit_equal(to_np(getattr(gpu, name).data), getattr(orc, name).data), name


@pytest.mark.parametrize("sweeps", [1, 3])
def test_multigrid_with_odd_sweep_counts_is_bit_exact(device, oracle32, sweeps):
    """An odd number of sweeps per level ends each level's ping-pong with a copy back (the tile sweeps, the LDS
    tail and the large levels' row prolongation)."""
    from oracle.sim import Multigrid as OracleMultigrid, project as oracle_project
    from plume_amd.sim.multigrid import Multigrid
    from plume_amd.sim.pressure import project

    gpu, orc = pair(device, oracle32, (64, 96, 64), DEFAULT_FACES)
    project(gpu, Multigrid(gpu, 2, sweeps))
    oracle_project(orc, OracleMultigrid(orc, 2, sweeps))
    assert bit_equal(to_np(gpu.pressure), orc.pressure)


@pytest.mark.parametrize("sweeps", [2, 3])
@pytest.mark.parametrize("faces", [DEFAULT_FACES, ALL_CLOSED, ALL_OPEN], ids=["default", "closed", "open"])
@pytest.mark.parametrize("cells", [(45, 38, 70), (90, 90, 127), (70, 66, 45)], ids=["odd", "k-full", "wide"])
def test_projection_with_mid_levels_is_bit_exact(device, oracle32, cells, faces, sweeps):
    """Grids with levels between the simulation grid and the LDS tail (their sweeps are tile launches with the
    coupling table, an odd count ending in a copy back): the whole projection equals the oracle's bit for bit on
    odd sizes, overhanging coarse cells, k extents that do not fill a 64-wide tile, and every face kind."""
    from oracle.sim import Multigrid as OracleMultigrid, project as oracle_project
    from plume_amd.sim.multigrid import Multigrid, PLM_LDS_CELLS, build_levels
    from plume_amd.sim.pressure import project

    gpu, orc = jet(cells, faces, device, oracle32)
    levels = build_levels(gpu)
    tail = 0
    mid = []
    for index in reversed(range(len(levels))):
        tail += int(np.prod(levels[index].cells))
        if tail > PLM_LDS_CELLS and index > 0:
            mid.append(levels[index].cells)
    assert mid, "the grid must have a level between the simulation grid and the tail"
    project(gpu, Multigrid(gpu, 2, sweeps))
    oracle_project(orc, OracleMultigrid(orc, 2, sweeps))
    assert bit_equal(to_np(gpu.pressure), orc.pressure), mid
    for name in ("u", "v", "w"):
        assert bit_equal(to_np(getattr(gpu, name).data), getattr(orc, name).data), name


def test_timed_sweeps_are_bit_exact_and_report_each_launch(device, oracle32):
    """While a launch family is timed, its sweeps launch with the event pair bound to the dispatch
    (hipExtLaunchKernelGGL). The solve must stay bit-identical to the oracle, and the events must give one
    positive duration per launch."""
    from oracle.sim import Multigrid as OracleMultigrid, project as oracle_project
    from plume_amd._lib import LIB
    from plume_amd.sim.multigrid import Multigrid
    from plume_amd.sim.pressure import project

    gpu, orc = pair(device, oracle32, (64, 96, 64), DEFAULT_FACES)
    solve = Multigrid(gpu, 2, 2)
    total, launches = ctypes.c_double(), ctypes.c_int64()
    assert LIB.plm_timing_enable(b"rb_sweep_tile_top") == 0
    try:
        project(gpu, solve)
        torch.cuda.synchronize()
        assert LIB.plm_timing_read(ctypes.byref(total), ctypes.byref(launches)) == 0, LIB.plm_last_error()
    finally:
        LIB.plm_timing_disable()
    oracle_project(orc, OracleMultigrid(orc, 2, 2))
    assert bit_equal(to_np(gpu.pressure), orc.pressure)
    assert launches.value > 0
    assert 0.0 < total.value / launches.value < 50.0  # ms per launch of a 64x96x64 sweep


def test_multigrid_on_a_larger_grid_is_bit_exact(device, oracle32):
    from oracle.sim import Multigrid as OracleMultigrid, project as oracle_project
    from plume_amd.sim.multigrid import Multigrid
    from plume_amd.sim.pressure import project

    gpu, orc = pair(device, oracle32, (64, 96, 64), DEFAULT_FACES)
    project(gpu, Multigrid(gpu, 2, 2))
    oracle_project(orc, OracleMultigrid(orc, 2, 2))
    assert bit_equal(to_np(gpu.pressure), orc.pressure)


TAIL_SMOOTH_CELLS = 2048  # a flat level below the simulation grid with more cells stays out of the LDS tail


def lds_smoothed_levels(levels, lds_cells):
    """The kinds ("big", "small", "thin") of the levels the single-grid V-cycle smooths in LDS tiles (two sweeps a
    side): flat levels below the simulation grid and above the tail. The tail starts at the first level, counted
    from the coarsest, at which the levels so far would exceed lds_cells, or which is flat with more than
    TAIL_SMOOTH_CELLS cells and a level below it."""
    total, start = 0, len(levels)
    for index in reversed(range(len(levels))):
        cells = int(np.prod(levels[index].cells))
        flat = all(f == 1.0 for f in levels[index].info.fill_hi)
        if total + cells > lds_cells or (0 < index < len(levels) - 1 and flat and cells > TAIL_SMOOTH_CELLS):
            break
        total += cells
        start = index
    kinds = []
    for index in range(1, start):
        if all(f == 1.0 for f in levels[index].info.fill_hi):
            n = levels[index].cells
            big = ((n[0] + 3) // 4) * ((n[1] + 7) // 8) * ((n[2] + 31) // 32) >= 128
            kinds.append("big" if big else "thin" if n[2] <= 16 else "small")
    return kinds


@pytest.mark.parametrize("cells, faces, kinds", [((96, 128, 96), ALL_OPEN, ["big", "small", "thin"]),
                                                 ((96, 128, 96), ALL_CLOSED, ["big", "small", "thin"]),
                                                 ((100, 72, 88), DEFAULT_FACES, ["big", "small"]),
                                                 ((100, 72, 88), ALL_OPEN, ["big", "small"]),
                                                 ((64, 96, 64), DEFAULT_FACES, ["small", "thin"])],
                         ids=["96x128x96-open", "96x128x96-closed", "overhang-default", "overhang-open", "64x96x64"])
def test_fused_last_sweep_and_restriction_is_bit_exact(device, oracle32, cells, faces, kinds):
    """Flat levels between the simulation grid and the LDS tail smooth in LDS tiles. A big level (at least 128
    tiles of 4 x 8 x 32) runs one sweep a launch: from zero, then the last pre-smoothing sweep with the residual
    and restriction, then the prolongation with the first post-smoothing sweep (the second is a tile sweep); a
    small one (2 x 4 x 32 tiles, 2 x 4 x 16 on a thin level of at most 16 cells along k) runs both pre-smoothing
    sweeps with the restriction as one launch, and the prolongation with both post-smoothing sweeps as another;
    flat levels of more than 2048 cells stay out of the LDS tail. 96 x 128 x 96 has one level of each kind
    (48 x 64 x 48, 24 x 32 x 24, 12 x 16 x 12); 100 x 72 x 88 has a big level whose tiles overhang its end on every
    axis (50 x 36 x 44) and a small level of odd extent (25 x 18 x 22) whose last coarse cells have fewer than
    eight children and whose prolongation clamps at the coarse level's end; 64 x 96 x 64 has 32 x 48 x 32 and the
    thin 16 x 24 x 16, whose tail is then 8 x 12 x 8 and below."""
    from oracle.sim import Multigrid as OracleMultigrid, project as oracle_project
    from plume_amd.sim.multigrid import Multigrid, PLM_LDS_CELLS, build_levels
    from plume_amd.sim.pressure import project

    gpu, orc = pair(device, oracle32, cells, faces)
    assert lds_smoothed_levels(build_levels(gpu), PLM_LDS_CELLS) == kinds
    project(gpu, Multigrid(gpu, 2, 2))
    oracle_project(orc, OracleMultigrid(orc, 2, 2))
    assert bit_equal(to_np(gpu.pressure), orc.pressure)
    for name in ("u", "v", "w"):
        assert bit_equal(to_np(getattr(gpu, name).data), getattr(orc, name).data), name


def test_tolerance_mode_reaches_the_target_and_matches_fixed_cycles(device, oracle32):
    from oracle.sim import Multigrid as OracleMultigrid, compute_divergence, poisson_residual
    from plume_amd.sim.multigrid import Multigrid, relative_residual
    from plume_amd.sim.pressure import compute_divergence as gpu_divergence, zero_closed_faces

    gpu, orc = pair(device, oracle32, (48, 64, 40), DEFAULT_FACES)
    zero_closed_faces(gpu)
    gpu_divergence(gpu)
    for axis, f in enumerate(orc.velocity_fields()):
        oracle32.zero_closed_faces(orc.info, f.data, axis)
    compute_divergence(orc)
    solve = Multigrid(gpu, 2, 2, tolerance=1e-5, max_cycles=20)
    solve(gpu)
    cycles = solve.last_cycles()
    ratio = solve.last_relative_residual()
    assert ratio <= 1e-5, (cycles, ratio)
    assert relative_residual(gpu.info, gpu.pressure, gpu.divergence) == pytest.approx(ratio, rel=1e-9)
    # the same number of fixed cycles on the oracle gives the same pressure, bit for bit
    OracleMultigrid(orc, cycles, 2)(orc)
    assert bit_equal(to_np(gpu.pressure), orc.pressure)
    assert poisson_residual(orc) == pytest.approx(ratio, rel=1e-6)


# ---------------------------------------------------------------------- budget and CFL

@pytest.mark.parametrize("cells, faces", GRIDS[:2])
def test_smoke_budget_matches_oracle(device, oracle32, cells, faces):
    from oracle.sim import SmokeBudget as OracleBudget, advect_maccormack as oracle_advect
    from plume_amd.sim.advect import advect_maccormack
    from plume_amd.sim.conserve import SmokeBudget
    from plume_amd.sim.reduce import TripleSum

    gpu, orc = pair(device, oracle32, cells, faces)
    advect_maccormack(gpu.smoke, tuple(f.data for f in gpu.velocity_fields()), gpu, 4 * DT)
    oracle_advect(orc.smoke, tuple(f.data for f in orc.velocity_fields()), orc, 4 * DT)
    sums = TripleSum(cells, device)(gpu.smoke.data, gpu.smoke.hat, gpu.smoke.next)
    budget, obudget = SmokeBudget(gpu), OracleBudget(orc)
    budget.hold(gpu)
    obudget.hold(orc)
    assert budget.last_totals == pytest.approx(obudget.last_totals, rel=1e-12)
    assert sums == pytest.approx(obudget.last_totals, rel=1e-12)
    assert budget.last_factor == pytest.approx(obudget.last_factor, rel=1e-12)
    compare(to_np(gpu.smoke.next), orc.smoke.next, what="held smoke")


def test_grid_speed_is_the_exact_maximum(device, oracle32):
    from plume_amd.sim.step_cap import StepCap
    from support import tiny_recipe

    gpu, orc = pair(device, oracle32, (31, 33, 29), DEFAULT_FACES)
    gpu.w.data[3, 4, 5] = -7.25
    expected = max(float(np.abs(f.data).max()) for f in orc.velocity_fields())
    assert StepCap(tiny_recipe()).grid_speed(gpu) == max(expected, 7.25)


def test_a_nan_velocity_stops_the_step_cap(device, oracle32):
    # NumPy's max propagates a NaN into math.ceil, which raises (sim/step_cap.py:38-55): the device maximum
    # must not drop it
    import math

    from plume_amd.sim.step_cap import StepCap
    from support import tiny_recipe

    gpu, _ = pair(device, oracle32, (31, 33, 29), DEFAULT_FACES)
    gpu.v.data[30, 1, 2] = float("nan")
    cap = StepCap(tiny_recipe())
    assert math.isnan(cap.grid_speed(gpu))
    with pytest.raises(ValueError, match="blown up"):
        cap.plan(gpu, 0.0, 0.1, 3, "frame 0")
    gpu.v.data[30, 1, 2] = float("-inf")
    assert cap.grid_speed(gpu) == math.inf


# ---------------------------------------------------------------------- rendering

@pytest.mark.parametrize("cells", [(16, 16, 16), (31, 33, 29)])
def test_light_volume_matches_oracle(device, oracle32, cells):
    from plume_amd.render.light_volume import LightVolume

    gpu, orc = pair(device, oracle32, cells, DEFAULT_FACES)
    direction = np.array([-0.5, 0.7, 0.5])
    volume = LightVolume(cells, direction, 4.0, device)
    light = to_np(volume.compute(gpu.smoke.data, 0.05))
    expected = np.zeros(cells, dtype=np.float32)
    oracle32.light_volume(orc.smoke.data, volume.toward_light, 0.05, 4.0, expected)
    compare(light, expected, rel=2e-6, what="light")


@pytest.mark.parametrize("layout", ["shade", "normal"])
@pytest.mark.parametrize("elevation, azimuth", [(0.0, 0.0), (25.0, 40.0), (90.0, 0.0)])
def test_raymarch_matches_oracle(device, oracle32, layout, elevation, azimuth):
    from oracle.render import render_oracle
    from plume_amd.recipe.schema import NORMAL_LAYOUT, SHADE_LAYOUT, RenderSettings
    from plume_amd.render.camera import make_camera, make_domain_view
    from plume_amd.render.frame_renderer import FrameRenderer

    cells = (24, 20, 16)
    gpu, orc = pair(device, oracle32, cells, DEFAULT_FACES, ambient=293.0)
    temperature = (293.0 + 900.0 * random_state(cells, 3)["temperature"]).astype(np.float32)
    gpu.temperature.data.copy_(torch.from_numpy(temperature))
    orc.temperature.data[...] = temperature
    channels = SHADE_LAYOUT if layout == "shade" else NORMAL_LAYOUT
    settings = RenderSettings(opacity_per_density=20.0, glow_start_k=600.0, smoke_hides_heat=0.7)
    view = make_domain_view(make_camera(elevation, azimuth), orc.origin_m, tuple(n * CELL_M for n in cells), CELL_M,
                            2.0)
    image = FrameRenderer(view, settings, cells, channels, device).render(gpu)
    expected = render_oracle(oracle32, orc, view, settings, channels)
    for channel in range(5):
        compare(image[..., channel], expected[..., channel], rel=2e-5, atol=1e-6, what=f"{layout} ch{channel}")
