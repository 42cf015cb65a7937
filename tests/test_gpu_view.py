"""The viewer's kernels against the numpy oracle (oracle/view.py), bit for bit: binary16 narrowing and widening
(random fields plus ties, subnormals, overflow and signed zeros), the fused cell-centre velocity, and the
six boundary layers; the frame cache's store / load round trip and its thinned plan on the GPU
(tests/view/test_frame_cache.py of the reference)."""

import math

import numpy as np
import pytest
import torch

from oracle.view import boundary_layers as oracle_layers, center_velocity_half, narrow

pytestmark = pytest.mark.gpu

SPECIAL = np.array([0.0, -0.0, 1.0, 1.0 + 2.0**-11, 1.0 + 3 * 2.0**-11, 2.0**-14, 2.0**-24, 2.0**-25, 3 * 2.0**-25,
                    -(2.0**-20), 65504.0, 65519.99, 65520.0, -70000.0, 1e-8, 0.1, 293.15, 1499.7],
                   dtype=np.float32)


def device_array(values, device):
    return torch.from_numpy(np.ascontiguousarray(values, dtype=np.float32)).to(device)


@pytest.mark.parametrize("cells", [(8, 6, 4), (5, 3, 3), (37, 29, 41)])
def test_a_stored_frame_is_the_field_rounded_to_16_bits(device, cells):
    from plume_amd.view.frame_cache import FrameCache

    rng = np.random.default_rng(5)
    smoke = rng.uniform(0.0, 3.0, cells).astype(np.float32)
    smoke.reshape(-1)[:min(SPECIAL.size, smoke.size)] = SPECIAL[:min(SPECIAL.size, smoke.size)]
    temperature = rng.uniform(290.0, 1500.0, cells).astype(np.float32)
    cache = FrameCache(cells, frames=4, budget_bytes=1e9, device=device)
    cache.store(0, device_array(smoke, device), device_array(temperature, device))
    torch.cuda.synchronize()
    stored = cache.stored_smoke(0).cpu().numpy()
    assert stored.dtype == np.float16
    assert np.array_equal(stored.view(np.uint16), narrow(smoke).view(np.uint16))
    wide_smoke, wide_temperature = cache.load(0, cache.scratch())
    torch.cuda.synchronize()
    assert np.array_equal(wide_smoke.cpu().numpy().view(np.uint32), narrow(smoke).astype(np.float32).view(np.uint32))
    assert np.array_equal(wide_temperature.cpu().numpy(), narrow(temperature).astype(np.float32))


@pytest.mark.parametrize("cells", [(8, 6, 4), (37, 29, 41)])
def test_a_kept_velocity_is_the_centered_velocity_rounded_to_16_bits(device, cells):
    from plume_amd.view.frame_cache import FrameCache, frame_bytes

    rng = np.random.default_rng(7)
    nx, ny, nz = cells
    u = rng.uniform(-2.0, 2.0, (nx + 1, ny, nz)).astype(np.float32)
    v = rng.uniform(-2.0, 2.0, (nx, ny + 1, nz)).astype(np.float32)
    w = rng.uniform(-2.0, 2.0, (nx, ny, nz + 1)).astype(np.float32)
    zeros = torch.zeros(cells, dtype=torch.float32, device=device)
    cache = FrameCache(cells, frames=2, budget_bytes=1e9, device=device)
    cache.store(0, zeros, zeros, (device_array(u, device), device_array(v, device), device_array(w, device)))
    want = center_velocity_half(u, v, w)
    assert cache.velocity_kept(0)
    assert np.array_equal(cache.stored_velocity(0).cpu().numpy().view(np.uint16), want.view(np.uint16))
    wide = cache.load_velocity(0, cache.velocity_scratch())
    torch.cuda.synchronize()
    assert np.array_equal(wide.cpu().numpy(), want.astype(np.float32))
    assert cache.stored_bytes == frame_bytes(cells, velocity=True)


def test_a_thinned_cache_skips_frames_and_keeps_no_velocity(device):
    from plume_amd.view.frame_cache import FrameCache, frame_bytes

    cells = (8, 6, 4)
    cache = FrameCache(cells, frames=8, budget_bytes=4 * frame_bytes(cells), device=device)
    zeros = torch.zeros(cells, dtype=torch.float32, device=device)
    for frame in range(8):
        cache.store(frame, zeros, zeros, (torch.zeros((9, 6, 4), device=device), torch.zeros((8, 7, 4), device=device),
                                          torch.zeros((8, 6, 5), device=device)))
    assert cache.kept_frames() == [0, 2, 4, 6] and cache.nearest_kept(5) == 4 and cache.nearest_kept(0) == 0
    assert not cache.has_velocity and not cache.velocity_kept(0)
    assert cache.stored_bytes == 4 * frame_bytes(cells)


@pytest.mark.parametrize("shape, extent", [((3, 4, 5), (5, 5)), ((38, 48, 38), (48, 48)), ((7, 2, 9), (4, 6))])
def test_the_boundary_layers_are_the_smokes_outer_cells(device, shape, extent):
    from plume_amd.view.live_run import boundary_layers

    smoke = np.random.default_rng(3).random(shape).astype(np.float32)
    layers = torch.full((6, *extent), -1.0, dtype=torch.float32, device=device)
    boundary_layers(device_array(smoke, device), layers)
    torch.cuda.synchronize()
    assert np.array_equal(layers.cpu().numpy(), oracle_layers(smoke, *extent))


def test_misaligned_pointers_are_refused_on_the_host(device):
    from plume_amd._lib import PlumeHipError, call, ptr, stream

    src = torch.zeros(9, dtype=torch.float32, device=device)
    dst = torch.zeros(8, dtype=torch.float16, device=device)
    with pytest.raises(PlumeHipError, match="aligned"):
        call("plm_narrow_f16", 8, ptr(src[1:]), ptr(dst, torch.float16), stream())


def test_the_puff_run_notes_the_ceiling_once_the_smoke_reaches_it(device):
    """EdgeWatch on a run's frames: no face while the puff rises, and only faces the recipe leaves open (or
    the ceiling) ever."""
    from plume_amd.recipe.derived import derive
    from plume_amd.sim.run import simulate_frames
    from plume_amd.view.live_run import FACE_NAMES, EdgeWatch
    from support import tiny_recipe

    recipe = tiny_recipe()
    derived = derive(recipe)
    watch = EdgeWatch(recipe, derived.cells, device)
    notes = [watch.faces_reached(grid.smoke.data) for _, grid in simulate_frames(recipe, derived, device)]
    assert notes[0] == ()
    faces = recipe.domain.faces
    for note in notes:
        assert all(name in FACE_NAMES and (name == "top" or getattr(faces, name) == "open") for name in note)
    assert math.isfinite(watch.layers.abs().max().item())
