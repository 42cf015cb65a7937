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


# ---------------------------------------------------------------------- the viewer's perspective march

PERSPECTIVE_CELLS = (16, 16, 16)
PERSPECTIVE_CELL_M = 0.1


def perspective_scene(device, oracle32, smoke, temperature=None):
    from support import ALL_CLOSED, gpu_grid, oracle_grid

    gpu = gpu_grid(PERSPECTIVE_CELLS, PERSPECTIVE_CELL_M, ALL_CLOSED, device=device)
    orc = oracle_grid(oracle32, PERSPECTIVE_CELLS, PERSPECTIVE_CELL_M, ALL_CLOSED)
    fields = {"smoke": smoke} if temperature is None else {"smoke": smoke, "temperature": temperature}
    for name, values in fields.items():
        values = np.ascontiguousarray(values, dtype=np.float32)
        getattr(gpu, name).data.copy_(torch.from_numpy(values))
        getattr(orc, name).data[...] = values
    return gpu, orc


def blob16(center, radius):
    idx = np.indices(PERSPECTIVE_CELLS).astype(np.float64) + 0.5
    return np.exp(-sum((idx[a] - center[a]) ** 2 for a in range(3)) / (2 * radius**2))


def perspective_params(view, distance_m):
    """A perspective camera `distance_m` out along the view's axis whose plane at the world origin matches the
    domain view's pixels (plane coordinates scale by 1 / distance), as the viewer sets it up
    (view/viewport.py:152,225)."""
    from plume_amd._lib import VIEW_PERSPECTIVE, float3
    from plume_amd.render.frame_renderer import domain_view_params

    camera = view.camera
    center_x = view.left_m + view.width_px * view.pixel_m / 2
    center_y = view.top_m - view.height_px * view.pixel_m / 2
    params = domain_view_params(view)
    params.perspective = VIEW_PERSPECTIVE
    params.eye = float3(camera.right * center_x + camera.up * center_y + camera.toward * distance_m)
    params.left_m = (view.left_m - center_x) / distance_m
    params.top_m = (view.top_m - center_y) / distance_m
    params.pixel_m = view.pixel_m / distance_m
    return params


def gpu_march(gpu, renderer, params, height, width, device):
    image = torch.zeros((height, width, 5), dtype=torch.float32, device=device)
    renderer.prepare(gpu.smoke.data, gpu.info.cell_m)
    renderer.march(gpu.smoke.data, gpu.temperature.data, gpu.info, params, image)
    torch.cuda.synchronize()
    return image.cpu().numpy()


@pytest.mark.parametrize("layout", ["shade", "normal"])
@pytest.mark.parametrize("distance_m, rel", [(3.0, 2e-5), (200.0, 3e-4)])
def test_the_perspective_march_matches_the_oracle(device, oracle32, layout, distance_m, rel):
    """plm_raymarch with perspective = 1 (render/raymarch.py:72-83: rays from the eye through the 1 m plane)
    against the oracle's march with the same view: within 2e-5 of each channel's scale with the eye 3 m out;
    with the eye 200 m out, sample positions carry FP32 rounding of 1.5e-5 m and this dense, hot blob turns
    the last-bit differences of exp and pow into more: the FP32 oracle itself differs from the FP64 oracle by
    8.4e-5 of the heat scale there, so the bar is 3e-4 (the reference's CPU-against-GPU image bar is 1-2%,
    tests/cli/test_cpu_gpu.py:21-33)."""
    from oracle.render import render_oracle
    from plume_amd.recipe.schema import NORMAL_LAYOUT, SHADE_LAYOUT, RenderSettings
    from plume_amd.render.camera import make_camera, make_domain_view
    from plume_amd.render.frame_renderer import FrameRenderer

    gpu, orc = perspective_scene(device, oracle32, 3.0 * blob16((8.0, 7.0, 9.0), 3.0),
                                 293.0 + 900.0 * blob16((8.0, 7.0, 9.0), 2.0))
    channels = SHADE_LAYOUT if layout == "shade" else NORMAL_LAYOUT
    view = make_domain_view(make_camera(25.0, 40.0), (-0.8, 0.0, -0.8), (1.6, 1.6, 1.6), PERSPECTIVE_CELL_M, 2.0)
    settings = RenderSettings()
    renderer = FrameRenderer(view, settings, PERSPECTIVE_CELLS, channels, device)
    params = perspective_params(view, distance_m)
    image = gpu_march(gpu, renderer, params, view.height_px, view.width_px, device)
    expected = render_oracle(oracle32, orc, view, settings, channels, params=params)
    for channel in range(5):
        scale = max(float(np.abs(expected[..., channel]).max()), 1e-30)
        worst = float(np.abs(image[..., channel] - expected[..., channel]).max())
        assert worst <= rel * scale + 1e-6, (channel, worst, scale)
    assert expected[..., 1].max() > 0.5


def test_a_far_narrow_perspective_camera_matches_the_orthographic_march(device, oracle32):
    """tests/render/test_view_cameras.py:42-56 on the HIP path: at 200 m the perspective image is within 0.01
    of the orthographic one (thickness, shade) and within 1% of the peak heat."""
    from plume_amd.recipe.schema import SHADE_LAYOUT, RenderSettings
    from plume_amd.render.camera import make_camera, make_domain_view
    from plume_amd.render.frame_renderer import FrameRenderer

    gpu, _ = perspective_scene(device, oracle32, 3.0 * blob16((8.0, 7.0, 9.0), 3.0),
                               293.0 + 900.0 * blob16((8.0, 7.0, 9.0), 2.0))
    view = make_domain_view(make_camera(25.0, 40.0), (-0.8, 0.0, -0.8), (1.6, 1.6, 1.6), PERSPECTIVE_CELL_M, 2.0)
    renderer = FrameRenderer(view, RenderSettings(), PERSPECTIVE_CELLS, SHADE_LAYOUT, device)
    ortho = gpu_march(gpu, renderer, renderer.view_params, view.height_px, view.width_px, device)
    persp = gpu_march(gpu, renderer, perspective_params(view, 200.0), view.height_px, view.width_px, device)
    assert np.abs(persp[..., 1] - ortho[..., 1]).max() < 0.01
    assert np.abs(persp[..., 2] - ortho[..., 2]).max() < 0.01
    assert np.abs(persp[..., 0] - ortho[..., 0]).max() < 0.01 * ortho[..., 0].max()
    assert ortho[..., 1].max() > 0.9


def test_an_eye_inside_the_smoke_sees_only_the_smoke_in_front_of_it(device, oracle32):
    """tests/render/test_view_cameras.py:59-69 on the HIP path: the eye 1.1 m from the back wall inside uniform
    smoke of density 0.5 sees thickness 1 - exp(-2 * 0.5 * 1.1)."""
    from plume_amd._lib import VIEW_PERSPECTIVE, float3
    from plume_amd.recipe.schema import SHADE_LAYOUT, RenderSettings
    from plume_amd.render.camera import make_camera, make_domain_view
    from plume_amd.render.frame_renderer import FrameRenderer, domain_view_params

    gpu, _ = perspective_scene(device, oracle32, np.full(PERSPECTIVE_CELLS, 0.5))
    view = make_domain_view(make_camera(0.0, 0.0), (-0.8, 0.0, -0.8), (1.6, 1.6, 1.6), PERSPECTIVE_CELL_M, 2.0)
    renderer = FrameRenderer(view, RenderSettings(opacity_per_density=2.0), PERSPECTIVE_CELLS, SHADE_LAYOUT, device)
    params = domain_view_params(view)
    params.perspective = VIEW_PERSPECTIVE
    params.eye = float3((0.0, 0.8, 0.3))  # 1.1 m from the back wall (z = -0.8), looking along -z
    params.left_m, params.top_m, params.pixel_m = -0.5, 0.5, 1.0 / 9.0
    image = gpu_march(gpu, renderer, params, 9, 9, device)
    assert image[4, 4, 1] == pytest.approx(1.0 - math.exp(-2.0 * 0.5 * 1.1), rel=1e-4)
