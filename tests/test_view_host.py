"""The viewer's frame cache plan and edge notes (host logic, CPU), and the viewer oracle's pins.

Follows tests/view/test_frame_cache.py and tests/view/test_live_run.py:78-107 of the reference.
"""

import numpy as np
import pytest

from oracle.view import boundary_layers, center_velocity_half, narrow, widen
from plume_amd.view.frame_cache import FrameCache, frame_bytes, frame_stride, keeps_velocity
from plume_amd.view.live_run import FACE_NAMES, faces_reached, visible_density
from support import tiny_recipe

CELLS = (8, 6, 4)


# ---------------------------------------------------------------------- oracle pins

@pytest.mark.parametrize("value, bits", [
    (1.0, 0x3C00), (-2.0, 0xC000), (65504.0, 0x7BFF), (65519.99, 0x7BFF), (65520.0, 0x7C00),  # tie -> 2^16: inf
    (1.0 + 2.0**-11, 0x3C00), (1.0 + 3 * 2.0**-11, 0x3C02),  # ties to even
    (2.0**-14, 0x0400), (2.0**-24, 0x0001), (2.0**-25, 0x0000), (3 * 2.0**-25, 0x0002),  # subnormals, ties
    (-0.0, 0x8000), (0.1, 0x2E66),  # 0.1 = 1.6 x 2^-4: mantissa round(0.6 x 1024) = 614 = 0x266
])
def test_narrowing_follows_ieee_binary16_round_to_nearest_even(value, bits):
    assert int(narrow(np.array([value], dtype=np.float32)).view(np.uint16)[0]) == bits
    assert widen(narrow(np.array([value], dtype=np.float32)))[0] == np.float32(
        np.uint16(bits).view(np.float16).astype(np.float32))


def test_the_centre_of_affine_face_fields_is_exact():
    nx, ny, nz = CELLS
    i = np.arange(nx + 1, dtype=np.float32)[:, None, None]
    u = np.broadcast_to(0.25 * i - 1.0, (nx + 1, ny, nz)).astype(np.float32)
    v = np.broadcast_to(0.5 * np.arange(ny + 1, dtype=np.float32)[None, :, None], (nx, ny + 1, nz)).astype(np.float32)
    w = np.broadcast_to(-0.125 * np.arange(nz + 1, dtype=np.float32), (nx, ny, nz + 1)).astype(np.float32)
    centre = center_velocity_half(u, v, w).astype(np.float64)
    ci, cj, ck = np.meshgrid(np.arange(nx) + 0.5, np.arange(ny) + 0.5, np.arange(nz) + 0.5, indexing="ij")
    assert np.array_equal(centre[..., 0], 0.25 * ci - 1.0)
    assert np.array_equal(centre[..., 1], 0.5 * cj)
    assert np.array_equal(centre[..., 2], -0.125 * ck)


def test_boundary_layers_are_the_smokes_outer_cells():
    smoke = np.arange(3 * 4 * 5, dtype=np.float32).reshape(3, 4, 5)
    got = boundary_layers(smoke, 5, 5)
    assert np.array_equal(got[0, :4, :5], smoke[0]) and np.array_equal(got[1, :4, :5], smoke[-1])
    assert np.array_equal(got[2, :3, :5], smoke[:, 0]) and np.array_equal(got[3, :3, :5], smoke[:, -1])
    assert np.array_equal(got[4, :3, :4], smoke[:, :, 0]) and np.array_equal(got[5, :3, :4], smoke[:, :, -1])
    assert got[0, 4].sum() == 0.0 and got[2, 3].sum() == 0.0
    assert got[5, 2, 3] == smoke[2, 3, 4] and got[1, 3, 4] == smoke[2, 3, 4]


# ---------------------------------------------------------------------- the cache's plan

def test_the_stride_doubles_until_the_kept_frames_fit():
    one = frame_bytes(CELLS)
    assert one == 2 * 2 * 8 * 6 * 4
    assert frame_stride(64, CELLS, 64 * one) == 1
    assert frame_stride(64, CELLS, 64 * one - 1) == 2
    assert frame_stride(64, CELLS, 16 * one) == 4
    assert frame_stride(64, CELLS, 1) == 64  # one frame is always kept


def test_velocity_is_dropped_before_any_frame():
    with_velocity = frame_bytes(CELLS, velocity=True)
    assert with_velocity == 5 * 2 * 8 * 6 * 4
    assert keeps_velocity(8, CELLS, 8 * with_velocity)
    assert not keeps_velocity(8, CELLS, 8 * with_velocity - 1)
    cache = FrameCache(CELLS, frames=8, budget_bytes=8 * with_velocity - 1, device="cpu")
    assert not cache.has_velocity and cache.stride == 1 and not cache.thinned
    assert "no velocity" in cache.thinning_note() and "every" not in cache.thinning_note()
    assert FrameCache(CELLS, frames=8, budget_bytes=8 * with_velocity, device="cpu").has_velocity
    thinned = FrameCache(CELLS, frames=8, budget_bytes=4 * frame_bytes(CELLS), device="cpu",
                         shared_with_compare=True)
    assert thinned.stride == 2 and thinned.keeps(4) and not thinned.keeps(5)
    assert "every 2nd frame and no velocity" in thinned.thinning_note()
    assert "the Compare run" in thinned.thinning_note()
    assert FrameCache(CELLS, frames=8, budget_bytes=frame_bytes(CELLS), device="cpu").thinning_note().startswith(
        "frame cache keeps every 8th frame")
    assert FrameCache(CELLS, frames=8, budget_bytes=1e9, device="cpu").thinning_note() is None
    assert FrameCache(CELLS, frames=8, budget_bytes=1e9, device="cpu").nearest_kept(3) is None


# ---------------------------------------------------------------------- edge notes

def test_every_frame_notes_the_open_faces_and_ceiling_its_smoke_reached():
    recipe = tiny_recipe(domain={"faces": {"left": "closed"}})
    layers = np.zeros((len(FACE_NAMES), 4, 4), dtype=np.float32)
    assert faces_reached(layers, recipe) == ()
    layers[0, 1, 1] = 1.0  # left: closed, so no mark
    layers[2, 0, 0] = 1.0  # bottom: closed
    layers[3, 2, 3] = 1.0  # top: the ceiling
    layers[5, 0, 1] = 1.0  # front: open
    assert faces_reached(layers, recipe) == ("top", "front")
    threshold = visible_density(recipe)
    assert threshold == pytest.approx(-np.log(1 - 1 / 255) / (recipe.render.opacity_per_density * recipe.domain.cell_m))
    layers[:] = 0.0
    layers[5, 0, 0] = np.nextafter(np.float32(threshold), np.float32(0))
    assert faces_reached(layers, recipe) == ()
