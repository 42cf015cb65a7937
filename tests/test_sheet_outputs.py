"""The sheet PNG and its JSON (sheet/pack.py:12-30, sheet/sheet_json.py:11-41 of the reference): the tiling
against a frame-by-frame placement, the JSON's key order and rounding, and equal bytes for equal pixels."""

import json

import numpy as np
import pytest

from plume_amd.sheet.outputs import TOOL, SheetFacts, recipe_source, sheet_image, sheet_json, write_png


def frames(count, px=6, seed=1):
    rng = np.random.default_rng(seed)
    return [rng.integers(0, 256, (px, px, 4), dtype=np.uint8) for _ in range(count)]


@pytest.mark.parametrize("count, columns", [(5, 3), (4, 2), (1, 1), (7, 7), (3, 5)])
def test_frames_tile_left_to_right_then_top_to_bottom(count, columns):
    tiles = frames(count)
    sheet = sheet_image(tiles, columns)
    px = tiles[0].shape[0]
    rows = -(-count // columns)
    assert sheet.shape == (rows * px, columns * px, 4) and sheet.dtype == np.uint8
    expected = np.zeros_like(sheet)
    for index, tile in enumerate(tiles):
        row, column = index // columns, index % columns
        expected[row * px:(row + 1) * px, column * px:(column + 1) * px] = tile
    np.testing.assert_array_equal(sheet, expected)
    with pytest.raises(ValueError):
        sheet_image([], 2)


def test_the_json_keeps_the_contract_order_and_six_decimals():
    facts = SheetFacts(frames=16, columns=4, frame_px=128, fps=24.0, frame_m=1.4625000004, seat=(0.5, 0.83333333),
                       channels="heat,thickness,shade,unused", source="recipes/smoke_puff.toml",
                       written_by="plume run recipes/smoke_puff.toml --preview")
    text = sheet_json(facts)
    assert text.endswith("}\n") and "\n  \"frames\": 16," in text
    table = json.loads(text)
    assert list(table) == ["frames", "columns", "frame_px", "fps", "frame_m", "seat", "channels", "source", "tool",
                           "written_by"]
    assert table["frame_m"] == 1.4625 and table["seat"] == [0.5, 0.833333] and table["tool"] == TOOL


def test_equal_pixels_give_equal_png_bytes(tmp_path):
    from PIL import Image

    image = sheet_image(frames(5), 3)
    write_png(tmp_path / "a.png", image)
    write_png(tmp_path / "b.png", image.copy())
    assert (tmp_path / "a.png").read_bytes() == (tmp_path / "b.png").read_bytes()
    np.testing.assert_array_equal(np.asarray(Image.open(tmp_path / "a.png")), image)


def test_the_source_is_relative_to_the_git_root(tmp_path):
    (tmp_path / "repo" / ".git").mkdir(parents=True)
    (tmp_path / "repo" / "recipes").mkdir()
    recipe = tmp_path / "repo" / "recipes" / "puff.toml"
    recipe.write_text("")
    assert recipe_source(recipe) == "recipes/puff.toml"
    outside = tmp_path / "loose.toml"
    outside.write_text("")
    assert recipe_source(outside) == outside.as_posix()
