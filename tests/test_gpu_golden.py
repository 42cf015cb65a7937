"""Recipe runs on the HIP path through the sheet frames and the numbers file.

- `plume run recipes/smoke_puff.toml --preview` (38 x 48 x 38 cells, 16 frames of 128 px) against the
  reference's golden numbers file, with the bounds of the reference's own golden-numbers test
  (tests/cli/test_golden_numbers.py:21-63), and its sheet frames against the FP32 oracle's;
- the tiny recipe in both channel layouts against the oracle's run with the bounds of the reference's
  CPU-against-GPU test (tests/cli/test_cpu_gpu.py:21-33: thickness bytes within 3, shade / normal bytes
  within 6, coverage and thickness within 2%), the layout checks of tests/cli/test_run.py:43-61, identical
  frames from identical runs, and the run's numbers.txt and settings log (which loads back to the recipe);
- the tiny recipe with a 900 K emitter, so that heat reaches the sheet, against the oracle's run.
"""

import numpy as np
import pytest

from golden_numbers import (GOLDEN, RECIPE, SOURCE, assert_matches_golden, assert_sheets_agree, largest_differences,
                            oracle_run)
from support import TINY

pytestmark = pytest.mark.gpu


def test_preview_run_matches_the_golden_numbers_and_the_oracle_sheet(device):
    from plume_amd.pipeline import run_recipe
    from plume_amd.report.numbers import read_numbers

    result = run_recipe(RECIPE, preview=True, device="cuda", source=SOURCE)
    golden = GOLDEN.read_text()
    print(result.numbers, flush=True)
    print("largest differences:", largest_differences(result.numbers, golden), flush=True)
    assert result.derived.cells == (38, 48, 38) and len(result.raw_frames) == 16
    assert_matches_golden(result.numbers, golden)
    summary, golden_summary = read_numbers(result.numbers)[0], read_numbers(golden)[0]
    assert float(summary["shade_divisor"]) == pytest.approx(float(golden_summary["shade_divisor"]), rel=1e-3)
    cpu, cpu_text = oracle_run(RECIPE, True, SOURCE)
    worst = assert_sheets_agree(result.assembled.frames, cpu.frames, result.numbers, cpu_text)
    print("largest sheet byte differences against the oracle:", worst, flush=True)
    assert result.assembled.crop == cpu.crop


@pytest.mark.parametrize("channels", ["heat,thickness,shade,unused", "heat,thickness,normal_x,normal_y"])
def test_tiny_recipe_sheets_agree_with_the_oracle(device, tmp_path, channels):
    from plume_amd.pipeline import run_recipe

    path = tmp_path / "tiny.toml"
    path.write_text(TINY + f'channels = "{channels}"\n')
    first = run_recipe(path, device="cuda", source="tiny.toml", out=tmp_path / "out")
    second = run_recipe(path, device="cuda", source="tiny.toml")
    # the run's files: numbers.txt as returned, and a settings log that loads back to the same recipe
    import tomli

    from plume_amd.recipe.load import parse_recipe

    assert (tmp_path / "out" / "numbers.txt").read_text() == first.numbers
    settings = tomli.loads((tmp_path / "out" / "settings.toml").read_text())
    assert parse_recipe(settings) == first.recipe
    resolved = settings["resolved"]
    assert resolved["total_steps"] == first.steps.total_steps == sum(resolved["steps_per_frame"])
    assert all(np.array_equal(a, b) for a, b in zip(first.assembled.frames, second.assembled.frames))
    assert first.numbers == second.numbers
    frames = np.stack(first.assembled.frames)
    assert frames.shape == (4, 64, 64, 4)
    if channels.endswith("unused"):
        assert np.all(frames[..., 3] == 255)
        covered = frames[..., 1] >= 128
        assert covered.any() and frames[covered, 2].max() > 0
    else:
        empty = (frames[..., 0] == 0) & (frames[..., 1] == 0)
        assert np.all(frames[empty, 2:] == 128)
        assert "shade_divisor" not in first.numbers
    cpu, cpu_text = oracle_run(path, False, "tiny.toml")
    worst = assert_sheets_agree(first.assembled.frames, cpu.frames, first.numbers, cpu_text)
    print(channels, "largest sheet byte differences against the oracle:", worst, flush=True)


def test_a_hot_tiny_recipe_sheet_agrees_with_the_oracle_in_heat(device, tmp_path):
    """The tiny recipe with a 900 K emitter (over the 600 K glow start), so the sheet's R channel carries heat:
    heat bytes held to the thickness bound (3)."""
    from plume_amd.pipeline import run_recipe
    from plume_amd.report.numbers import read_numbers

    path = tmp_path / "tiny.toml"
    assert "temperature_k = 500.0" in TINY
    path.write_text(TINY.replace("temperature_k = 500.0", "temperature_k = 900.0"))
    gpu = run_recipe(path, device="cuda", source="tiny.toml")
    cpu, cpu_text = oracle_run(path, False, "tiny.toml")
    assert gpu.assembled.heat_divisor > 0.0
    assert gpu.assembled.heat_divisor == pytest.approx(cpu.heat_divisor, rel=1e-4)
    worst = assert_sheets_agree(gpu.assembled.frames, cpu.frames, gpu.numbers, cpu_text)
    print("hot tiny: largest sheet byte differences against the oracle:", worst, flush=True)
    assert worst["R"] <= 3
    _, rows = read_numbers(gpu.numbers)
    assert max(float(row["peak_heat"]) for row in rows) == pytest.approx(1.0, abs=1.0 / 255)


def test_a_variation_runs_like_the_recipe_it_resolves_to(device, tmp_path):
    """A child recipe (extends the tiny recipe, overrides its seed and, by --set, its emitter radius) gives the
    same frames and numbers as the merged recipe written out whole, and its settings log names the parent and
    the set."""
    import tomli

    from plume_amd.pipeline import run_recipe

    (tmp_path / "tiny.toml").write_text(TINY)
    child = tmp_path / "child.toml"
    child.write_text('extends = "tiny.toml"\nseed = 11\n')
    varied = run_recipe(child, device="cuda", source="child.toml", out=tmp_path / "out",
                        sets=("emitter.puff.radius_m=0.25",))
    merged_text = TINY.replace("seed = 3", "seed = 11").replace("radius_m = 0.2", "radius_m = 0.25")
    assert merged_text != TINY
    merged = tmp_path / "child_merged.toml"
    merged.write_text(merged_text)
    whole = run_recipe(merged, device="cuda", source="child.toml")
    assert varied.recipe.seed == 11 and varied.recipe.emitter[0].radius_m == 0.25
    assert all(np.array_equal(a, b) for a, b in zip(varied.assembled.frames, whole.assembled.frames))
    assert varied.numbers == whole.numbers
    resolved = tomli.loads((tmp_path / "out" / "settings.toml").read_text())["resolved"]
    assert resolved["extends_chain"] == ["tiny.toml"] and resolved["sets"] == ["emitter.puff.radius_m=0.25"]


@pytest.mark.parametrize("channels", ["heat,thickness,shade,unused", "heat,thickness,normal_x,normal_y"])
def test_a_run_writes_the_sheet_and_its_json_and_rewrites_them_byte_for_byte(device, tmp_path, channels):
    """tests/cli/test_run.py:27-74 of the reference on the HIP path: the sheet PNG and JSON beside numbers.txt and
    settings.toml, the JSON's facts, the sheet holding the assembled frames, and identical bytes on a rerun;
    a preview goes into a preview folder."""
    import json

    from PIL import Image

    from plume_amd.pipeline import run_recipe

    path = tmp_path / "tiny.toml"
    path.write_text(TINY + f'channels = "{channels}"\n')
    first = run_recipe(path, device="cuda", out=tmp_path / "one")
    second = run_recipe(path, device="cuda", out=tmp_path / "two")
    names = ("tiny.png", "tiny.json", "numbers.txt", "settings.toml")
    assert sorted(p.name for p in first.files.values()) == sorted(names)
    for name in names:
        assert (tmp_path / "one" / name).is_file(), name
        if name == "tiny.json":
            continue  # written_by names the output folder
        assert (tmp_path / "one" / name).read_bytes() == (tmp_path / "two" / name).read_bytes(), name
    sheet = json.loads((tmp_path / "one" / "tiny.json").read_text())
    assert sheet["frames"] == 4 and sheet["columns"] == 2 and sheet["frame_px"] == 64 and sheet["fps"] == 4.0
    assert sheet["channels"] == channels and sheet["written_by"].startswith("plume run ")
    assert sheet["frame_m"] == round(first.assembled.frame_m, 6)
    image = np.asarray(Image.open(tmp_path / "one" / "tiny.png"))
    assert image.shape == (128, 128, 4)
    for index, frame in enumerate(first.assembled.frames):
        row, column = divmod(index, 2)
        np.testing.assert_array_equal(image[row * 64:(row + 1) * 64, column * 64:(column + 1) * 64], frame)
    if channels.endswith("unused"):
        assert np.all(image[..., 3] == 255)
    else:
        empty = (image[..., 0] == 0) & (image[..., 1] == 0)
        assert np.all(image[empty, 2:] == 128)
    preview = run_recipe(path, preview=True, device="cuda", out=tmp_path / "p")
    assert json.loads((tmp_path / "p" / "preview" / "tiny.json").read_text())["frames"] == len(preview.assembled.frames)
