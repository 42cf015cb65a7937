"""The reference's approved preview of the smoke puff and the bounds its CLI test allows
(tests/cli/test_golden_numbers.py:21-63 of the reference), the bounds of its CPU-against-GPU sheet test
(tests/cli/test_cpu_gpu.py:21-33), and a recipe run on the oracle for either comparison.

`tests/golden/smoke_puff_preview.numbers.txt` is the reference's golden numbers file (a Warp CPU run of
`plume run recipes/smoke_puff.toml --preview`: 38 x 48 x 38 cells, 16 frames of 128 px) and
`tests/golden/smoke_puff.toml` the recipe it was made from, both kept here unchanged as fixtures.
"""

from pathlib import Path

import numpy as np
import pytest

GOLDEN_DIR = Path(__file__).resolve().parent / "golden"
RECIPE = GOLDEN_DIR / "smoke_puff.toml"
GOLDEN = GOLDEN_DIR / "smoke_puff_preview.numbers.txt"
SOURCE = "recipes/smoke_puff.toml"  # the recipe path the golden file names

SAME_SUMMARY_KEYS = ("recipe", "grid", "frames")
CLOSE_SUMMARY_KEYS = ("fps", "frame_m")  # relative 1e-6
FRAME_KEYS = ("peak_frame", "fade_frame")
BOX_KEYS = ("left", "top", "right", "bottom")
AMOUNT_KEYS = ("coverage", "mean_thick", "total_thick")
HEAT_KEYS = ("mean_heat", "peak_heat")
RELATIVE = 0.02  # amounts within 2 percent
AMOUNT_FLOOR = 1e-3  # amounts this close count as equal however small
BOX_FRACTION = 0.02  # box edges within about 2 px of a 128 px frame
HEAT_FRACTION = 0.02  # heat values are fractions of the hottest frame
FRAME_SLACK = 1  # peak and fade frames may move by one frame


def assert_matches_golden(actual_text: str, golden_text: str) -> None:
    """Every bound of the reference's golden-numbers test."""
    from plume_amd.report.numbers import read_numbers

    actual, actual_rows = read_numbers(actual_text)
    golden, golden_rows = read_numbers(golden_text)
    for key in SAME_SUMMARY_KEYS:
        assert actual[key] == golden[key], key
    for key in CLOSE_SUMMARY_KEYS:
        assert float(actual[key]) == pytest.approx(float(golden[key]), rel=1e-6), key
    for key in FRAME_KEYS:
        assert (actual[key] == "never") == (golden[key] == "never"), key
        if golden[key] != "never":
            assert abs(int(actual[key]) - int(golden[key])) <= FRAME_SLACK, key
    assert len(actual_rows) == len(golden_rows)
    for row, want in zip(actual_rows, golden_rows):
        where = f"frame {want['frame']}"
        assert row["frame"] == want["frame"]
        assert float(row["time_s"]) == pytest.approx(float(want["time_s"])), where
        for key in BOX_KEYS:
            assert (row[key] == "-") == (want[key] == "-"), f"{where} {key}"
            if want[key] != "-":
                assert float(row[key]) == pytest.approx(float(want[key]), abs=BOX_FRACTION), f"{where} {key}"
        for key in AMOUNT_KEYS:
            assert float(row[key]) == pytest.approx(float(want[key]), rel=RELATIVE, abs=AMOUNT_FLOOR), \
                f"{where} {key}"
        for key in HEAT_KEYS:
            assert float(row[key]) == pytest.approx(float(want[key]), abs=HEAT_FRACTION), f"{where} {key}"


def largest_differences(actual_text: str, golden_text: str) -> dict[str, float]:
    """Per numeric column, the largest |actual - golden| over the rows (for reports and tighter checks)."""
    from plume_amd.report.numbers import read_numbers

    _, actual_rows = read_numbers(actual_text)
    _, golden_rows = read_numbers(golden_text)
    worst = {}
    for row, want in zip(actual_rows, golden_rows):
        for key in BOX_KEYS + AMOUNT_KEYS + HEAT_KEYS:
            if want[key] != "-" and row[key] != "-":
                worst[key] = max(worst.get(key, 0.0), abs(float(row[key]) - float(want[key])))
    return worst


SHEET_G_BYTES = 3  # thickness bytes of the two runs' sheets within 3
SHEET_BA_BYTES = 6  # shade / normal bytes within 6
SHEET_RELATIVE = 0.02  # coverage, mean and total thickness within 2 percent


def oracle_run(recipe_path, preview: bool, source: str):
    """(AssembledFrames, numbers text) of a recipe run on the FP32 oracle (simulation and frame render), then the
    package's sheet assembly and measurement."""
    from oracle import Oracle
    from oracle.render import render_oracle
    from oracle.sim import simulate_frames
    from plume_amd.pipeline import sheet_numbers
    from plume_amd.recipe.derived import derive
    from plume_amd.recipe.load import load_recipe
    from plume_amd.render.camera import sheet_domain_view

    recipe = load_recipe(Path(recipe_path), preview=preview)
    derived = derive(recipe)
    o = Oracle("f32")
    view = sheet_domain_view(recipe, derived)
    raws = [render_oracle(o, grid, view, recipe.render, recipe.output.channels)
            for _, grid in simulate_frames(o, recipe, derived)]
    return sheet_numbers(raws, recipe, derived, "cpu", source)


def assert_sheets_agree(gpu_frames, cpu_frames, gpu_text: str, cpu_text: str) -> dict[str, int]:
    """The CPU-against-GPU bounds on two runs' sheet frames and numbers; returns the largest byte differences
    per channel."""
    from plume_amd.report.numbers import read_numbers

    assert len(gpu_frames) == len(cpu_frames)
    gpu = np.stack(gpu_frames).astype(int)
    cpu = np.stack(cpu_frames).astype(int)
    assert gpu.shape == cpu.shape
    worst = {name: int(np.abs(gpu[..., c] - cpu[..., c]).max()) for c, name in enumerate("RGBA")}
    assert worst["G"] <= SHEET_G_BYTES, worst
    assert max(worst["B"], worst["A"]) <= SHEET_BA_BYTES, worst
    _, gpu_rows = read_numbers(gpu_text)
    _, cpu_rows = read_numbers(cpu_text)
    for gpu_row, cpu_row in zip(gpu_rows, cpu_rows):
        for key in ("coverage", "mean_thick", "total_thick"):
            assert float(gpu_row[key]) == pytest.approx(float(cpu_row[key]), rel=SHEET_RELATIVE), key
    return worst
