"""`plume run recipes/smoke_puff.toml --preview` on the HIP path (38 x 48 x 38 cells, 16 frames of 128 px)
against the reference's golden numbers file, with the bounds of the reference's own golden-numbers test
(tests/cli/test_golden_numbers.py:21-63): the HIP simulation and ray march, then the package's sheet assembly
and measurement, must give the approved preview's coverage, boxes, thickness and heat."""

import pytest

from golden_numbers import GOLDEN, RECIPE, SOURCE, assert_matches_golden, largest_differences

pytestmark = pytest.mark.gpu


def test_preview_run_matches_the_golden_numbers(device):
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
