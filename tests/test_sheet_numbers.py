"""The sheet frames and the numbers file (host code, CPU), and the oracle's preview of the smoke puff against
the reference's golden numbers file.

Unit checks follow tests/sheet/test_assemble.py, tests/sheet/test_shade_sheet.py and
tests/report/test_reports.py of the reference, on frames built here, plus closed forms of the blur and the
resampler. The golden check runs `plume run recipes/smoke_puff.toml --preview` on the FP32 oracle (38 x 48 x
38 cells, 16 frames) through the package's sheet and report code and applies the reference's own bounds.
"""

import numpy as np
import pytest

from golden_numbers import GOLDEN, RECIPE, SOURCE, assert_matches_golden, largest_differences, oracle_run
from plume_amd.recipe.schema import NORMAL_LAYOUT, SHADE_LAYOUT, RenderSettings
from plume_amd.render.camera import make_camera, make_domain_view
from plume_amd.render.raw_frame import CHANNELS, HEAT, NORMAL_X, NORMAL_Y, NORMAL_Z, SHADE, THICKNESS
from plume_amd.report.numbers import RunFacts, measure, numbers_text, peak_and_fade, read_numbers
from plume_amd.sheet.assemble import GutterError, assemble
from plume_amd.sheet.crop import CONTENT_LEVEL, CropSquare, NothingVisibleError, content_box, square_crop
from plume_amd.sheet.quantize import quantize, to_byte
from plume_amd.sheet.resample import resample_frame
from plume_amd.sheet.shade import ease_thin_pixels, shade_divisor
from plume_amd.sheet.soften import blur, edge_clamped_operator, gaussian_weights, soften

VIEW = make_domain_view(make_camera(0.0, 0.0), (-1.0, 0.0, -1.0), (2.0, 2.0, 2.0), 0.0625, 2.0)  # 64 x 64 px


def disc_frame(center, radius, thickness=0.8, heat=0.0, shade=None, size=64, layout=NORMAL_LAYOUT):
    frame = np.zeros((size, size, CHANNELS), dtype=np.float32)
    if layout == NORMAL_LAYOUT:
        frame[..., NORMAL_Z] = 1.0
    ys, xs = np.mgrid[0:size, 0:size] + 0.5
    inside = (xs - center[0]) ** 2 + (ys - center[1]) ** 2 < radius * radius
    frame[inside, THICKNESS] = thickness
    frame[inside, HEAT] = heat
    if shade is not None:
        frame[inside, SHADE] = thickness * shade(xs[inside], ys[inside])
    return frame


def square_frame(g_byte, r_byte=0, size=32, lo=8, hi=24):
    frame = np.zeros((size, size, 4), dtype=np.uint8)
    frame[..., 2:] = 128
    frame[lo:hi, lo:hi, 0] = r_byte
    frame[lo:hi, lo:hi, 1] = g_byte
    return frame


# ---------------------------------------------------------------------- closed forms

def test_gaussian_weights_are_normalized_and_reach_three_sigma():
    for sigma in (0.75, 1.3, 4.0):
        w = gaussian_weights(sigma)
        assert len(w) == 2 * int(np.ceil(3 * sigma)) + 1
        assert w.sum() == pytest.approx(1.0, abs=1e-15)
        assert np.array_equal(w, w[::-1]) and np.argmax(w) == len(w) // 2
        assert w[0] / w[len(w) // 2] == pytest.approx(np.exp(-np.ceil(3 * sigma) ** 2 / (2 * sigma**2)))


def test_edge_clamped_operator_rows_sum_to_one_and_match_a_padded_convolution():
    w = gaussian_weights(1.3)
    op = edge_clamped_operator(11, w)
    assert np.allclose(op.sum(axis=1), 1.0, atol=1e-15)
    x = np.random.default_rng(1).random(11)
    r = len(w) // 2
    padded = np.concatenate([np.full(r, x[0]), x, np.full(r, x[-1])])
    direct = np.array([np.dot(w, padded[i:i + len(w)]) for i in range(11)])
    assert np.allclose(op @ x, direct, rtol=0, atol=1e-15)


def test_blur_keeps_constants_spreads_an_impulse_as_the_weights_and_conserves_it():
    image = np.full((20, 24, CHANNELS), 0.37)
    assert np.allclose(blur(image, 1.1), 0.37, atol=1e-15)
    impulse = np.zeros((21, 21, 2))
    impulse[10, 10, 1] = 1.0
    blurred = blur(impulse, 0.75)
    w = gaussian_weights(0.75)
    r = len(w) // 2
    assert np.allclose(blurred[10 - r:11 + r, 10 - r:11 + r, 1], np.outer(w, w), atol=1e-16)
    assert blurred[..., 1].sum() == pytest.approx(1.0)
    assert not blurred[..., 0].any()
    assert blur(impulse, 0.0) is impulse


def test_soften_renormalizes_normals_toward_the_camera():
    frame = np.zeros((8, 8, CHANNELS))
    frame[..., NORMAL_X] = 3.0
    frame[..., NORMAL_Z] = -4.0
    frame[0, 0, NORMAL_X:NORMAL_Z + 1] = 0.0
    out = soften(frame, 0.0, NORMAL_LAYOUT)
    assert np.allclose(out[1, 1, NORMAL_X:NORMAL_Z + 1], [0.6, 0.0, 0.8])
    assert np.array_equal(out[0, 0, NORMAL_X:NORMAL_Z + 1], [0.0, 0.0, 1.0])
    assert soften(frame, 0.0, SHADE_LAYOUT) is frame


def test_resampling_an_aligned_unit_scale_crop_copies_the_pixels():
    # bicubic at scale 1 on pixel-aligned boxes has weights (0, 1, 0): the frame is the raw crop, clamped
    raw = np.random.default_rng(2).random((40, 50, CHANNELS)).astype(np.float32) * 1.4 - 0.2
    frame = resample_frame(raw, CropSquare(left=7.0, top=5.0, side=24.0), 24, NORMAL_LAYOUT)
    want = raw[5:29, 7:31].astype(np.float64)
    assert np.array_equal(frame[..., NORMAL_X:], want[..., NORMAL_X:])
    assert np.array_equal(frame[..., THICKNESS], np.clip(want[..., THICKNESS], 0.0, 1.0))
    assert np.array_equal(frame[..., HEAT], np.maximum(want[..., HEAT], 0.0))


def test_resampling_past_the_view_reads_empty_space():
    raw = np.zeros((16, 16, CHANNELS), dtype=np.float32)
    raw[..., NORMAL_Z] = 1.0
    raw[..., THICKNESS] = 0.5
    frame = resample_frame(raw, CropSquare(left=-16.0, top=0.0, side=16.0), 16, NORMAL_LAYOUT)
    assert not frame[:, :13, THICKNESS].any()  # three pixels of the left neighbour of the view reach in
    assert np.allclose(frame[..., NORMAL_Z][:, :13], 1.0)


def test_resampling_keeps_shade_within_its_thickness():
    frame = disc_frame((32, 32), 12, thickness=0.9, shade=lambda x, y: 0.9 - 0.6 * (x - 20) / 24,
                       layout=SHADE_LAYOUT)
    out = resample_frame(frame, CropSquare(left=3.3, top=2.7, side=57.0), 40, SHADE_LAYOUT)
    assert np.all(out[..., SHADE] >= 0.0) and np.all(out[..., SHADE] <= out[..., THICKNESS])
    assert np.all(out[..., THICKNESS] <= 1.0) and np.all(out[..., HEAT] >= 0.0)


def test_content_box_and_square_crop():
    frames = [disc_frame((20, 40), 4), disc_frame((40, 20), 3)]
    assert content_box(frames) == (16, 17, 43, 44)  # pixel centres x + 0.5 inside (16, 24) and (37, 43)
    faint = np.zeros((64, 64, CHANNELS), dtype=np.float32)
    faint[5, 6, THICKNESS] = CONTENT_LEVEL
    faint[50, 60, THICKNESS] = np.nextafter(np.float32(CONTENT_LEVEL), np.float32(0))
    assert content_box([faint]) == (6, 5, 7, 6)
    hot = np.zeros((64, 64, CHANNELS), dtype=np.float32)
    hot[10, 10, HEAT] = 100.0
    hot[30, 40, HEAT] = 100.0 * CONTENT_LEVEL
    assert content_box([hot]) == (10, 10, 41, 31)
    crop = square_crop((10, 20, 30, 60), 0.1)
    assert (crop.left, crop.top, crop.side) == pytest.approx((-4.0, 16.0, 48.0))
    with pytest.raises(NothingVisibleError):
        content_box([np.zeros((8, 8, CHANNELS), dtype=np.float32)])


def test_quantize_rounds_half_up_and_encodes_normals_with_y_up():
    assert list(to_byte(np.array([-0.1, 0.5 / 255, np.nextafter(0.5 / 255, 0), 0.5, 1.2]))) == [0, 1, 0, 128, 255]
    frame = np.zeros((2, 2, CHANNELS))
    frame[..., THICKNESS] = 0.5
    frame[0, 0, NORMAL_Z] = 1.0
    frame[0, 1, NORMAL_Y] = 1.0
    frame[1, 0, NORMAL_X] = -1.0
    frame[1, 1, THICKNESS] = 0.0
    rgba = quantize(frame, 0.0, NORMAL_LAYOUT)
    assert tuple(rgba[0, 0]) == (0, 128, 128, 128)
    assert rgba[0, 1, 3] == 255 and rgba[1, 0, 2] == 0
    assert tuple(rgba[1, 1]) == (0, 0, 128, 128)
    shade = np.zeros((1, 2, CHANNELS))
    shade[..., HEAT] = [2.0, 4.0]
    shade[..., SHADE] = [0.25, 0.5]
    assert [tuple(p) for p in quantize(shade, 4.0, SHADE_LAYOUT, 0.5)[0]] == [(128, 0, 128, 255), (255, 0, 255, 255)]


# ---------------------------------------------------------------------- shade easing and divisor

def test_easing_keeps_thick_pixels_and_leaves_an_empty_frame_at_zero():
    frame = disc_frame((32, 32), 12, thickness=0.9, shade=lambda x, y: 0.9 - 0.6 * (x - 20) / 24,
                       layout=SHADE_LAYOUT).astype(np.float64)
    eased = ease_thin_pixels(frame, 0.1, 4.0)
    thick = frame[..., THICKNESS] >= 0.1
    assert np.allclose(eased[thick, SHADE], frame[thick, SHADE] / frame[thick, THICKNESS])
    assert not ease_thin_pixels(np.zeros_like(frame), 0.1, 4.0)[..., SHADE].any()


def test_an_empty_pixel_takes_the_neighbourhood_formula():
    frame = np.zeros((16, 16, CHANNELS))
    frame[8, 8, THICKNESS] = 0.5
    frame[8, 8, SHADE] = 0.5 * 0.6
    eased = ease_thin_pixels(frame, 0.1, 1.5)
    w = gaussian_weights(1.5)
    r = len(w) // 2
    g = w[r + 2] * w[r + 1]  # pixel (10, 9) is (2, 1) from the smoke
    m = 1.0 / 255.0
    assert eased[10, 9, SHADE] == pytest.approx((0.3 * g + m * 0.6) / (0.5 * g + m), rel=1e-12)
    far = (0.0 + m * 0.6) / (0.0 + m)
    assert eased[0, 0, SHADE] == pytest.approx(far)


def test_shade_divisor_falls_back_when_nothing_is_well_covered():
    thin = np.zeros((4, 4, CHANNELS))
    thin[1, 1, THICKNESS] = 0.01
    thin[1, 1, SHADE] = 0.5
    assert shade_divisor([thin], 0.1) == pytest.approx(0.5)
    thick = thin.copy()
    thick[2, 2, THICKNESS] = 0.2
    thick[2, 2, SHADE] = 0.3
    assert shade_divisor([thin, thick], 0.1) == pytest.approx(0.3)
    assert shade_divisor([np.zeros((4, 4, CHANNELS))], 0.1) == 1.0


# ---------------------------------------------------------------------- assembly

def rising():
    return [disc_frame((32, 50 - 8 * i), 6 + 2 * i, heat=10.0 / (i + 1)) for i in range(4)]


def test_assembled_frames_keep_an_empty_gutter_and_normalize_heat():
    result = assemble(rising(), VIEW, np.array([0.0, 0.2, 0.0]), 64, RenderSettings(), NORMAL_LAYOUT)
    for frame in result.frames:
        assert frame.shape == (64, 64, 4) and frame.dtype == np.uint8
        assert not frame[:2, :, :2].any() and not frame[-2:, :, :2].any()
        assert not frame[:, :2, :2].any() and not frame[:, -2:, :2].any()
    assert max(int(f[..., 0].max()) for f in result.frames) == 255
    assert result.frames[0][..., 0].max() > result.frames[3][..., 0].max()
    assert result.shade_divisor is None


def test_frame_m_and_seat_follow_the_crop():
    result = assemble(rising(), VIEW, np.array([0.0, 0.0, 0.0]), 64,
                      RenderSettings(soften_px=0.0, crop_margin=0.0), NORMAL_LAYOUT)
    # the union of the discs spans rows 14..55 (42 px); with no margin their edge lands in the gutter, so the
    # margin grows by 0.02 until it clears
    assert result.margin == pytest.approx(0.06)
    assert result.crop.side == pytest.approx(42.0 * (1.0 + 2.0 * result.margin))
    assert result.frame_m == pytest.approx(result.crop.side * VIEW.pixel_m)
    seat_x, seat_y = result.seat
    assert seat_x == pytest.approx(0.5, abs=0.02)
    assert seat_y > 1.0


def test_margin_grows_until_the_gutter_is_empty():
    frame = disc_frame((32, 32), 10)
    result = assemble([frame], VIEW, np.zeros(3), 32, RenderSettings(soften_px=0.75, crop_margin=0.0), NORMAL_LAYOUT)
    assert result.margin > 0.0
    with pytest.raises(GutterError):
        assemble([frame], VIEW, np.zeros(3), 32, RenderSettings(soften_px=12.0, crop_margin=0.0), NORMAL_LAYOUT)


def test_shade_layout_alpha_is_unused_and_the_brightest_thick_smoke_maps_to_255():
    frame = disc_frame((32, 32), 12, thickness=0.9, shade=lambda x, y: 0.9 - 0.6 * (x - 20) / 24,
                       layout=SHADE_LAYOUT)
    result = assemble([frame], VIEW, np.zeros(3), 64, RenderSettings(), SHADE_LAYOUT)
    out = result.frames[0]
    assert np.all(out[..., 3] == 255)
    dense = out[..., 1] >= 0.1 * 255
    assert out[dense, 2].max() == 255
    assert result.shade_divisor == pytest.approx(0.9, abs=0.05)
    empty = out[..., 1] == 0
    left, right = empty & (np.arange(64) < 32), empty & (np.arange(64) >= 32)
    assert out[left, 2].mean() > out[right, 2].mean()  # the lit side's edge stays lit


# ---------------------------------------------------------------------- numbers

def test_measure_reports_coverage_box_and_means():
    row = measure([square_frame(204, r_byte=51)], fps=20.0)[0]
    assert row.time_s == pytest.approx(0.05)
    assert row.coverage == pytest.approx(0.25)
    assert row.box == pytest.approx((0.25, 0.25, 0.75, 0.75))
    assert row.mean_thick == pytest.approx(0.8) and row.mean_heat == pytest.approx(0.2)
    assert row.total_thick == pytest.approx(256 * 0.8) and row.peak_heat == pytest.approx(0.2)
    faint = measure([square_frame(7)], fps=1.0)[0]  # under G = 8: nothing counts as covered
    assert faint.coverage == 0.0 and faint.box is None and faint.total_thick == pytest.approx(256 * 7 / 255)


def test_peak_and_fade_frames():
    frames = [square_frame(b) for b in (50, 200, 150, 30, 10)]
    assert peak_and_fade(measure(frames, 10.0)) == (1, 4)
    assert peak_and_fade(measure(frames[:3], 10.0)) == (1, None)


def test_numbers_text_round_trips_and_keeps_the_golden_layout():
    frames = [square_frame(b) for b in (50, 200, 0)]
    facts = RunFacts(recipe="recipes/x.toml", backend="cpu", grid="8 x 8 x 8 cells of 0.100000 m", fps=10.0,
                     frame_m=1.5, heat_divisor=0.0, channels=SHADE_LAYOUT, shade_divisor=0.75)
    text = numbers_text(measure(frames, 10.0), facts)
    summary, rows = read_numbers(text)
    assert summary["shade_divisor"] == "0.750000" and summary["peak_frame"] == "1"
    assert summary["fade_frame"] == "2" and summary["grid"] == "8 x 8 x 8 cells of 0.100000 m"
    assert rows[2]["left"] == "-"
    assert float(rows[1]["mean_thick"]) == pytest.approx(200 / 255, abs=1e-4)
    normal = RunFacts(**{**facts.__dict__, "channels": NORMAL_LAYOUT, "shade_divisor": None})
    assert "shade_divisor" not in read_numbers(numbers_text(measure(frames, 10.0), normal))[0]
    # the golden file's own lines are this layout: its header and rows re-render unchanged
    golden = GOLDEN.read_text().splitlines()
    header = golden.index("") + 1
    from plume_amd.report.numbers import COLUMNS, row_text

    assert golden[header] == row_text(COLUMNS)
    assert golden[header + 1] == row_text(golden[header + 1].split())


# ---------------------------------------------------------------------- the golden preview on the oracle

@pytest.fixture(scope="module")
def oracle_preview():
    return oracle_run(RECIPE, True, SOURCE)


def test_oracle_preview_matches_the_golden_numbers(oracle_preview):
    assembled, text = oracle_preview
    golden = GOLDEN.read_text()
    assert_matches_golden(text, golden)
    # tighter than the reference's bounds: the FP32 oracle reproduces the Warp CPU run to the printed digits
    # but for single G steps (measured: one pixel's byte in one frame's total_thick)
    summary, _ = read_numbers(text)
    assert summary["frame_m"] == read_numbers(golden)[0]["frame_m"]
    assert float(summary["shade_divisor"]) == pytest.approx(float(read_numbers(golden)[0]["shade_divisor"]),
                                                             abs=2e-6)
    worst = largest_differences(text, golden)
    assert max(worst[k] for k in ("left", "top", "right", "bottom")) <= 1.0 / 128 + 1e-9
    assert worst["total_thick"] <= 4.0 / 255 and worst["coverage"] <= 4.0 / 128**2 + 1e-9
