"""The numbers file (report/numbers.py of the reference): a block of `key value` lines, a blank line, then one
fixed-width row of measurements per frame, measured from the quantized sheet frames (what a game gets).

Row fields: frame, time_s = (frame + 1) / fps, coverage = fraction of pixels with G >= 8, the box of those
pixels (left, top, right, bottom as frame fractions, right and bottom exclusive; "-" when none), mean heat and
mean thickness (R / 255 and G / 255 over covered pixels), total thickness (G / 255 over all pixels) and peak
heat (largest R / 255). The summary names the frame of greatest total thickness and the first later frame
under a tenth of it ("never" when none).
"""

from dataclasses import dataclass

import numpy as np

COVERED_LEVEL = 8
FADE_FRACTION = 0.1
KEY_WIDTH = 18
COLUMNS = ("frame", "time_s", "coverage", "left", "top", "right", "bottom", "mean_heat", "mean_thick",
           "total_thick", "peak_heat")
WIDTHS = (5, 8, 9, 7, 7, 7, 7, 10, 11, 12, 10)


@dataclass(frozen=True)
class FrameNumbers:
    frame: int
    time_s: float
    coverage: float
    box: tuple[float, float, float, float] | None  # left, top, right, bottom as frame fractions
    mean_heat: float
    mean_thick: float
    total_thick: float
    peak_heat: float


@dataclass(frozen=True)
class RunFacts:
    recipe: str
    backend: str
    grid: str
    fps: float
    frame_m: float
    heat_divisor: float
    channels: str
    shade_divisor: float | None  # None in the normal layout, which has no shade


def measure_frame(index: int, frame: np.ndarray, time_s: float) -> FrameNumbers:
    """One row from a (size, size, 4) uint8 frame (report/numbers.py:84-102)."""
    heat = frame[..., 0] / 255.0
    thick = frame[..., 1] / 255.0
    covered = frame[..., 1] >= COVERED_LEVEL
    size = frame.shape[0]
    box, mean_heat, mean_thick = None, 0.0, 0.0
    if covered.any():
        rows = np.flatnonzero(covered.any(axis=1))
        columns = np.flatnonzero(covered.any(axis=0))
        box = (columns[0] / size, rows[0] / size, (columns[-1] + 1) / size, (rows[-1] + 1) / size)
        mean_heat = float(heat[covered].mean())
        mean_thick = float(thick[covered].mean())
    return FrameNumbers(frame=index, time_s=time_s, coverage=float(covered.mean()), box=box, mean_heat=mean_heat,
                        mean_thick=mean_thick, total_thick=float(thick.sum()), peak_heat=float(heat.max()))


def measure(frames: list[np.ndarray], fps: float) -> list[FrameNumbers]:
    """A row per quantized RGBA frame, frame i at time (i + 1) / fps (report/numbers.py:44-46)."""
    return [measure_frame(index, frame, (index + 1) / fps) for index, frame in enumerate(frames)]


def peak_and_fade(rows: list[FrameNumbers]) -> tuple[int, int | None]:
    """The first frame of greatest total thickness and the first later frame under FADE_FRACTION of it
    (report/numbers.py:49-56)."""
    totals = np.array([row.total_thick for row in rows])
    peak = int(np.argmax(totals))
    later = [row.frame for row in rows[peak + 1:] if row.total_thick < FADE_FRACTION * totals[peak]]
    return peak, (later[0] if later else None)


def row_text(values) -> str:
    return " ".join(str(value).rjust(width) for value, width in zip(values, WIDTHS)).rstrip()


def row_values(row: FrameNumbers) -> list[str]:
    box = [f"{edge:.4f}" for edge in row.box] if row.box else ["-"] * 4
    return [str(row.frame), f"{row.time_s:.4f}", f"{row.coverage:.4f}", *box, f"{row.mean_heat:.4f}",
            f"{row.mean_thick:.4f}", f"{row.total_thick:.4f}", f"{row.peak_heat:.4f}"]


def numbers_text(rows: list[FrameNumbers], facts: RunFacts) -> str:
    """The whole file (report/numbers.py:59-81)."""
    peak, fade = peak_and_fade(rows)
    summary = [("recipe", facts.recipe), ("backend", facts.backend), ("grid", facts.grid),
               ("frames", str(len(rows))), ("fps", f"{facts.fps:.6f}"), ("frame_m", f"{facts.frame_m:.6f}"),
               ("heat_divisor", f"{facts.heat_divisor:.6f} K m"), ("channels", facts.channels)]
    if facts.shade_divisor is not None:
        summary.append(("shade_divisor", f"{facts.shade_divisor:.6f}"))
    summary += [("peak_frame", str(peak)), ("fade_frame", "never" if fade is None else str(fade))]
    lines = [key.ljust(KEY_WIDTH) + value for key, value in summary]
    lines += ["", row_text(COLUMNS)] + [row_text(row_values(row)) for row in rows]
    return "\n".join(lines) + "\n"


def read_numbers(text: str) -> tuple[dict[str, str], list[dict[str, str]]]:
    """The summary (key -> value text) and the rows (column -> value text) of a numbers file
    (report/numbers.py:117-127)."""
    head, _, table = text.partition("\n\n")
    summary = {}
    for line in head.splitlines():
        key, _, value = line.partition(" ")
        summary[key] = value.strip()
    lines = table.strip().splitlines()
    names = lines[0].split()
    return summary, [dict(zip(names, line.split())) for line in lines[1:]]
