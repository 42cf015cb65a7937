"""The sheet image and its JSON, the two files a game reads from a `plume run` (the interface of
sheet/pack.py:12-30 and sheet/sheet_json.py:11-41, written from cli/run.py:98-115).

The sheet holds the frames left to right, then top to bottom, in a grid `columns` wide; grid slots past the
last frame are transparent black. It is an 8-bit RGBA PNG written with a fixed compression level and no
metadata, so equal pixels always give equal bytes. The JSON beside it lists, in this order: frames, columns,
frame_px, fps, frame_m, seat (x, y in frame fractions), channels, source, tool and written_by, floats rounded
to 6 decimals, indented by 2, with a trailing newline.
"""

import json
import math
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from plume_amd import __version__

PNG_COMPRESS_LEVEL = 6
JSON_DECIMALS = 6
TOOL = f"plume_amd {__version__}"


def sheet_image(frames: list[np.ndarray], columns: int) -> np.ndarray:
    """The frames tiled into one (rows px, columns px, channels) image, rows = ceil(frames / columns)."""
    if not frames or columns < 1:
        raise ValueError("a sheet needs at least one frame and one column")
    px, _, channels = frames[0].shape
    rows = math.ceil(len(frames) / columns)
    slots = np.zeros((rows * columns, px, px, channels), dtype=frames[0].dtype)
    slots[: len(frames)] = np.stack(frames)
    # (row, column, y, x, c) -> (row, y, column, x, c): each grid row's frames side by side
    return slots.reshape(rows, columns, px, px, channels).transpose(0, 2, 1, 3, 4).reshape(rows * px, columns * px,
                                                                                          channels)


def write_png(path: Path, image: np.ndarray) -> None:
    from PIL import Image

    Image.fromarray(np.ascontiguousarray(image, dtype=np.uint8)).save(path, format="PNG",
                                                                      compress_level=PNG_COMPRESS_LEVEL)


@dataclass(frozen=True)
class SheetFacts:
    """What the sheet's JSON states about it (SheetInfo, sheet/sheet_json.py:11-23)."""

    frames: int
    columns: int
    frame_px: int
    fps: float
    frame_m: float
    seat: tuple[float, float]
    channels: str
    source: str
    written_by: str
    tool: str = TOOL


def sheet_json(facts: SheetFacts) -> str:
    def rounded(x: float) -> float:
        return round(float(x), JSON_DECIMALS)

    ordered = (("frames", facts.frames), ("columns", facts.columns), ("frame_px", facts.frame_px),
               ("fps", rounded(facts.fps)), ("frame_m", rounded(facts.frame_m)),
               ("seat", [rounded(facts.seat[0]), rounded(facts.seat[1])]), ("channels", facts.channels),
               ("source", facts.source), ("tool", facts.tool), ("written_by", facts.written_by))
    return json.dumps(dict(ordered), indent=2) + "\n"


def recipe_source(recipe_path: Path) -> str:
    """The recipe's path as the JSON names it: relative to the root of the git repository holding it, with
    forward slashes, or as given outside one (sheet/sheet_json.py:36-41)."""
    resolved = Path(recipe_path).resolve()
    root = next((folder for folder in resolved.parents if (folder / ".git").exists()), None)
    return resolved.relative_to(root).as_posix() if root is not None else Path(recipe_path).as_posix()
