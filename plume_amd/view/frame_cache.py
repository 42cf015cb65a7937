"""The frame cache (view/frame_cache.py of the reference): each kept frame's smoke and temperature, and its
velocity at cell centres, as IEEE binary16 on the GPU, under a byte budget.

The plan, fixed before a run from the frame count and grid size: every frame with velocity when that fits;
else every frame without velocity; else every second frame, every fourth, and so on (the smallest power-of-two
stride that fits), without velocity. Playback shows the latest kept frame at or before the one asked for.
Reading a frame widens it into the reader's own FP32 scratch, since the ray march reads FP32 fields.

Kernels: plm_narrow_f16 / plm_widen_f16 (4 values per thread) and plm_narrow_center_velocity (the cell-centre
velocity and its rounding to binary16 in one pass, 3 halves per cell).
"""

import ctypes
import math

import torch

from plume_amd._lib import GridInfo, call, int3, ptr, stream

SCALAR_FIELDS = 2  # smoke and temperature
VELOCITY_VALUES = 3  # one vector per cell
BYTES_PER_VALUE = 2  # binary16


def frame_bytes(cells: tuple[int, int, int], velocity: bool = False) -> int:
    """Bytes one kept frame takes, with or without its velocity (view/frame_cache.py:52-55)."""
    return (SCALAR_FIELDS + (VELOCITY_VALUES if velocity else 0)) * BYTES_PER_VALUE * math.prod(cells)


def frame_stride(frames: int, cells: tuple[int, int, int], budget_bytes: float) -> int:
    """The smallest power of two s whose kept frames 0, s, 2s, ... fit the budget without velocity; one frame is
    always kept (view/frame_cache.py:58-64)."""
    stride = 1
    while stride < frames and math.ceil(frames / stride) * frame_bytes(cells) > budget_bytes:
        stride *= 2
    return stride


def keeps_velocity(frames: int, cells: tuple[int, int, int], budget_bytes: float) -> bool:
    """Whether every frame fits with its velocity; velocity is dropped before any frame (view/frame_cache.py:67-69)."""
    return frames * frame_bytes(cells, velocity=True) <= budget_bytes


def _ordinal(number: int) -> str:
    teen = number % 100 in (11, 12, 13)
    return f"{number}{'th' if teen else {1: 'st', 2: 'nd', 3: 'rd'}.get(number % 10, 'th')}"


def _cells_info(cells: tuple[int, int, int]) -> GridInfo:
    """A plm_grid_info carrying only the cell counts (the centring kernel reads nothing else)."""
    info = GridInfo()
    info.dims = int3(cells)
    info.cell_m = 1.0
    return info


class FrameCache:
    """The kept frames of one run, filled in frame order (view/frame_cache.py:72-182)."""

    def __init__(self, cells: tuple[int, int, int], frames: int, budget_bytes: float, device: str,
                 shared_with_compare: bool = False):
        self.cells = tuple(int(c) for c in cells)
        self.frames = frames
        self.budget_bytes = budget_bytes
        self.device = device
        self.shared_with_compare = shared_with_compare  # the budget is what the Compare run left
        self.has_velocity = keeps_velocity(frames, self.cells, budget_bytes)
        self.stride = 1 if self.has_velocity else frame_stride(frames, self.cells, budget_bytes)
        self._stored: dict[int, tuple[torch.Tensor, torch.Tensor]] = {}
        self._velocity: dict[int, torch.Tensor] = {}
        self._info = _cells_info(self.cells)

    def keeps(self, frame: int) -> bool:
        return frame % self.stride == 0

    def store(self, frame: int, smoke: torch.Tensor, temperature: torch.Tensor,
              velocity: tuple[torch.Tensor, torch.Tensor, torch.Tensor] | None = None) -> None:
        """Rounds a kept frame's fields to binary16 and keeps them (other frames are ignored); with the grid's
        face arrays `velocity` = (u, v, w), also its cell-centre velocity when the plan keeps velocity."""
        if not self.keeps(frame):
            return
        count = math.prod(self.cells)
        kept = []
        for source in (smoke, temperature):
            target = torch.empty(self.cells, dtype=torch.float16, device=self.device)
            call("plm_narrow_f16", count, ptr(source), ptr(target, torch.float16), stream())
            kept.append(target)
        self._stored[frame] = (kept[0], kept[1])
        if self.has_velocity and velocity is not None:
            centre = torch.empty((*self.cells, 3), dtype=torch.float16, device=self.device)
            u, v, w = velocity
            call("plm_narrow_center_velocity", ctypes.byref(self._info), ptr(u), ptr(v), ptr(w),
                 ptr(centre, torch.float16), stream())
            self._velocity[frame] = centre

    def kept_frames(self) -> list[int]:
        return sorted(self._stored)

    def nearest_kept(self, frame: int) -> int | None:
        """The latest kept frame at or before `frame`, else the first kept frame; None when empty."""
        kept = self.kept_frames()
        if not kept:
            return None
        earlier = [index for index in kept if index <= frame]
        return earlier[-1] if earlier else kept[0]

    @property
    def stored_bytes(self) -> int:
        velocity_extra = frame_bytes(self.cells, velocity=True) - frame_bytes(self.cells)
        return len(self._stored) * frame_bytes(self.cells) + len(self._velocity) * velocity_extra

    @property
    def thinned(self) -> bool:
        return self.stride > 1

    def scratch(self) -> tuple[torch.Tensor, torch.Tensor]:
        """An FP32 smoke and temperature pair for one reader's loads."""
        return (torch.zeros(self.cells, dtype=torch.float32, device=self.device),
                torch.zeros(self.cells, dtype=torch.float32, device=self.device))

    def velocity_scratch(self) -> torch.Tensor:
        """An FP32 cell-centre velocity array (nx, ny, nz, 3) for one reader's `load_velocity`."""
        return torch.zeros((*self.cells, 3), dtype=torch.float32, device=self.device)

    def load(self, frame: int, into: tuple[torch.Tensor, torch.Tensor]) -> tuple[torch.Tensor, torch.Tensor]:
        """A kept frame's smoke and temperature widened to FP32 into the reader's pair (from `scratch`)."""
        count = math.prod(self.cells)
        for source, target in zip(self._stored[frame], into, strict=True):
            call("plm_widen_f16", count, ptr(source, torch.float16), ptr(target), stream())
        return into

    def velocity_kept(self, frame: int) -> bool:
        return frame in self._velocity

    def load_velocity(self, frame: int, into: torch.Tensor) -> torch.Tensor:
        """A kept frame's cell-centre velocity (m/s) widened to FP32 into `into` (from `velocity_scratch`)."""
        call("plm_widen_f16", 3 * math.prod(self.cells), ptr(self._velocity[frame], torch.float16), ptr(into),
             stream())
        return into

    def stored_smoke(self, frame: int) -> torch.Tensor:
        return self._stored[frame][0]

    def stored_velocity(self, frame: int) -> torch.Tensor:
        return self._velocity[frame]

    def thinning_note(self) -> str | None:
        """A status line when the budget made the cache drop velocity or skip frames, else None."""
        if self.has_velocity:
            return None
        gigabytes = self.budget_bytes / 2**30
        room = "the room the Compare run leaves of --cache-gb" if self.shared_with_compare else "--cache-gb"
        if self.stride == 1:
            return (f"frame cache keeps no velocity (the slice plane offers smoke and temperature only): "
                    f"{self.frames} frames at {frame_bytes(self.cells, velocity=True) / 2**20:.1f} MB each with "
                    f"velocity do not fit {gigabytes:.3g} GB ({room})")
        return (f"frame cache keeps every {_ordinal(self.stride)} frame and no velocity: {self.frames} frames at "
                f"{frame_bytes(self.cells) / 2**20:.1f} MB each do not fit {gigabytes:.3g} GB ({room})")
