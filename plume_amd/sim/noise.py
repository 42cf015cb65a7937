"""Seeded value noise (sim/noise.py of the reference), evaluated by the GPU kernels' own device functions.

The kernels use it inside emitter coverage and turbulence; these probes expose it for parity tests of the
integer hash and the interpolation.
"""

import torch

from plume_amd._lib import call, ptr, stream

LATTICE_OFFSET = 1048576


def lattice_values(seed: int, stream_id: int, cells: torch.Tensor) -> torch.Tensor:
    """lattice_value (sim/noise.py:24-32) at int32 (i, j, k) rows of `cells` (on the GPU)."""
    out = torch.empty(cells.shape[0], dtype=torch.float32, device=cells.device)
    call("plm_noise_probe", int(seed), int(stream_id), None, ptr(cells.contiguous(), torch.int32), cells.shape[0],
         ptr(out), stream())
    return out


def value_noise(seed: int, stream_id: int, points: torch.Tensor) -> torch.Tensor:
    """value_noise (sim/noise.py:40-57) at float32 (x, y, z) rows of `points` (on the GPU)."""
    out = torch.empty(points.shape[0], dtype=torch.float32, device=points.device)
    call("plm_noise_probe", int(seed), int(stream_id), ptr(points.contiguous()), None, points.shape[0], ptr(out),
         stream())
    return out
