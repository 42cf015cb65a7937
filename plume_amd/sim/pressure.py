"""Pressure projection (sim/pressure.py of the reference): solve laplacian(q) = divergence(u) for a scaled
pressure q, then subtract gradient(q) from the face velocities. Closed faces are walls; open faces hold
pressure 0 just outside. The solve is passed in: `GaussSeidel` here or `Multigrid` (sim/multigrid.py).
"""

import ctypes
from typing import Protocol

import torch

from plume_amd._lib import call, ptr, stream

# Over-relaxation factor of the Gauss-Seidel solve (sim/pressure.py:18).
RELAXATION = 1.5


class PressureSolve(Protocol):
    def __call__(self, grid) -> None:
        """Improves `grid.pressure` toward the solution of laplacian(q) = grid.divergence."""


def red_black_sweeps(q: torch.Tensor, rhs: torch.Tensor, info, cells, device, sweeps: int, relaxation: float,
                     scratch: torch.Tensor | None = None) -> None:
    """`sweeps` red-black sweeps of laplacian(q) = rhs on the grid `info` describes (sim/pressure.py:135-141).
    With a level-sized `scratch` tensor both colours of a sweep run in one launch (bit-identical)."""
    call("plm_rb_sweeps", ctypes.byref(info), ptr(q), ptr(rhs), ptr(scratch), int(sweeps), relaxation, stream())


def red_black_sweep(q: torch.Tensor, rhs: torch.Tensor, info, color: int, relaxation: float) -> None:
    """One colour of red_black_sweep_kernel (sim/pressure.py:80-92)."""
    call("plm_rb_sweep", ctypes.byref(info), ptr(q), ptr(rhs), int(color), relaxation, stream())


class GaussSeidel:
    """A fixed number of red-black Gauss-Seidel sweeps with over-relaxation (sim/pressure.py:124-132)."""

    def __init__(self, iterations: int):
        self.iterations = iterations
        self._scratch = None

    def __call__(self, grid) -> None:
        if self._scratch is None or self._scratch.shape != grid.pressure.shape:
            self._scratch = torch.empty_like(grid.pressure)
        red_black_sweeps(grid.pressure, grid.divergence, grid.info, grid.cells, grid.device, self.iterations,
                         RELAXATION, self._scratch)


def zero_closed_faces(grid) -> None:
    """zero_closed_faces_kernel on u, v and w (sim/pressure.py:21-28, 146-149)."""
    u, v, w = (f.data for f in grid.velocity_fields())
    call("plm_zero_closed_faces", ctypes.byref(grid.info), ptr(u), ptr(v), ptr(w), stream())


def compute_divergence(grid) -> torch.Tensor:
    """Fills and returns `grid.divergence` from the current face velocities (sim/pressure.py:157-162)."""
    u, v, w = (f.data for f in grid.velocity_fields())
    call("plm_divergence", ctypes.byref(grid.info), ptr(u), ptr(v), ptr(w), ptr(grid.divergence), stream())
    return grid.divergence


def close_faces_and_divergence(grid) -> torch.Tensor:
    """zero_closed_faces then compute_divergence in one launch (sim/pressure.py:146-149): the same faces and
    divergence. Fills and returns `grid.divergence`."""
    u, v, w = (f.data for f in grid.velocity_fields())
    call("plm_close_divergence", ctypes.byref(grid.info), ptr(u), ptr(v), ptr(w), ptr(grid.divergence), stream())
    return grid.divergence


def subtract_gradient(grid) -> None:
    """subtract_gradient_kernel on u, v and w (sim/pressure.py:95-116, 152-154)."""
    u, v, w = (f.data for f in grid.velocity_fields())
    call("plm_subtract_gradient", ctypes.byref(grid.info), ptr(u), ptr(v), ptr(w), ptr(grid.pressure), stream())


def project(grid, solve: PressureSolve) -> None:
    """Makes the face velocities divergence-free (to the solver's accuracy) and walls impermeable
    (sim/pressure.py:144-154)."""
    close_faces_and_divergence(grid)
    solve(grid)
    subtract_gradient(grid)


def boundary_flags(info, device) -> torch.Tensor:
    """Per-cell boundary flags (uint16; two bits per neighbour x-, x+, y-, y+, z-, z+: 0 in-domain, 1 open
    ghost, 2 closed; bits 12-14 mark the last cell per axis): the integer content of neighbor_term's
    branches (sim/pressure.py:52-69)."""
    flags = torch.empty(tuple(info.dims), dtype=torch.int16, device=device)
    call("plm_build_flags", ctypes.byref(info), ptr(flags, torch.int16), stream())
    return flags.view(torch.uint16) if hasattr(torch, "uint16") else flags
