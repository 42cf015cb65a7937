"""Vorticity confinement (Fedkiw, Stam and Jensen 2001), sim/vorticity.py of the reference.

Four passes: centre velocity, curl, the push strength * cell_m * (N x curl), and each face adding the average
push of its two cells. By default the centre-velocity pass is folded into the curl pass (one launch and one
vec3 array fewer, same arithmetic); `fuse=False` runs the reference's four launches.
"""

import ctypes

import torch

from plume_amd._lib import call, ptr, stream


class VorticityConfinement:
    """Scratch fields for confinement on one grid, allocated once per run (sim/vorticity.py:87-93)."""

    def __init__(self, grid, fuse: bool = True):
        shape = tuple(grid.cells) + (3,)
        self.fuse = fuse
        self.center = torch.zeros(shape, dtype=torch.float32, device=grid.device)
        self.curl = torch.zeros(shape, dtype=torch.float32, device=grid.device)
        self.push = torch.zeros(shape, dtype=torch.float32, device=grid.device)

    def apply(self, grid, strength: float, dt: float) -> None:
        """Adds strength * cell_m * (N x curl) * dt to the face velocities; nothing when strength is 0
        (sim/vorticity.py:95-106)."""
        if strength == 0.0:
            return
        u, v, w = (f.data for f in grid.velocity_fields())
        call("plm_confinement", ctypes.byref(grid.info), ptr(u), ptr(v), ptr(w), ptr(self.center), ptr(self.curl),
             ptr(self.push), strength, dt, 1 if self.fuse else 0, stream())

    def compute_curl(self, grid) -> torch.Tensor:
        """Fills and returns the curl from `self.center` (sim/vorticity.py:108-111)."""
        call("plm_curl", ctypes.byref(grid.info), ptr(self.center), ptr(self.curl), stream())
        return self.curl

    def measure_curl(self, grid) -> torch.Tensor:
        """The curl of the grid's current velocity at cell centres (sim/vorticity.py:113-117)."""
        u, v, w = (f.data for f in grid.velocity_fields())
        call("plm_center_velocity", ctypes.byref(grid.info), ptr(u), ptr(v), ptr(w), ptr(self.center), stream())
        return self.compute_curl(grid)
