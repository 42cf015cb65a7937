"""The simulation grid: a MAC (staggered) grid of float32 torch tensors on the GPU (sim/grid.py of the
reference).

Positions inside kernels are in cell units with 0 at the domain's minimum corner; a field's samples sit at
`index + offset`. Arrays are C-order [i, j, k] with k contiguous: cells (nx, ny, nz), faces u (nx+1, ny, nz),
v (nx, ny+1, nz), w (nx, ny, nz+1).
"""

from dataclasses import dataclass

import torch

from plume_amd._lib import GridInfo, float3, int3

CENTER = (0.5, 0.5, 0.5)
U_FACE = (0.0, 0.5, 0.5)
V_FACE = (0.5, 0.0, 0.5)
W_FACE = (0.5, 0.5, 0.0)
FACE_OFFSETS = (U_FACE, V_FACE, W_FACE)


def grid_info(dims, cell_m, origin_m=(0.0, 0.0, 0.0), open_lo=(0, 0, 0), open_hi=(0, 0, 0),
              fill_hi=(1.0, 1.0, 1.0), ghost_gap=0.5) -> GridInfo:
    """A GridInfo (sim/grid.py:21-33); float fields are stored as float32, as Warp's struct does."""
    info = GridInfo()
    info.dims = int3(dims)
    info.cell_m = float(cell_m)
    info.origin_m = float3(origin_m)
    info.open_lo = int3(open_lo)
    info.open_hi = int3(open_hi)
    info.fill_hi = float3(fill_hi)
    info.ghost_gap = float(ghost_gap)
    return info


class AdvectedField:
    """One advected quantity: its current values and the scratch arrays MacCormack needs
    (sim/grid.py:36-51). `back` is allocated for the reference's three-pass form; the fused form leaves it
    unused."""

    def __init__(self, shape, offset, outside: float, uses_outside: bool, device: str, initial: float = 0.0):
        self.offset = offset
        self.outside = outside
        self.uses_outside = 1 if uses_outside else 0
        self.data = torch.full(shape, float(initial), dtype=torch.float32, device=device)
        self.hat = torch.zeros(shape, dtype=torch.float32, device=device)
        self.back = torch.zeros(shape, dtype=torch.float32, device=device)
        self.next = torch.zeros(shape, dtype=torch.float32, device=device)
        # the clamp range the forward pass carries to the correction: lo, then hi
        self.bounds = torch.zeros((2,) + tuple(shape), dtype=torch.float32, device=device)

    @property
    def shape(self):
        return tuple(self.data.shape)

    def swap(self) -> None:
        """Makes the freshly advected values current."""
        self.data, self.next = self.next, self.data


@dataclass
class Grid:
    info: GridInfo
    device: str
    cells: tuple[int, int, int]
    u: AdvectedField
    v: AdvectedField
    w: AdvectedField
    smoke: AdvectedField
    temperature: AdvectedField
    pressure: torch.Tensor
    divergence: torch.Tensor

    def velocity_fields(self) -> tuple[AdvectedField, AdvectedField, AdvectedField]:
        return self.u, self.v, self.w


def face_flags(faces) -> tuple[tuple[int, int, int], tuple[int, int, int]]:
    """open_lo, open_hi from FaceSettings (sim/grid.py:78-79, 97-98)."""
    is_open = lambda kind: 1 if kind == "open" else 0  # noqa: E731
    return ((is_open(faces.left), is_open(faces.bottom), is_open(faces.back)),
            (is_open(faces.right), is_open(faces.top), is_open(faces.front)))


def make_grid(derived, cell_m: float, faces, ambient_k: float, device: str) -> Grid:
    """Allocates a grid at rest, filled with ambient air (sim/grid.py:71-94)."""
    nx, ny, nz = derived.cells
    open_lo, open_hi = face_flags(faces)
    info = grid_info((nx, ny, nz), cell_m, derived.domain_min_m, open_lo, open_hi)
    center_shape = (nx, ny, nz)
    return Grid(
        info=info, device=device, cells=(nx, ny, nz),
        u=AdvectedField((nx + 1, ny, nz), U_FACE, 0.0, False, device),
        v=AdvectedField((nx, ny + 1, nz), V_FACE, 0.0, False, device),
        w=AdvectedField((nx, ny, nz + 1), W_FACE, 0.0, False, device),
        smoke=AdvectedField(center_shape, CENTER, 0.0, True, device),
        temperature=AdvectedField(center_shape, CENTER, ambient_k, True, device, initial=ambient_k),
        pressure=torch.zeros(center_shape, dtype=torch.float32, device=device),
        divergence=torch.zeros(center_shape, dtype=torch.float32, device=device),
    )
