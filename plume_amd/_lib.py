"""ctypes binding of libplume_hip.so (include/plume_hip.h) and the tensor plumbing every operator uses.

The library is the only compute path: importing this module fails loudly when the shared library is
missing, and every call raises `PlumeHipError` with the library's message when it returns non-zero.
"""

import ctypes
from pathlib import Path

import torch

LIBRARY = Path(__file__).resolve().parent / "libplume_hip.so"

_int3 = ctypes.c_int32 * 3
_float3 = ctypes.c_float * 3


class PlumeHipError(RuntimeError):
    """A libplume_hip call failed (bad argument or failed launch)."""


class GridInfo(ctypes.Structure):
    """plm_grid_info: GridInfo of sim/grid.py:21-33, passed by value to every kernel."""

    _fields_ = [("dims", _int3), ("cell_m", ctypes.c_float), ("origin_m", _float3), ("open_lo", _int3),
                ("open_hi", _int3), ("fill_hi", _float3), ("ghost_gap", ctypes.c_float)]

    def copy(self) -> "GridInfo":
        other = GridInfo()
        ctypes.memmove(ctypes.byref(other), ctypes.byref(self), ctypes.sizeof(GridInfo))
        return other

    def __repr__(self) -> str:
        return (f"GridInfo(dims={tuple(self.dims)}, cell_m={self.cell_m}, origin_m={tuple(self.origin_m)}, "
                f"open_lo={tuple(self.open_lo)}, open_hi={tuple(self.open_hi)}, fill_hi={tuple(self.fill_hi)}, "
                f"ghost_gap={self.ghost_gap})")


class EmitterParams(ctypes.Structure):
    """plm_emitter: EmitterParams of sim/emitters.py:23-38."""

    _fields_ = [("shape", ctypes.c_int32), ("center_m", _float3), ("radius_m", ctypes.c_float),
                ("half_size_m", _float3), ("edge_softness", ctypes.c_float), ("start_s", ctypes.c_float),
                ("end_s", ctypes.c_float), ("smoke_per_s", ctypes.c_float), ("temperature_k", ctypes.c_float),
                ("velocity_mps", _float3), ("outward", ctypes.c_int32), ("noise_scale_m", ctypes.c_float),
                ("noise_strength", ctypes.c_float), ("stream", ctypes.c_int32)]


class ViewParams(ctypes.Structure):
    """plm_view: ViewParams of render/raymarch.py:30-42."""

    _fields_ = [("right", _float3), ("up", _float3), ("toward", _float3), ("left_m", ctypes.c_float),
                ("top_m", ctypes.c_float), ("pixel_m", ctypes.c_float), ("perspective", ctypes.c_int32),
                ("eye", _float3)]


class LookParams(ctypes.Structure):
    """plm_look: LookParams of render/raymarch.py:45-52."""

    _fields_ = [("opacity_per_density", ctypes.c_float), ("glow_start_k", ctypes.c_float),
                ("smoke_hides_heat", ctypes.c_float), ("normal_fade_thickness", ctypes.c_float),
                ("shade_layout", ctypes.c_int32), ("shade_fill", ctypes.c_float)]


_P = ctypes.c_void_p
_F = ctypes.c_float
_I = ctypes.c_int32
_L = ctypes.c_int64
_D = ctypes.c_double
_INFO = ctypes.POINTER(GridInfo)

# name -> argument types, in include/plume_hip.h order
SIGNATURES = {
    "plm_advect_sl": [_INFO, _P, _P, _P, _P, _F, _I, _P, _P, _P, _F, _P],
    "plm_maccormack_correct": [_INFO, _P, _P, _P, _P, _P, _P, _F, _I, _P, _P, _P, _F, _P],
    "plm_advect_maccormack": [_INFO, _P, _P, _P, _P, _P, _P, _F, _I, _P, _P, _P, _F, _I, _P],
    "plm_advect_velocity": [_INFO, _P, _P, _P, _P, _P, _P, _P, _P, _P, _F, _P],
    "plm_advect_velocity_bounded": [_INFO, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _F, _P],
    "plm_advect_scalars": [_INFO, _P, _P, _P, _P, _F, _P, _P, _P, _P, _F, _P, _P, _P, _F, _P],
    "plm_buoyancy": [_INFO, _P, _P, _P, _F, _F, _F, _F, _F, _P],
    "plm_drag": [_INFO, _P, _P, _P, _P, _F, _P],
    "plm_center_velocity": [_INFO, _P, _P, _P, _P, _P],
    "plm_curl": [_INFO, _P, _P, _P],
    "plm_confinement_push": [_INFO, _P, _P, _F, _P],
    "plm_add_face_push": [_INFO, _P, _P, _P, _P, _F, _P],
    "plm_confinement": [_INFO, _P, _P, _P, _P, _P, _P, _F, _F, _I, _P],
    "plm_emit": [_INFO, _P, _P, _P, _P, _P, _P, _I, _I, _F, _F, _P],
    "plm_emit_box": [_INFO, _P, _P, _P, _P, _P, _P, _I, _I, _F, _F, _P, _P, _P],
    "plm_noise_probe": [_I, _I, _P, _P, _I, _P, _P],
    "plm_turbulence": [_INFO, _P, _P, _P, _I, _F, _F, _F, _F, _P],
    "plm_dissipate": [_L, _P, _P, _F, _F, _F, _P],
    "plm_zero_closed_faces": [_INFO, _P, _P, _P, _P],
    "plm_divergence": [_INFO, _P, _P, _P, _P, _P],
    "plm_close_divergence": [_INFO, _P, _P, _P, _P, _P],
    "plm_build_flags": [_INFO, _P, _P],
    "plm_rb_sweep": [_INFO, _P, _P, _I, _F, _P],
    "plm_rb_sweeps": [_INFO, _P, _P, _P, _I, _F, _P],
    "plm_subtract_gradient": [_INFO, _P, _P, _P, _P, _P],
    "plm_autotune_sweep": [_INFO, _P, _P, _P, _I, _P, _P, _P],
    "plm_set_sweep_tile": [_P, _I],
    "plm_residual": [_INFO, _P, _P, _P, _P],
    "plm_restrict": [_P, _P, _P, _P, _P],
    "plm_prolong_add": [_P, _P, _P, _P, _P],
    "plm_zero": [_L, _P, _P],
    "plm_vcycle": [_I, _P, _P, _P, _P, _I, _I, _D, _I, _P, _P, _P],
    "plm_mg_batch_create": [_I, _P, _P, _P, _P, _P, _I, _D, _P, _P, _P, ctypes.POINTER(_P)],
    "plm_mg_batch_solve": [_P, _I, _P],
    "plm_mg_batch_destroy": [_P],
    "plm_mg_batch_autotune": [_P, _I, _P, _P],
    "plm_residual_norms": [_INFO, _P, _P, _P, _P, _P],
    "plm_vcycle_lp": [_I, _P, _I, _P, _P, _P, _I, _I, _I, _P],
    "plm_rb_sweeps_lp": [_P, _I, _P, _P, _P, _I, _I, _P],
    "plm_set_f16x2_shape": [_I],
    "plm_residual_restrict_lp": [_P, _I, _P, _P, _P, _P],
    "plm_prolong_add_lp": [_P, _I, _P, _P, _P],
    "plm_to_storage": [_L, _P, _F, _I, _P, _P],
    "plm_add_storage": [_L, _P, _P, _I, _F, _P],
    "plm_refine_residual": [_INFO, _P, _P, _F, _I, _P, _P, _P, _F, _P, _P, _P],
    "plm_adj_buoyancy": [_INFO, _P, _P, _F, _F, _F, _F, _F, _P, _P, _P, _P],
    "plm_adj_confinement": [_INFO, _P, _P, _P, _F, _F, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "plm_adj_advect_velocity": [_INFO] + [_P] * 15 + [_F, _P],
    "plm_adj_advect_scalars": [_INFO, _P, _P, _P, _P, _P, _F, _F] + [_P] * 11 + [_F, _P],
    "plm_adj_gradient_t": [_INFO, _P, _P, _P, _P, _P],
    "plm_adj_divergence_t": [_INFO, _P, _P, _P, _P, _P],
    "plm_adj_raymarch": [_INFO, _P, _P, ctypes.POINTER(ViewParams), ctypes.POINTER(LookParams), _P, _I, _I, _I, _P,
                         _P, _P, _P],
    "plm_line_sums": [_P, _P, _P, _P, _P, _P],
    "plm_smoke_budget": [_P, _P, _P, _P, _P, _P, _P],
    "plm_smoke_budget_noalloc": [_P, _P, _P, _P, _P, _P],
    "plm_scale": [_L, _P, _F, _P],
    "plm_max_abs3": [_INFO, _P, _P, _P, _P, _P],
    "plm_light_volume": [_P, _P, _P, _F, _F, _I, _P, _P],
    "plm_blur_axis": [_P, _P, _P, _P, _I, _I, _P],
    "plm_gradient": [_P, _P, _P, _P],
    "plm_raymarch": [_INFO, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _I, _I, _P],
    "plm_narrow_f16": [_L, _P, _P, _P],
    "plm_widen_f16": [_L, _P, _P, _P],
    "plm_narrow_center_velocity": [_INFO, _P, _P, _P, _P, _P],
    "plm_boundary_layers": [_P, _P, _P, _I, _I, _P],
    "plm_timing_enable": [ctypes.c_char_p],
    "plm_timing_disable": [],
    "plm_timing_read": [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64)],
}


def _load() -> ctypes.CDLL:
    if not LIBRARY.is_file():
        raise ImportError(f"{LIBRARY} is missing: build it with `python -m plume_amd.build` (hipcc, gfx950). "
                          "plume_amd has no other compute path.")
    lib = ctypes.CDLL(str(LIBRARY))
    lib.plm_last_error.restype = ctypes.c_char_p
    lib.plm_last_error.argtypes = []
    lib.plm_abi_version.restype = ctypes.c_int
    lib.plm_norm_scratch_doubles.restype = ctypes.c_int64
    lib.plm_norm_scratch_doubles.argtypes = [_INFO]
    lib.plm_adj_raymarch_record_stride.restype = ctypes.c_int32
    lib.plm_adj_raymarch_record_stride.argtypes = [_INFO, _I]
    lib.plm_mg_batch_depth.restype = ctypes.c_int32
    lib.plm_mg_batch_depth.argtypes = [_P]
    lib.plm_sweep_tile_shapes.restype = ctypes.c_int32
    lib.plm_sweep_tile_shapes.argtypes = [_P, _P, _P, _I]
    for name, argtypes in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = ctypes.c_int
        fn.argtypes = argtypes
    return lib


LIB = _load()
ABI_VERSION = 2

# plm_view.perspective (include/plume_hip.h)
VIEW_ORTHO, VIEW_PERSPECTIVE, VIEW_ORTHO_NEAR = 0, 1, 2
if LIB.plm_abi_version() != ABI_VERSION:
    raise ImportError(f"{LIBRARY} has ABI {LIB.plm_abi_version()}, expected {ABI_VERSION}: rebuild it")


def call(name: str, *args) -> None:
    """Calls a libplume_hip entry point; raises PlumeHipError on a non-zero return."""
    rc = getattr(LIB, name)(*args)
    if rc != 0:
        message = LIB.plm_last_error().decode(errors="replace")
        raise PlumeHipError(f"{name} failed ({rc}): {message}")


def stream() -> int:
    """torch's current HIP stream, on which every call is queued."""
    return torch.cuda.current_stream().cuda_stream


def ptr(t: torch.Tensor | None, dtype: torch.dtype = torch.float32):
    """The device pointer of a contiguous tensor of `dtype` on the GPU (None passes NULL)."""
    if t is None:
        return None
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"expected a torch tensor, got {type(t).__name__}")
    if t.device.type != "cuda":
        raise ValueError(f"plume_amd fields live on the GPU; got a tensor on {t.device}")
    if t.dtype != dtype:
        raise TypeError(f"expected {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError("plume_amd fields must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


def int3(values) -> ctypes.Array:
    return _int3(*(int(v) for v in values))


def float3(values) -> ctypes.Array:
    return _float3(*(float(v) for v in values))


def shape3(t: torch.Tensor) -> ctypes.Array:
    return _int3(*t.shape[:3])
