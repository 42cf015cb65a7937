"""CPU oracle for the Plume substep and ray march: ctypes bindings to `plume_oracle.c`.

TEST INFRASTRUCTURE ONLY: imported by `tests/`, `__graft_entry__.smoke()` and the CPU-baseline leg of
`bench.py`, and nowhere else. `Oracle("f32")` follows the reference's FP32 arithmetic; `Oracle("f64")` is
the same algorithm in double precision, the reference point for error statements.

Each method takes C-contiguous numpy arrays of the oracle's dtype and writes its outputs in place, like the
reference kernels it restates (sim/*.py and render/*.py of the reference).
"""

import ctypes
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIBS = {"f32": HERE / "liborc_f32.so", "f64": HERE / "liborc_f64.so"}

_int3 = ctypes.c_int32 * 3


def _structs(real):
    real3 = real * 3

    class Info(ctypes.Structure):
        _fields_ = [("dims", _int3), ("cell_m", real), ("origin_m", real3), ("open_lo", _int3),
                    ("open_hi", _int3), ("fill_hi", real3), ("ghost_gap", real)]

    class Emitter(ctypes.Structure):
        _fields_ = [("shape", ctypes.c_int32), ("center_m", real3), ("radius_m", real), ("half_size_m", real3),
                    ("edge_softness", real), ("start_s", real), ("end_s", real), ("smoke_per_s", real),
                    ("temperature_k", real), ("velocity_mps", real3), ("outward", ctypes.c_int32),
                    ("noise_scale_m", real), ("noise_strength", real), ("stream", ctypes.c_int32)]

    class View(ctypes.Structure):
        _fields_ = [("right", real3), ("up", real3), ("toward", real3), ("left_m", real), ("top_m", real),
                    ("pixel_m", real), ("perspective", ctypes.c_int32), ("eye", real3)]

    class Look(ctypes.Structure):
        _fields_ = [("opacity_per_density", real), ("glow_start_k", real), ("smoke_hides_heat", real),
                    ("normal_fade_thickness", real), ("shade_layout", ctypes.c_int32), ("shade_fill", real)]

    return Info, Emitter, View, Look


class Oracle:
    """The oracle's kernels in one precision."""

    def __init__(self, precision: str = "f32", library: Path | None = None):
        """`library`: another build of the same source in this precision (the CPU baseline's native build)."""
        if precision not in LIBS:
            raise ValueError(f"precision must be one of {tuple(LIBS)}")
        path = LIBS[precision] if library is None else Path(library)
        if not path.is_file():
            from oracle.build import build

            build()
        self.precision = precision
        self.dtype = np.float32 if precision == "f32" else np.float64
        self.real = ctypes.c_float if precision == "f32" else ctypes.c_double
        self.lib = ctypes.CDLL(str(path))
        self.Info, self.Emitter, self.View, self.Look = _structs(self.real)
        self.lib.orc_hash_u32.restype = ctypes.c_uint32
        self.lib.orc_hash_u32.argtypes = [ctypes.c_uint32]
        self.lib.orc_lattice_value.restype = self.real
        self.lib.orc_lattice_value.argtypes = [ctypes.c_int32] * 5
        self.lib.orc_value_noise.restype = self.real
        self.lib.orc_value_noise.argtypes = [ctypes.c_int32, ctypes.c_int32, self.real, self.real, self.real]

    # ------------------------------------------------------------------ helpers
    def _p(self, a: np.ndarray, dtype=None):
        dtype = self.dtype if dtype is None else dtype
        if not isinstance(a, np.ndarray) or a.dtype != dtype or not a.flags.c_contiguous:
            raise TypeError(f"oracle arrays must be C-contiguous {np.dtype(dtype).name}")
        return a.ctypes.data_as(ctypes.c_void_p)

    def _r(self, x):
        return self.real(float(x))

    @staticmethod
    def _shape(a):
        return _int3(*a.shape[:3])

    def _vec(self, values):
        return (self.real * 3)(*(float(x) for x in values))

    def info(self, dims, cell_m, origin_m=(0.0, 0.0, 0.0), open_lo=(0, 0, 0), open_hi=(0, 0, 0),
             fill_hi=(1.0, 1.0, 1.0), ghost_gap=0.5):
        """The oracle's GridInfo (sim/grid.py:21-33)."""
        return self.Info(_int3(*dims), float(cell_m), self._vec(origin_m), _int3(*open_lo), _int3(*open_hi),
                         self._vec(fill_hi), float(ghost_gap))

    def emitters(self, params):
        """An array of EmitterParams (sim/emitters.py:23-38) from dicts with the struct's field names."""
        array = (self.Emitter * max(1, len(params)))()
        for slot, p in zip(array, params):
            slot.shape = int(p["shape"])
            slot.center_m = self._vec(p["center_m"])
            slot.radius_m = float(p["radius_m"])
            slot.half_size_m = self._vec(p["half_size_m"])
            for key in ("edge_softness", "start_s", "end_s", "smoke_per_s", "temperature_k", "noise_scale_m",
                        "noise_strength"):
                setattr(slot, key, float(p[key]))
            slot.velocity_mps = self._vec(p["velocity_mps"])
            slot.outward = int(p["outward"])
            slot.stream = int(p["stream"])
        return array

    def view(self, right, up, toward, left_m, top_m, pixel_m, perspective=0, eye=(0.0, 0.0, 0.0)):
        return self.View(self._vec(right), self._vec(up), self._vec(toward), float(left_m), float(top_m),
                         float(pixel_m), int(perspective), self._vec(eye))

    def look(self, opacity_per_density, glow_start_k, smoke_hides_heat, normal_fade_thickness, shade_layout,
             shade_fill):
        return self.Look(float(opacity_per_density), float(glow_start_k), float(smoke_hides_heat),
                         float(normal_fade_thickness), int(shade_layout), float(shade_fill))

    # ------------------------------------------------------------------ noise (sim/noise.py)
    def set_threads(self, threads: int) -> None:
        """OpenMP threads of the following calls."""
        self.lib.orc_set_threads(int(threads))

    def hash_u32(self, x: int) -> int:
        return int(self.lib.orc_hash_u32(x & 0xFFFFFFFF))

    def lattice_value(self, seed, stream, i, j, k) -> float:
        return float(self.lib.orc_lattice_value(seed, stream, i, j, k))

    def value_noise(self, seed, stream, p) -> float:
        return float(self.lib.orc_value_noise(seed, stream, *(self._r(x) for x in p)))

    # ------------------------------------------------------------------ advection (sim/advect.py)
    def advect_sl(self, info, src, dst, offset, outside, uses_outside, u, v, w, dt_over_cell):
        self.lib.orc_advect_sl(ctypes.byref(info), self._p(src), self._p(dst), self._shape(src), self._vec(offset),
                               self._r(outside), int(uses_outside), self._p(u), self._p(v), self._p(w),
                               self._r(dt_over_cell))

    def maccormack_correct(self, info, original, hat, back, dst, offset, outside, uses_outside, u, v, w,
                           dt_over_cell):
        self.lib.orc_maccormack_correct(ctypes.byref(info), self._p(original), self._p(hat), self._p(back),
                                        self._p(dst), self._shape(original), self._vec(offset), self._r(outside),
                                        int(uses_outside), self._p(u), self._p(v), self._p(w),
                                        self._r(dt_over_cell))

    def advect_maccormack(self, info, data, hat, back, nxt, offset, outside, uses_outside, u, v, w, dt):
        """`dt` in seconds; dt / cell_m is formed in double and rounded once (sim/advect.py:62)."""
        self.lib.orc_advect_maccormack(ctypes.byref(info), self._p(data), self._p(hat), self._p(back),
                                       self._p(nxt), self._shape(data), self._vec(offset), self._r(outside),
                                       int(uses_outside), self._p(u), self._p(v), self._p(w),
                                       self._r(float(dt) / info.cell_m))

    # ------------------------------------------------------------------ forces (sim/forces.py, vorticity.py)
    def buoyancy(self, info, v, smoke, temperature, gravity, heat_buoyancy, smoke_weight, ambient_k, dt):
        self.lib.orc_buoyancy(ctypes.byref(info), self._p(v), self._p(smoke), self._p(temperature),
                              self._r(gravity), self._r(heat_buoyancy), self._r(smoke_weight), self._r(ambient_k),
                              self._r(dt))

    def drag(self, f, wind, keep):
        self.lib.orc_drag(self._p(f), ctypes.c_int64(f.size), self._r(wind), self._r(keep))

    def center_velocity(self, info, u, v, w, center):
        self.lib.orc_center_velocity(ctypes.byref(info), self._p(u), self._p(v), self._p(w), self._p(center))

    def curl(self, info, center, curl):
        self.lib.orc_curl(ctypes.byref(info), self._p(center), self._p(curl))

    def confinement_push(self, info, curl, push, strength):
        self.lib.orc_confinement_push(ctypes.byref(info), self._p(curl), self._p(push), self._r(strength))

    def add_face_push(self, info, f, push, axis, dt):
        self.lib.orc_add_face_push(ctypes.byref(info), self._p(f), self._p(push), int(axis), self._r(dt))

    def emit_scalars(self, info, smoke, temperature, emitters, n, seed, time_s, dt):
        self.lib.orc_emit_scalars(ctypes.byref(info), self._p(smoke), self._p(temperature), emitters, int(n),
                                  int(seed), self._r(time_s), self._r(dt))

    def emit_velocity(self, info, component, axis, emitters, n, seed, time_s, dt):
        self.lib.orc_emit_velocity(ctypes.byref(info), self._p(component), int(axis), emitters, int(n), int(seed),
                                   self._r(time_s), self._r(dt))

    def turbulence(self, info, f, axis, seed, scale_m, strength, pattern_time, dt):
        self.lib.orc_turbulence(ctypes.byref(info), self._p(f), int(axis), int(seed), self._r(scale_m),
                                self._r(strength), self._r(pattern_time), self._r(dt))

    # ------------------------------------------------------------------ pressure (sim/pressure.py, multigrid.py)
    def zero_closed_faces(self, info, f, axis):
        self.lib.orc_zero_closed_faces(ctypes.byref(info), self._p(f), int(axis))

    def divergence(self, info, u, v, w, div):
        self.lib.orc_divergence(ctypes.byref(info), self._p(u), self._p(v), self._p(w), self._p(div))

    def build_flags(self, info, flags):
        self.lib.orc_build_flags(ctypes.byref(info), self._p(flags, np.uint16))

    def rb_sweep(self, info, q, rhs, color, relaxation):
        self.lib.orc_rb_sweep(ctypes.byref(info), self._p(q), self._p(rhs), int(color), self._r(relaxation))

    def rb_sweeps(self, info, q, rhs, sweeps, relaxation):
        self.lib.orc_rb_sweeps(ctypes.byref(info), self._p(q), self._p(rhs), int(sweeps), self._r(relaxation))

    def subtract_gradient(self, info, f, q, axis):
        self.lib.orc_subtract_gradient(ctypes.byref(info), self._p(f), self._p(q), int(axis))

    def residual(self, info, q, rhs, residual):
        self.lib.orc_residual(ctypes.byref(info), self._p(q), self._p(rhs), self._p(residual))

    def restrict(self, fine, coarse):
        self.lib.orc_restrict(self._p(fine), self._shape(fine), self._p(coarse), self._shape(coarse))

    def prolong_add(self, coarse, fine):
        self.lib.orc_prolong_add(self._p(coarse), self._shape(coarse), self._p(fine), self._shape(fine))

    def multigrid(self, infos, q, rhs, residual, cycles, sweeps):
        n = len(infos)
        info_array = (self.Info * n)(*infos)
        ptrs = ctypes.c_void_p * n
        self.lib.orc_multigrid(n, info_array, ptrs(*(self._p(a) for a in q)), ptrs(*(self._p(a) for a in rhs)),
                               ptrs(*(self._p(a) for a in residual)), int(cycles), int(sweeps))

    # ------------------------------------------------------------------ budget (sim/reduce.py, conserve.py)
    def line_sums(self, a, b, c, lines):
        self.lib.orc_line_sums(self._shape(a), self._p(a), self._p(b), self._p(c), self._p(lines))

    def scale(self, f, factor):
        self.lib.orc_scale(self._p(f), ctypes.c_int64(f.size), self._r(factor))

    def dissipate(self, smoke, temperature, smoke_keep, heat_keep, ambient_k):
        self.lib.orc_dissipate(self._p(smoke), self._p(temperature), ctypes.c_int64(smoke.size),
                               self._r(smoke_keep), self._r(heat_keep), self._r(ambient_k))

    # ------------------------------------------------------------------ rendering (render/*.py)
    def light_volume(self, smoke, toward_light, cell_m, opacity_per_density, light, max_steps=0):
        self.lib.orc_light_volume(self._shape(smoke), self._p(smoke), self._vec(toward_light), self._r(cell_m),
                                  self._r(opacity_per_density), int(max_steps), self._p(light))

    def blur_axis(self, src, dst, weights, radius, axis):
        self.lib.orc_blur_axis(self._shape(src), self._p(src), self._p(dst), self._p(weights), int(radius),
                               int(axis))

    def gradient(self, src, gradient):
        self.lib.orc_gradient(self._shape(src), self._p(src), self._p(gradient))

    def raymarch(self, info, smoke, temperature, gradient, light, view, look, image, max_steps=0):
        height, width = image.shape[:2]
        self.lib.orc_raymarch(ctypes.byref(info), self._p(smoke), self._p(temperature), self._p(gradient),
                              self._shape(gradient), self._p(light), self._shape(light), ctypes.byref(view),
                              ctypes.byref(look), self._p(image), int(height), int(width), int(max_steps))
