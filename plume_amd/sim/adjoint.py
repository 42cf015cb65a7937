"""Differentiable rollout: the substep and the ray march's heat and thickness with hand-written adjoints.

The reference runs no backward pass (sim/device.py:28 sets enable_backward = False); this addition makes a
rollout differentiable with respect to per-cell emitter sources and the buoyancy coefficient alpha, so that
emitter and force parameters can be fitted to a target flipbook.

The differentiable substep is the plain substep's operators in the reference's order (sim/stepper.py:28-44),
with the emitter given as per-cell source fields:

1. emission: smoke += dt S_s, temperature += dt S_T, v += dt S_v (sources on cells, cells, v faces);
2. buoyancy (sim/forces.py:13-31) with smoke_weight = alpha / g;
3. vorticity confinement (sim/vorticity.py:95-106);
4. clamped MacCormack of u, v, w by the old velocity, then swap (sim/stepper.py:46-52);
5. projection (sim/pressure.py:144-154) with multigrid to a relative residual;
6. clamped MacCormack of smoke and temperature by the projected velocity, then swap (sim/stepper.py:54-60).

Left out on purpose: the smoke budget's global rescale (a hold of totals, not differentiable where it binds)
and the CFL cap (the step length is fixed). Every forward operation is the library's ordinary kernel; the
backward runs the adjoint kernels of libplume_hip in reverse (plm_adj_*), the projection's transpose with one
more multigrid solve of the same symmetric operator. The rollout renders the heat and thickness images at the
end of every `checkpoint` substeps (the flipbook) and stores the state only there; the backward pass re-runs
each segment forward once, keeping the substeps' intermediates, and walks it back.
"""

import ctypes
from dataclasses import dataclass
from types import SimpleNamespace

import numpy as np
import torch

from plume_amd._lib import LIB, VIEW_ORTHO_NEAR, GridInfo, LookParams, ViewParams, call, float3, ptr, stream
from plume_amd.render.camera import make_camera
from plume_amd.render.frame_renderer import look_params
from plume_amd.render.raymarch import march
from plume_amd.sim.advect import advect_scalars, advect_velocity
from plume_amd.sim.forces import apply_buoyancy
from plume_amd.sim.grid import make_grid
from plume_amd.sim.multigrid import Multigrid
from plume_amd.sim.pressure import project
from plume_amd.sim.vorticity import VorticityConfinement

FIELDS = ("u", "v", "w", "smoke", "temperature")


@dataclass
class Sources:
    """Per-cell emitter sources: smoke and temperature per second on cells, upward acceleration on v faces."""

    smoke: torch.Tensor
    temperature: torch.Tensor
    v: torch.Tensor

    def tensors(self):
        return self.smoke, self.temperature, self.v


def sphere_sources(grid, center_cells, radius_cells: float, smoke_per_s: float = 2.0, heat_per_s: float = 4.0,
                   lift_mps2: float = 1.0, softness: float = 0.6, falloff: str = "smoothstep") -> Sources:
    """Sources on a sphere: strength times a coverage that is 0 beyond the radius. falloff "smoothstep" is the
    reference's soft edge (sim/emitters.py:41-66), 1 inside and falling to 0 over the outer `softness` of the
    radius; "gaussian" is exp(-d^2 / (2 (radius / 2)^2)), which has no flat top."""
    if falloff not in ("smoothstep", "gaussian"):
        raise ValueError(f"unknown falloff {falloff!r}")
    nx, ny, nz = grid.cells

    def coverage(shape, offset):
        coords = [torch.arange(shape[a], dtype=torch.float32, device=grid.device) + offset[a] for a in range(3)]
        X, Y, Z = torch.meshgrid(*coords, indexing="ij")
        d = torch.sqrt((X - center_cells[0]) ** 2 + (Y - center_cells[1]) ** 2 + (Z - center_cells[2]) ** 2)
        if falloff == "gaussian":
            sigma = 0.5 * radius_cells
            return torch.where(d < radius_cells, torch.exp(-d * d / (2.0 * sigma * sigma)), torch.zeros_like(d))
        x = ((radius_cells - d) / (softness * radius_cells)).clamp(0.0, 1.0)
        return x * x * (3.0 - 2.0 * x)

    cells = coverage((nx, ny, nz), (0.5, 0.5, 0.5))
    faces = coverage((nx, ny + 1, nz), (0.5, 0.0, 0.5))
    return Sources(smoke_per_s * cells, heat_per_s * cells, lift_mps2 * faces)


def _square_view(grid, camera, pixels: int) -> ViewParams:
    """An orthographic view of `pixels` x `pixels` over the domain's projection (render/frame_renderer.py:72-83
    with a square pixel grid), its rays started on the camera plane (VIEW_ORTHO_NEAR): the reference's start
    1e4 m out rounds FP32 sample positions by ~5e-4 m, a twentieth of a centimetre cell, which the gradient
    would inherit."""
    cell = grid.info.cell_m
    origin = np.array(grid.info.origin_m, dtype=np.float64)
    size = np.array(grid.cells, dtype=np.float64) * cell
    corners = np.array([[origin[0] + size[0] * a, origin[1] + size[1] * b, origin[2] + size[2] * c]
                        for a in (0, 1) for b in (0, 1) for c in (0, 1)])
    xs, ys = corners @ np.asarray(camera.right), corners @ np.asarray(camera.up)
    span = max(xs.max() - xs.min(), ys.max() - ys.min())
    pixel_m = span / pixels
    view = ViewParams()
    view.right, view.up, view.toward = float3(camera.right), float3(camera.up), float3(camera.toward)
    view.left_m = float((xs.min() + xs.max()) / 2 - span / 2)
    view.top_m = float((ys.min() + ys.max()) / 2 + span / 2)
    view.pixel_m = float(pixel_m)
    view.perspective = VIEW_ORTHO_NEAR
    view.eye = float3((0.0, 0.0, 0.0))
    return view


class DiffSim:
    """The scene of a differentiable rollout: one grid, its solver and scratch, the camera and the adjoint
    buffers."""

    def __init__(self, recipe, derived, device: str, pixels: int = 128, march_steps: int = 128,
                 tolerance: float = 1e-6, max_cycles: int = 40):
        self.recipe = recipe
        self.grid = make_grid(derived, recipe.domain.cell_m, recipe.domain.faces, recipe.domain.ambient_k, device)
        g = self.grid
        self.dt = float(derived.dt_s)
        self.dto = self.dt / g.info.cell_m
        forces = recipe.forces
        self.gravity = float(forces.gravity_mps2)
        self.heat_buoyancy = float(forces.heat_buoyancy)
        self.ambient = float(recipe.domain.ambient_k)
        self.strength = float(forces.vorticity_confinement)
        self.solve = Multigrid(g, 2, recipe.solver.multigrid_sweeps, tolerance=tolerance, max_cycles=max_cycles)
        self.confinement = VorticityConfinement(g) if self.strength > 0 else None
        camera = make_camera(recipe.camera.elevation_deg, recipe.camera.azimuth_deg)
        self.view = _square_view(g, camera, pixels)
        self.look: LookParams = look_params(recipe.render, shade_layout=True)
        self.pixels = pixels
        self.march_steps = int(march_steps)  # samples per ray, fixed (every ray the same count)
        self.image = torch.zeros((pixels, pixels, 5), dtype=torch.float32, device=device)
        self.light = torch.ones((1, 1, 1), dtype=torch.float32, device=device)
        self.gradient = torch.zeros((1, 1, 1, 3), dtype=torch.float32, device=device)
        self._adjoint = None

    # ------------------------------------------------------------------ state

    def reset(self, initial: dict | None = None) -> None:
        """Rest (zero velocity and smoke, ambient temperature), or the fields of `initial`."""
        g = self.grid
        if initial is not None:
            self.load({**initial, "pressure": initial.get("pressure", torch.zeros_like(g.pressure))})
            return
        for name in ("u", "v", "w", "smoke"):
            getattr(g, name).data.zero_()
        g.temperature.data.fill_(self.ambient)
        g.pressure.zero_()

    def state(self) -> dict:
        g = self.grid
        out = {name: getattr(g, name).data.clone() for name in FIELDS}
        out["pressure"] = g.pressure.clone()
        return out

    def load(self, state: dict) -> None:
        g = self.grid
        for name in FIELDS:
            getattr(g, name).data.copy_(state[name])
        g.pressure.copy_(state["pressure"])

    # ------------------------------------------------------------------ forward

    def substep(self, sources: Sources, alpha: float, keep: bool = False):
        """One differentiable substep; with `keep`, returns what its adjoint needs."""
        g, dt = self.grid, self.dt
        g.smoke.data.add_(sources.smoke * dt)
        g.temperature.data.add_(sources.temperature * dt)
        g.v.data.add_(sources.v * dt)
        kept = {}
        if keep:
            kept["smoke"] = g.smoke.data.clone()
            kept["temperature"] = g.temperature.data.clone()
        apply_buoyancy(g, self.gravity, self.heat_buoyancy, alpha / self.gravity, self.ambient, dt)
        if keep:
            kept["buoyant"] = [f.data.clone() for f in g.velocity_fields()]
        if self.confinement is not None:
            self.confinement.apply(g, self.strength, dt)
        if keep:
            kept["confined"] = [f.data.clone() for f in g.velocity_fields()]
        advect_velocity(g, dt)
        for f in g.velocity_fields():
            f.swap()
        project(g, self.solve)
        if keep:
            kept["projected"] = [f.data.clone() for f in g.velocity_fields()]
        advect_scalars(g, tuple(f.data for f in g.velocity_fields()), dt)
        g.smoke.swap()
        g.temperature.swap()
        return kept if keep else None

    def render(self) -> torch.Tensor:
        """(heat, thickness) of the current state, (pixels, pixels, 2)."""
        g = self.grid
        march(g.smoke.data, g.temperature.data, self.gradient, self.light, g.info, self.view, self.look, self.image,
              self.march_steps)
        return self.image[..., :2].clone()

    # ------------------------------------------------------------------ adjoint

    def _buffers(self):
        if self._adjoint is None:
            g = self.grid
            dev = g.device
            faces = [tuple(f.data.shape) for f in g.velocity_fields()]
            cells = tuple(g.cells)
            z = lambda shape: torch.zeros(shape, dtype=torch.float32, device=dev)  # noqa: E731
            stride = LIB.plm_adj_raymarch_record_stride(ctypes.byref(g.info), self.march_steps)
            holder = SimpleNamespace(pressure=z(cells), divergence=z(cells), cells=cells, info=g.info, device=dev)
            self._adjoint = SimpleNamespace(
                hat=[z(s) for s in faces], hat_bar=[z(s) for s in faces], in_bar=[z(s) for s in faces],
                scalar_hat=[z(cells), z(cells)], scalar_hat_bar=[z(cells), z(cells)],
                curl=z(cells + (3,)), push_bar=z(cells + (3,)), curl_bar=z(cells + (3,)), spin_bar=z(cells),
                center_bar=z(cells + (3,)), weight_terms=z(faces[1]), holder=holder,
                solve=Multigrid(g, 2, self.solve.sweeps, tolerance=self.solve.tolerance,
                                max_cycles=self.solve.max_cycles),
                record=z((self.pixels * self.pixels * stride * 4,)), image_bar=z((self.pixels, self.pixels, 5)))
        return self._adjoint

    def render_adjoint(self, heat_thickness_bar: torch.Tensor, bar: dict) -> None:
        """bar['smoke'], bar['temperature'] += the adjoint of render() given the images' adjoint."""
        a, g = self._buffers(), self.grid
        a.image_bar.zero_()
        a.image_bar[..., :2].copy_(heat_thickness_bar)
        call("plm_adj_raymarch", ctypes.byref(g.info), ptr(g.smoke.data), ptr(g.temperature.data),
             ctypes.byref(self.view), ctypes.byref(self.look), ptr(a.image_bar), self.pixels, self.pixels,
             self.march_steps, ptr(a.record), ptr(bar["smoke"]), ptr(bar["temperature"]), stream())

    def substep_adjoint(self, bar: dict, kept: dict, sources_bar: Sources, alpha: float) -> float:
        """Replaces bar (the adjoint of the substep's output state) by the adjoint of its input state; adds the
        sources' adjoint into sources_bar and returns alpha's."""
        a, g, dt = self._buffers(), self.grid, self.dt
        info = ctypes.byref(g.info)
        U5 = kept["projected"]
        ub = [bar["u"], bar["v"], bar["w"]]
        s_bar = torch.zeros_like(bar["smoke"])
        t_bar = torch.zeros_like(bar["temperature"])
        # 6. scalar advection: the scalars' adjoint, and the projected velocity's part added into ub
        call("plm_adj_advect_scalars", info, *(ptr(x) for x in U5), ptr(kept["smoke"]), ptr(kept["temperature"]),
             0.0, self.ambient, ptr(a.scalar_hat[0]), ptr(a.scalar_hat[1]), ptr(bar["smoke"]), ptr(bar["temperature"]),
             ptr(s_bar), ptr(t_bar), *(ptr(x) for x in ub), ptr(a.scalar_hat_bar[0]), ptr(a.scalar_hat_bar[1]),
             self.dto, stream())
        # 5. projection: ub = Z (ub - D^T y), y the Poisson solve of G^T ub
        holder = a.holder
        call("plm_adj_gradient_t", info, *(ptr(x) for x in ub), ptr(holder.divergence), stream())
        holder.pressure.zero_()
        a.solve(holder)
        call("plm_adj_divergence_t", info, ptr(holder.pressure), *(ptr(x) for x in ub), stream())
        # 4. velocity advection from the confined velocity
        for x in a.in_bar:
            x.zero_()
        U3 = kept["confined"]
        call("plm_adj_advect_velocity", info, *(ptr(x) for x in U3), *(ptr(x) for x in a.hat), *(ptr(x) for x in ub),
             *(ptr(x) for x in a.in_bar), *(ptr(x) for x in a.hat_bar), self.dto, stream())
        for dst, src in zip(ub, a.in_bar):
            dst.copy_(src)
        # 3. confinement at the buoyant velocity
        U2 = kept["buoyant"]
        if self.confinement is not None:
            call("plm_adj_confinement", info, *(ptr(x) for x in U2), self.strength, dt, *(ptr(x) for x in ub),
                 ptr(a.curl), ptr(a.push_bar), ptr(a.curl_bar), ptr(a.spin_bar), ptr(a.center_bar), stream())
        # 2. buoyancy: v passes through; the scalars and alpha receive their parts
        call("plm_adj_buoyancy", info, ptr(ub[1]), ptr(kept["smoke"]), self.gravity, self.heat_buoyancy,
             alpha / self.gravity, self.ambient, dt, ptr(s_bar), ptr(t_bar), ptr(a.weight_terms), stream())
        alpha_bar = float(a.weight_terms.double().sum().item()) / self.gravity
        # 1. emission
        sources_bar.smoke.add_(s_bar * dt)
        sources_bar.temperature.add_(t_bar * dt)
        sources_bar.v.add_(ub[1] * dt)
        bar["smoke"].copy_(s_bar)
        bar["temperature"].copy_(t_bar)
        return alpha_bar


class Rollout(torch.autograd.Function):
    """images = rollout(S_s, S_T, S_v, alpha): `substeps` differentiable substeps from rest (or from the
    fields of `initial`), the (heat, thickness) image after every `checkpoint` of them, shape
    (substeps / checkpoint, pixels, pixels, 2). Only the state at each checkpoint is stored; backward re-runs
    one segment at a time."""

    @staticmethod
    def forward(ctx, src_smoke, src_temperature, src_v, alpha, sim: DiffSim, substeps: int, checkpoint: int,
                initial: dict | None):
        if substeps % checkpoint != 0:
            raise ValueError("substeps must be a multiple of checkpoint")
        sources = Sources(src_smoke.detach(), src_temperature.detach(), src_v.detach())
        a = float(alpha)
        sim.reset(initial)
        states, images = [], []
        for _ in range(substeps // checkpoint):
            states.append(sim.state())
            for _ in range(checkpoint):
                sim.substep(sources, a)
            images.append(sim.render())
        ctx.sim, ctx.states, ctx.sources, ctx.alpha, ctx.checkpoint = sim, states, sources, a, checkpoint
        return torch.stack(images)

    @staticmethod
    def backward(ctx, images_bar):
        sim, sources, a, k = ctx.sim, ctx.sources, ctx.alpha, ctx.checkpoint
        bar = {name: torch.zeros_like(getattr(sim.grid, name).data) for name in FIELDS}
        sources_bar = Sources(*(torch.zeros_like(t) for t in sources.tensors()))
        alpha_bar = 0.0
        for segment in reversed(range(len(ctx.states))):
            sim.load(ctx.states[segment])
            kept = [sim.substep(sources, a, keep=True) for _ in range(k)]
            sim.render_adjoint(images_bar[segment].contiguous(), bar)
            for saved in reversed(kept):
                alpha_bar += sim.substep_adjoint(bar, saved, sources_bar, a)
        alpha_grad = torch.tensor(alpha_bar, dtype=torch.float64)
        return sources_bar.smoke, sources_bar.temperature, sources_bar.v, alpha_grad, None, None, None, None


def rollout(sources: Sources, alpha: torch.Tensor, sim: DiffSim, substeps: int, checkpoint: int,
            initial: dict | None = None) -> torch.Tensor:
    return Rollout.apply(sources.smoke, sources.temperature, sources.v, alpha, sim, substeps, checkpoint, initial)
