"""The hand-written adjoints against float64 finite differences of the oracle (the C restatement in double).

For an operator F and a random output cotangent w, input point x and direction d, the GPU adjoint's
<F^T w, d> is compared with the oracle's central difference (<w, F(x + h d)> - <w, F(x - h d)>) / 2h in FP64.
Then the whole differentiable rollout (sources -> 8 substeps -> heat and thickness images, checkpointed) is
checked the same way along random directions of the emitter sources and of alpha."""

import ctypes

import numpy as np
import pytest

from support import DEFAULT_FACES, gpu_grid, oracle_grid

pytestmark = pytest.mark.gpu

CELLS = (10, 14, 9)
CELL_M = 0.1
DT = 0.04


def to_np(t):
    return t.detach().cpu().numpy()


def smooth_field(rng, shape, scale=1.0, offset=0.0):
    idx = np.indices(shape).astype(np.float64)
    f = np.zeros(shape)
    for _ in range(3):
        k = rng.uniform(0.2, 0.7, size=3)
        phase = rng.uniform(0, 2 * np.pi)
        f += np.sin(sum(k[a] * idx[a] for a in range(3)) + phase)
    return offset + scale * (f / 3.0 + 0.05 * rng.standard_normal(shape))


def velocity(rng):
    nx, ny, nz = CELLS
    return [smooth_field(rng, (nx + 1, ny, nz), 0.6), smooth_field(rng, (nx, ny + 1, nz), 0.6),
            smooth_field(rng, (nx, ny, nz + 1), 0.6)]


def set_gpu(gpu, fields):
    import torch

    for name, values in fields.items():
        target = gpu.pressure if name == "pressure" else getattr(gpu, name).data
        target.copy_(torch.from_numpy(np.asarray(values, dtype=np.float32)))


def set_orc(orc, fields):
    for name, values in fields.items():
        target = orc.pressure if name == "pressure" else getattr(orc, name).data
        target[...] = np.asarray(values, dtype=np.float32).astype(np.float64)


def fd_check(adjoint_dot, evaluate, fields, direction, h=1e-3, rel=1e-3):
    """evaluate(fields) -> <w, F(fields)> in FP64; direction: same keys as a subset of fields."""
    plus = {k: v + h * direction[k] if k in direction else v for k, v in fields.items()}
    minus = {k: v - h * direction[k] if k in direction else v for k, v in fields.items()}
    fd = (evaluate(plus) - evaluate(minus)) / (2 * h)
    assert abs(adjoint_dot - fd) <= rel * max(abs(fd), 1e-12), (adjoint_dot, fd)
    return fd


def test_buoyancy_adjoint(device, oracle64):
    import torch

    from plume_amd._lib import call, ptr, stream

    rng = np.random.default_rng(1)
    gpu = gpu_grid(CELLS, CELL_M, DEFAULT_FACES, device=device)
    orc = oracle_grid(oracle64, CELLS, CELL_M, DEFAULT_FACES)
    s, t = smooth_field(rng, CELLS, 0.5, 0.6), smooth_field(rng, CELLS, 0.4, 1.5)
    w = smooth_field(rng, tuple(gpu.v.data.shape))
    g, hb, amb, alpha = 9.81, 1.0 / 9.81, 1.0, 0.1
    set_gpu(gpu, {"smoke": s, "temperature": t})
    sb = torch.zeros_like(gpu.smoke.data)
    tb = torch.zeros_like(gpu.temperature.data)
    terms = torch.zeros_like(gpu.v.data)
    wt = torch.from_numpy(w.astype(np.float32)).to(device)
    call("plm_adj_buoyancy", ctypes.byref(gpu.info), ptr(wt), ptr(gpu.smoke.data), g, hb, alpha / g, amb, DT, ptr(sb),
         ptr(tb), ptr(terms), stream())
    ds, dtemp, dalpha = rng.standard_normal(CELLS), rng.standard_normal(CELLS), 0.7
    dot = float((to_np(sb) * ds).sum() + (to_np(tb) * dtemp).sum() + to_np(terms).astype(np.float64).sum() / g * dalpha)

    def evaluate(f):
        orc.v.data[...] = 0.0
        orc.smoke.data[...] = f["smoke"]
        orc.temperature.data[...] = f["temperature"]
        oracle64.buoyancy(orc.info, orc.v.data, orc.smoke.data, orc.temperature.data, g, hb, f["alpha"] / g, amb, DT)
        return float((orc.v.data * w).sum())

    fd_check(dot, evaluate, {"smoke": s, "temperature": t, "alpha": alpha},
             {"smoke": ds, "temperature": dtemp, "alpha": dalpha}, rel=1e-4)


def test_confinement_adjoint(device, oracle64):
    import torch

    from oracle.sim import VorticityConfinement
    from plume_amd._lib import call, ptr, stream

    rng = np.random.default_rng(2)
    gpu = gpu_grid(CELLS, CELL_M, DEFAULT_FACES, device=device)
    orc = oracle_grid(oracle64, CELLS, CELL_M, DEFAULT_FACES)
    U = velocity(rng)
    W = [rng.standard_normal(u.shape) for u in U]
    D = [rng.standard_normal(u.shape) for u in U]
    set_gpu(gpu, dict(zip("uvw", U)))
    bars = [torch.from_numpy(x.astype(np.float32)).to(device) for x in W]
    scratch = [torch.zeros(s, dtype=torch.float32, device=device)  # held: each is its own buffer
               for s in (CELLS + (3,), CELLS + (3,), CELLS + (3,), CELLS, CELLS + (3,))]
    call("plm_adj_confinement", ctypes.byref(gpu.info), *(ptr(f.data) for f in gpu.velocity_fields()), 0.3, DT,
         *(ptr(b) for b in bars), *(ptr(x) for x in scratch), stream())
    dot = sum(float((to_np(b) * d).sum()) for b, d in zip(bars, D))
    conf = VorticityConfinement(orc)

    def evaluate(f):
        for name, field in zip("uvw", orc.velocity_fields()):
            field.data[...] = f[name]
        conf.apply(orc, 0.3, DT)
        return sum(float((field.data * w).sum()) for field, w in zip(orc.velocity_fields(), W))

    fd_check(dot, evaluate, dict(zip("uvw", U)), dict(zip("uvw", D)), rel=2e-3)


def test_velocity_advection_adjoint(device, oracle64):
    import torch

    from oracle.sim import advect_maccormack
    from plume_amd._lib import call, ptr, stream

    rng = np.random.default_rng(3)
    gpu = gpu_grid(CELLS, CELL_M, DEFAULT_FACES, device=device)
    orc = oracle_grid(oracle64, CELLS, CELL_M, DEFAULT_FACES)
    U = velocity(rng)
    W = [rng.standard_normal(u.shape) for u in U]
    D = [rng.standard_normal(u.shape) for u in U]
    set_gpu(gpu, dict(zip("uvw", U)))
    nb = [torch.from_numpy(x.astype(np.float32)).to(device) for x in W]
    ib = [torch.zeros_like(b) for b in nb]
    scratch = [torch.zeros_like(b) for b in nb] + [torch.zeros_like(b) for b in nb]
    call("plm_adj_advect_velocity", ctypes.byref(gpu.info), *(ptr(f.data) for f in gpu.velocity_fields()),
         *(ptr(x) for x in scratch[:3]), *(ptr(x) for x in nb), *(ptr(x) for x in ib), *(ptr(x) for x in scratch[3:]),
         DT / CELL_M, stream())
    dot = sum(float((to_np(b) * d).sum()) for b, d in zip(ib, D))

    def evaluate(f):
        fields = orc.velocity_fields()
        for name, field in zip("uvw", fields):
            field.data[...] = f[name]
        old = tuple(field.data.copy() for field in fields)
        total = 0.0
        for field, w in zip(fields, W):
            advect_maccormack(field, old, orc, DT)
            total += float((field.next * w).sum())
        return total

    # MacCormack's clamp makes the map piecewise smooth: the difference needs a step that crosses no kink
    fd_check(dot, evaluate, dict(zip("uvw", U)), dict(zip("uvw", D)), h=1e-6, rel=1e-3)


def test_scalar_advection_adjoint(device, oracle64):
    import torch

    from oracle.sim import advect_maccormack
    from plume_amd._lib import call, ptr, stream

    rng = np.random.default_rng(4)
    gpu = gpu_grid(CELLS, CELL_M, DEFAULT_FACES, ambient=1.0, device=device)
    orc = oracle_grid(oracle64, CELLS, CELL_M, DEFAULT_FACES, ambient=1.0)
    U = velocity(rng)
    s, t = smooth_field(rng, CELLS, 0.5, 0.6), smooth_field(rng, CELLS, 0.4, 1.5)
    Ws, Wt = rng.standard_normal(CELLS), rng.standard_normal(CELLS)
    set_gpu(gpu, dict(zip("uvw", U)) | {"smoke": s, "temperature": t})
    z = lambda: torch.zeros(CELLS, dtype=torch.float32, device=device)  # noqa: E731
    sb, tb, h0, h1, hb0, hb1 = z(), z(), z(), z(), z(), z()
    ub = [torch.zeros_like(f.data) for f in gpu.velocity_fields()]
    ws = torch.from_numpy(Ws.astype(np.float32)).to(device)
    wt = torch.from_numpy(Wt.astype(np.float32)).to(device)
    call("plm_adj_advect_scalars", ctypes.byref(gpu.info), *(ptr(f.data) for f in gpu.velocity_fields()),
         ptr(gpu.smoke.data), ptr(gpu.temperature.data), 0.0, 1.0, ptr(h0), ptr(h1), ptr(ws), ptr(wt), ptr(sb),
         ptr(tb), *(ptr(b) for b in ub), ptr(hb0), ptr(hb1), DT / CELL_M, stream())
    Ds, Dt = rng.standard_normal(CELLS), rng.standard_normal(CELLS)
    DU = [rng.standard_normal(u.shape) for u in U]
    dot = float((to_np(sb) * Ds).sum() + (to_np(tb) * Dt).sum()) + sum(float((to_np(b) * d).sum())
                                                                      for b, d in zip(ub, DU))

    def evaluate(f):
        for name, field in zip("uvw", orc.velocity_fields()):
            field.data[...] = f[name]
        orc.smoke.data[...] = f["smoke"]
        orc.temperature.data[...] = f["temperature"]
        vel = tuple(field.data for field in orc.velocity_fields())
        advect_maccormack(orc.smoke, vel, orc, DT)
        advect_maccormack(orc.temperature, vel, orc, DT)
        return float((orc.smoke.next * Ws).sum() + (orc.temperature.next * Wt).sum())

    fields = dict(zip("uvw", U)) | {"smoke": s, "temperature": t}
    fd_check(dot, evaluate, fields, dict(zip("uvw", DU)) | {"smoke": Ds, "temperature": Dt}, h=1e-6, rel=1e-3)


def test_projection_adjoint(device, oracle64):
    import torch

    from oracle.sim import Multigrid as OracleMultigrid, poisson_residual, project
    from plume_amd._lib import call, ptr, stream
    from plume_amd.sim.multigrid import Multigrid

    rng = np.random.default_rng(5)
    gpu = gpu_grid(CELLS, CELL_M, DEFAULT_FACES, device=device)
    orc = oracle_grid(oracle64, CELLS, CELL_M, DEFAULT_FACES)
    U = velocity(rng)
    W = [rng.standard_normal(u.shape) for u in U]
    D = [rng.standard_normal(u.shape) for u in U]
    ub = [torch.from_numpy(x.astype(np.float32)).to(device) for x in W]
    call("plm_adj_gradient_t", ctypes.byref(gpu.info), *(ptr(b) for b in ub), ptr(gpu.divergence), stream())
    gpu.pressure.zero_()
    Multigrid(gpu, 2, 2, tolerance=1e-7, max_cycles=60)(gpu)
    call("plm_adj_divergence_t", ctypes.byref(gpu.info), ptr(gpu.pressure), *(ptr(b) for b in ub), stream())
    dot = sum(float((to_np(b) * d).sum()) for b, d in zip(ub, D))

    def solve(grid):
        s = OracleMultigrid(grid, 1, 2)
        for _ in range(60):
            s(grid)
            if poisson_residual(grid) < 1e-13:
                break

    def evaluate(f):
        for name, field in zip("uvw", orc.velocity_fields()):
            field.data[...] = f[name]
        orc.pressure[...] = 0.0
        project(orc, solve)
        return sum(float((field.data * w).sum()) for field, w in zip(orc.velocity_fields(), W))

    fd_check(dot, evaluate, dict(zip("uvw", U)), dict(zip("uvw", D)), h=1e-2, rel=1e-3)


def test_raymarch_adjoint(device, oracle64):
    import torch

    from plume_amd._lib import LookParams, call, ptr, stream
    from plume_amd.render.camera import make_camera
    from plume_amd.render.raymarch import march
    from plume_amd.sim.adjoint import _square_view

    # 1 cm cells with the near-origin orthographic view of the differentiable rollout (rays started on the
    # camera plane, not 1e4 m out, so FP32 sample positions carry no ~5e-4 m rounding); the opacity per
    # density is that of a 1 m grid scaled to keep the optical depth per step
    rng = np.random.default_rng(6)
    gpu = gpu_grid(CELLS, 0.01, DEFAULT_FACES, ambient=1.0, device=device)
    orc = oracle_grid(oracle64, CELLS, 0.01, DEFAULT_FACES, ambient=1.0)
    s, t = smooth_field(rng, CELLS, 0.3, 0.5), smooth_field(rng, CELLS, 0.3, 1.5)
    set_gpu(gpu, {"smoke": s, "temperature": t})
    pixels, steps = 24, 40
    view = _square_view(gpu, make_camera(20.0, 30.0), pixels)
    look = LookParams()
    look.opacity_per_density, look.glow_start_k, look.smoke_hides_heat = 20.0, 0.5, 1.0
    look.normal_fade_thickness, look.shade_layout, look.shade_fill = 0.5, 1, 0.3
    W = rng.standard_normal((pixels, pixels, 5))
    W[..., 2:] = 0.0
    wb = torch.from_numpy(W.astype(np.float32)).to(device)
    stride = steps
    record = torch.zeros(pixels * pixels * stride * 4, dtype=torch.float32, device=device)
    sb, tb = torch.zeros_like(gpu.smoke.data), torch.zeros_like(gpu.temperature.data)
    call("plm_adj_raymarch", ctypes.byref(gpu.info), ptr(gpu.smoke.data), ptr(gpu.temperature.data),
         ctypes.byref(view), ctypes.byref(look), ptr(wb), pixels, pixels, steps, ptr(record), ptr(sb), ptr(tb),
         stream())
    Ds, Dt = rng.standard_normal(CELLS), rng.standard_normal(CELLS)
    dot = float((to_np(sb) * Ds).sum() + (to_np(tb) * Dt).sum())
    # the GPU forward and the oracle render the same view
    image = torch.zeros((pixels, pixels, 5), dtype=torch.float32, device=device)
    march(gpu.smoke.data, gpu.temperature.data, torch.zeros((1, 1, 1, 3), device=device),
          torch.ones((1, 1, 1), device=device), gpu.info, view, look, image, steps)
    oview = oracle64.view(tuple(view.right), tuple(view.up), tuple(view.toward), view.left_m, view.top_m, view.pixel_m,
                          view.perspective, (0.0, 0.0, 0.0))
    olook = oracle64.look(20.0, 0.5, 1.0, 0.5, 1, 0.3)
    light = np.ones((1, 1, 1))
    grad = np.zeros((1, 1, 1, 3))

    def evaluate(f):
        out = np.zeros((pixels, pixels, 5))
        oracle64.raymarch(orc.info, f["smoke"], f["temperature"], grad, light, oview, olook, out, steps)
        return float((out * W).sum())

    base = np.zeros((pixels, pixels, 5))
    oracle64.raymarch(orc.info, s.astype(np.float32).astype(np.float64), t.astype(np.float32).astype(np.float64),
                      grad, light, oview, olook, base, steps)
    assert np.allclose(to_np(image)[..., :2], base[..., :2], rtol=1e-3, atol=1e-3)
    fields = {"smoke": s.astype(np.float32).astype(np.float64), "temperature": t.astype(np.float32).astype(np.float64)}
    fd_check(dot, evaluate, fields, {"smoke": Ds, "temperature": Dt}, h=1e-6, rel=1e-3)


# ---------------------------------------------------------------------- the whole rollout

ROLL_CELLS = (16, 24, 16)
SUBSTEPS, CHECKPOINT, PIXELS, MARCH = 8, 4, 24, 48


def diff_scene(device):
    """A puff scene on 1 m cells (see test_raymarch_adjoint) started from a smooth, asymmetric state with
    Gaussian sources off the grid's mirror planes. From rest the clamped MacCormack meets exact ties (zero
    velocity, ambient temperature, a flat-topped emitter, mirror-symmetric stencils) where it has no derivative:
    there its directional derivative is not linear in the direction, and no adjoint can match a finite
    difference. A generic state has no ties."""
    import torch

    from plume_amd.recipe.derived import derive
    from plume_amd.recipe.load import fill_dependent_defaults, parse_recipe
    from plume_amd.sim.adjoint import DiffSim, Sources
    from plume_amd.sim.scenes import puff_table

    table = puff_table(ROLL_CELLS, frames=2, radius_cells=4)
    table["domain"]["cell_m"] = 1.0
    table["domain"]["size_m"] = [float(n) for n in ROLL_CELLS]
    table["emitter"][0]["center_m"] = [0.0, 10.0, 0.0]
    table["emitter"][0]["radius_m"] = 4.0
    table["render"]["glow_start_k"] = 0.5  # below ambient: the heat channel's max(T - glow, 0) stays smooth
    table["render"]["opacity_per_density"] = 0.5
    recipe = fill_dependent_defaults(parse_recipe(table), "adjoint")
    derived = derive(recipe)
    sim = DiffSim(recipe, derived, device, pixels=PIXELS, march_steps=MARCH, tolerance=1e-6)
    nx, ny, nz = ROLL_CELLS
    center = (7.7, 10.3, 8.2)

    def gaussian(shape, offset, strength):
        idx = np.indices(shape).astype(np.float64)
        d2 = sum((idx[a] + offset[a] - center[a]) ** 2 for a in range(3))
        return (strength * np.exp(-d2 / (2 * 2.5 ** 2))).astype(np.float32)

    host = [gaussian((nx, ny, nz), (0.5, 0.5, 0.5), 2.0), gaussian((nx, ny, nz), (0.5, 0.5, 0.5), 4.0),
            gaussian((nx, ny + 1, nz), (0.5, 0.0, 0.5), 1.0)]
    sources = Sources(*(torch.from_numpy(x).to(device) for x in host))
    rng = np.random.default_rng(21)
    initial = {"u": smooth_field(rng, (nx + 1, ny, nz), 1.5), "v": smooth_field(rng, (nx, ny + 1, nz), 1.5),
               "w": smooth_field(rng, (nx, ny, nz + 1), 1.5), "smoke": smooth_field(rng, (nx, ny, nz), 0.3, 0.6),
               "temperature": smooth_field(rng, (nx, ny, nz), 1.0, recipe.domain.ambient_k + 2.0)}
    initial = {k: v.astype(np.float32) for k, v in initial.items()}
    return recipe, derived, sim, sources, initial


def gpu_initial(initial, device):
    import torch

    return {k: torch.from_numpy(v).to(device) for k, v in initial.items()}


def oracle_images(o, recipe, derived, sim, sources64, alpha, initial):
    """The same rollout in FP64 on the oracle from `initial` (tests/adjoint_oracle.py)."""
    from adjoint_oracle import Camera, rollout_images

    return rollout_images(o, recipe, derived, sim.dt, Camera.of(sim), sources64, alpha, SUBSTEPS, CHECKPOINT,
                          PIXELS, MARCH, initial)


def test_rollout_gradient_matches_fp64_finite_differences(device, oracle64):
    import torch

    from plume_amd.sim.adjoint import Sources, rollout

    recipe, derived, sim, sources, initial = diff_scene(device)
    start = gpu_initial(initial, device)
    with torch.no_grad():
        target = rollout(Sources(*(1.1 * t for t in sources.tensors())), torch.tensor(0.1, dtype=torch.float64),
                         sim, SUBSTEPS, CHECKPOINT, start).clone()
    params = [t.clone().requires_grad_(True) for t in sources.tensors()]
    alpha = torch.tensor(0.1, dtype=torch.float64, requires_grad=True)
    images = rollout(Sources(*params), alpha, sim, SUBSTEPS, CHECKPOINT, start)
    loss = ((images - target) ** 2).sum()
    loss.backward()
    base64 = [to_np(t).astype(np.float64) for t in sources.tensors()]
    # the forward images agree with the oracle's FP64 rollout
    images64 = oracle_images(oracle64, recipe, derived, sim, base64, 0.1, initial)
    assert np.abs(to_np(images) - images64).max() <= 2e-3 * max(1.0, np.abs(images64).max())
    # The adjoint is the vector-Jacobian product J^T bar with bar = dloss/dimages of the FP32 images; the check
    # differentiates <bar, images(sources, alpha)> in FP64 with that same bar. (Differencing the loss itself
    # would weight the FP64 Jacobian by the FP64 residual, whose difference from the FP32 one - the images'
    # rounding - is not small against a target only 10% away.)
    bar64 = 2.0 * (to_np(images).astype(np.float64) - to_np(target).astype(np.float64))

    def functional(srcs, a):
        return float((oracle_images(oracle64, recipe, derived, sim, srcs, a, initial) * bar64).sum())

    rng = np.random.default_rng(9)
    results = {}
    for which in range(4):
        dirs = [np.zeros_like(b) for b in base64]
        dalpha = 0.0
        if which < 3:
            dirs[which] = rng.standard_normal(base64[which].shape) * base64[which]
        else:
            dalpha = 0.1
        adj = sum(float((to_np(p.grad).astype(np.float64) * d).sum()) for p, d in zip(params, dirs))
        adj += float(alpha.grad) * dalpha
        fds = []
        for h in (1e-3, 1e-5):
            plus = functional([b + h * d for b, d in zip(base64, dirs)], 0.1 + h * dalpha)
            minus = functional([b - h * d for b, d in zip(base64, dirs)], 0.1 - h * dalpha)
            fds.append((plus - minus) / (2 * h))
        results[("smoke", "temperature", "lift", "alpha")[which]] = (adj, *fds)
    print("directional derivatives (adjoint, FP64 central difference h=1e-3, h=1e-5):", results)
    # h = 1e-5 (relative to the sources): a step of 1e-3 already crosses kinks of MacCormack's clamp
    for name, (adj, fd3, fd5) in results.items():
        assert abs(adj - fd5) <= 1e-3 * abs(fd5) + 1e-12, (name, adj, fd3, fd5)
