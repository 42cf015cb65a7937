"""Writes tests/golden/adjoint_fd_64x96x64.json: FP64 central differences of the differentiable rollout at
64 x 96 x 64 on the CPU oracle, the checker of tests/test_gpu_configs.py's gradient test.

The scene is plume_amd.sim.scenes.differentiable_puff at its full size: the puff from rest, 32 substeps,
(heat, thickness) images of 128 x 128 pixels with 192 samples per ray after every 8 substeps, Gaussian-falloff
sphere sources (smoke 2 /s, heat 4 K/s, lift 1 m/s^2 on v faces) and alpha = 0.1. The functional is
F(sources, alpha) = <W, images> with the fixed weights W of `weights()` (a smooth pattern, exactly
representable in FP32), so the GPU test gets the same vector-Jacobian product from images.backward(W). For each
direction (each source scaled by itself; alpha scaled by itself) the file holds (F(+h) - F(-h)) / 2h at
h = 1e-3 and h = 1e-5 (relative steps), from FP64 rollouts of the oracle (tests/adjoint_oracle.py).

    python tests/golden/make_adjoint_fd.py        # ~15 minutes on 8 cores

Test infrastructure: nothing here is imported by the package.
"""

import json
import sys
import time
from pathlib import Path
from types import SimpleNamespace

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))
sys.path.insert(0, str(HERE.parent))

CELLS = (64, 96, 64)
SUBSTEPS, CHECKPOINT, PIXELS, MARCH = 32, 8, 128, 192
ALPHA = 0.1
RADIUS_CELLS = 6.0
STRENGTHS = (2.0, 4.0, 1.0)  # smoke /s, heat K/s, lift m/s^2 (sphere_sources defaults)
NAMES = ("smoke", "temperature", "lift")
STEPS = (1e-3, 1e-5)
OUT = HERE / "adjoint_fd_64x96x64.json"


def weights() -> np.ndarray:
    """(checkpoints, pixels, pixels, 2) smooth weights, float32 values (exact in FP64 too)."""
    c, y, x, ch = np.meshgrid(np.arange(SUBSTEPS // CHECKPOINT), np.arange(PIXELS), np.arange(PIXELS), np.arange(2),
                              indexing="ij")
    w = 0.5 + np.sin(0.05 * x + 0.3 * c + 1.1 * ch) * np.cos(0.07 * y - 0.2 * c)
    return w.astype(np.float32)


def gaussian_sources(cells):
    """sphere_sources(..., falloff="gaussian") of plume_amd.sim.adjoint, in float32 numpy: the centre 10 cells
    above the floor centre, exp(-d^2 / (2 (r/2)^2)) inside the radius."""
    nx, ny, nz = cells
    center = (nx / 2.0, 10.0, nz / 2.0)

    def coverage(shape, offset):
        coords = [np.arange(shape[a], dtype=np.float32) + np.float32(offset[a]) for a in range(3)]
        X, Y, Z = np.meshgrid(*coords, indexing="ij")
        d = np.sqrt((X - np.float32(center[0])) ** 2 + (Y - np.float32(center[1])) ** 2
                    + (Z - np.float32(center[2])) ** 2)
        sigma = np.float32(0.5 * RADIUS_CELLS)
        return np.where(d < np.float32(RADIUS_CELLS), np.exp(-d * d / (np.float32(2.0) * sigma * sigma)),
                        np.float32(0.0)).astype(np.float32)

    cells_cov = coverage((nx, ny, nz), (0.5, 0.5, 0.5))
    faces_cov = coverage((nx, ny + 1, nz), (0.5, 0.0, 0.5))
    return [np.float32(STRENGTHS[0]) * cells_cov, np.float32(STRENGTHS[1]) * cells_cov,
            np.float32(STRENGTHS[2]) * faces_cov]


def scene():
    """(recipe, derived, dt, Camera) of the differentiable puff, built on the host only."""
    from adjoint_oracle import Camera
    from plume_amd.render.camera import make_camera
    from plume_amd.render.frame_renderer import look_params
    from plume_amd.sim.adjoint import _square_view
    from plume_amd.sim.grid import grid_info
    from plume_amd.sim.scenes import differentiable_recipe

    recipe, derived = differentiable_recipe(CELLS, RADIUS_CELLS)
    info = grid_info(CELLS, recipe.domain.cell_m, derived.domain_min_m)
    camera = make_camera(recipe.camera.elevation_deg, recipe.camera.azimuth_deg)
    view = _square_view(SimpleNamespace(info=info, cells=CELLS), camera, PIXELS)
    look = look_params(recipe.render, shade_layout=True)
    return recipe, derived, float(derived.dt_s), Camera.of(SimpleNamespace(view=view, look=look))


def main() -> None:
    from adjoint_oracle import rollout_images
    from oracle import Oracle

    o = Oracle("f64")
    recipe, derived, dt, camera = scene()
    base = [s.astype(np.float64) for s in gaussian_sources(CELLS)]
    w = weights().astype(np.float64)

    def functional(sources, alpha):
        images = rollout_images(o, recipe, derived, dt, camera, sources, alpha, SUBSTEPS, CHECKPOINT, PIXELS, MARCH)
        return float((images * w).sum()), images

    start = time.perf_counter()
    f0, images = functional(base, ALPHA)
    print(f"base rollout {time.perf_counter() - start:.1f} s, F = {f0:.12e}", flush=True)
    fd = {}
    for which, name in enumerate(NAMES + ("alpha",)):
        fd[name] = {}
        for h in STEPS:
            signed = []
            for sign in (1.0, -1.0):
                sources = [b + sign * h * b if k == which else b for k, b in enumerate(base)]
                alpha = ALPHA * (1.0 + sign * h) if name == "alpha" else ALPHA
                signed.append(functional(sources, alpha)[0])
            fd[name][f"{h:g}"] = (signed[0] - signed[1]) / (2.0 * h)
            print(f"{name} h={h:g}: {fd[name][f'{h:g}']:.12e}", flush=True)
    record = {
        "scene": f"differentiable puff {CELLS[0]}x{CELLS[1]}x{CELLS[2]} from rest, {SUBSTEPS} substeps, (heat, "
                 f"thickness) {PIXELS}x{PIXELS} px every {CHECKPOINT} substeps, {MARCH} samples per ray, "
                 f"alpha {ALPHA}",
        "functional": "sum(W * images), W = make_adjoint_fd.weights()",
        "directions": "each source scaled by itself; alpha scaled by itself (d alpha = alpha)",
        "oracle": "FP64 C oracle rollout (tests/adjoint_oracle.py), projection solved to 1e-13",
        "F": f0,
        "image_sums": images.sum(axis=(1, 2)).tolist(),
        "central_differences": fd,
    }
    OUT.write_text(json.dumps(record, indent=1) + "\n")
    print(f"wrote {OUT} in {time.perf_counter() - start:.0f} s", flush=True)


if __name__ == "__main__":
    main()
