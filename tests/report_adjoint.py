"""Gradient report of the differentiable rollout at 64 x 96 x 64 against FP64 central differences.

The scene (plume_amd.sim.scenes.differentiable_puff): the puff from rest, 32 substeps, the (heat, thickness)
images of 128 x 128 pixels after every 8 substeps (the state is stored only there), the L2 loss against the
images of the same rollout with the smoke, heat and lift sources 10% stronger. Gradients: the three per-cell
source fields and alpha.

Two phases, so that the FP64 rollouts (about a minute each on 8 cores) use no GPU time:

    python tests/report_adjoint.py gpu  --npz OUT.npz [--warm N]   # on a GPU: the adjoint's derivatives
    python tests/report_adjoint.py fp64 --npz OUT.npz              # anywhere: FP64 central differences

With --warm N the rollout starts from the state N plain substeps of the same sources reach (the same start for
the target): from rest the clamped MacCormack meets exact ties (still air, zero velocity) where the rollout
has no derivative along directions of mixed sign, while a developed plume has none in the sources' reach.

The check differentiates <bar, images(sources, alpha)> on the FP64 oracle with bar = dloss/dimages of the
FP32 images, the vector-Jacobian product that the adjoint computes. Directions: each source scaled by
itself (the direction of the target's perturbation), alpha, and each source times a smooth random field.
"""

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE))

CELLS = (64, 96, 64)
SUBSTEPS, CHECKPOINT, PIXELS, MARCH = 32, 8, 128, 192
ALPHA = 0.1
NAMES = ("smoke", "temperature", "lift")


def smooth(rng, shape):
    """A smooth random field of order 1 (three random plane waves)."""
    idx = np.indices(shape).astype(np.float64)
    f = np.zeros(shape)
    for _ in range(3):
        k = rng.uniform(0.05, 0.3, size=3)
        f += np.sin(sum(k[a] * idx[a] for a in range(3)) + rng.uniform(0, 2 * np.pi))
    return f / 3.0


def directions(sources64):
    """[(name, source directions (3 arrays), alpha direction)]."""
    out = []
    for k, name in enumerate(NAMES):
        d = [np.zeros_like(s) for s in sources64]
        d[k] = sources64[k].copy()
        out.append((f"{name} x itself", d, 0.0))
    out.append(("alpha", [np.zeros_like(s) for s in sources64], ALPHA))
    rng = np.random.default_rng(20261203)
    for k in (0, 2):
        d = [np.zeros_like(s) for s in sources64]
        d[k] = sources64[k] * smooth(rng, sources64[k].shape)
        out.append((f"{NAMES[k]} x smooth random", d, 0.0))
    return out


def gpu_phase(npz: Path, warm: int) -> None:
    import torch

    from adjoint_oracle import Camera
    from plume_amd.sim.adjoint import Sources, rollout
    from plume_amd.sim.device import init_device
    from plume_amd.sim.scenes import differentiable_puff

    device = init_device("cuda")
    recipe, derived, sim, sources = differentiable_puff(device, CELLS, PIXELS, MARCH)
    initial = None
    if warm > 0:
        sim.reset()
        with torch.no_grad():
            for _ in range(warm):
                sim.substep(sources, ALPHA)
        initial = {k: v for k, v in sim.state().items() if k != "pressure"}
    with torch.no_grad():
        target = rollout(Sources(*(1.1 * t for t in sources.tensors())), torch.tensor(ALPHA, dtype=torch.float64),
                         sim, SUBSTEPS, CHECKPOINT, initial).clone()
    params = [t.clone().requires_grad_(True) for t in sources.tensors()]
    alpha = torch.tensor(ALPHA, dtype=torch.float64, requires_grad=True)
    torch.cuda.synchronize()
    start = time.perf_counter()
    images = rollout(Sources(*params), alpha, sim, SUBSTEPS, CHECKPOINT, initial)
    loss = ((images - target) ** 2).sum()
    loss.backward()
    torch.cuda.synchronize()
    seconds = time.perf_counter() - start
    sources64 = [t.detach().cpu().numpy().astype(np.float64) for t in sources.tensors()]
    grads = [p.grad.detach().cpu().numpy().astype(np.float64) for p in params]
    adjoint = []
    for name, dirs, dalpha in directions(sources64):
        value = sum(float((g * d).sum()) for g, d in zip(grads, dirs)) + float(alpha.grad) * dalpha
        adjoint.append(value)
        print(f"[gpu] {name}: adjoint {value:.9e}", flush=True)
    start_fields = {f"initial_{k}": v.cpu().numpy() for k, v in (initial or {}).items()}
    np.savez_compressed(npz, smoke=sources64[0].astype(np.float32), temperature=sources64[1].astype(np.float32),
                        lift=sources64[2].astype(np.float32), images=images.detach().cpu().numpy(),
                        target=target.cpu().numpy(), camera=Camera.of(sim).as_array(), dt=sim.dt,
                        adjoint=np.array(adjoint), loss=float(loss.detach()), seconds=seconds, warm=warm,
                        **start_fields)
    print(f"[gpu] loss {float(loss.detach()):.6e}, forward + backward {seconds:.3f} s; wrote {npz}", flush=True)


def fp64_phase(npz: Path, steps: list[float], out: Path | None) -> None:
    from adjoint_oracle import Camera, rollout_images
    from oracle import Oracle
    from plume_amd.sim.scenes import differentiable_recipe

    data = np.load(npz)
    recipe, derived = differentiable_recipe(CELLS)
    o = Oracle("f64")
    camera = Camera.from_array(data["camera"])
    dt = float(data["dt"])
    sources64 = [data[n].astype(np.float64) for n in ("smoke", "temperature", "lift")]
    bar64 = 2.0 * (data["images"].astype(np.float64) - data["target"].astype(np.float64))

    initial = {k[len("initial_"):]: data[k] for k in data.files if k.startswith("initial_")} or None

    def images(srcs, a):
        return rollout_images(o, recipe, derived, dt, camera, srcs, a, SUBSTEPS, CHECKPOINT, PIXELS, MARCH,
                              initial)

    start = time.perf_counter()
    base = images(sources64, ALPHA)
    print(f"[fp64] base rollout {time.perf_counter() - start:.1f} s; FP32 vs FP64 images: max |diff| "
          f"{np.abs(data['images'] - base).max():.3e} of max {np.abs(base).max():.3e}", flush=True)
    rows = []
    for (name, dirs, dalpha), adj in zip(directions(sources64), data["adjoint"]):
        row = {"direction": name, "adjoint": float(adj)}
        for h in steps:
            plus = float((images([s + h * d for s, d in zip(sources64, dirs)], ALPHA + h * dalpha) * bar64).sum())
            minus = float((images([s - h * d for s, d in zip(sources64, dirs)], ALPHA - h * dalpha) * bar64).sum())
            fd = (plus - minus) / (2 * h)
            row[f"fd_h{h:g}"] = fd
            row[f"rel_err_h{h:g}"] = abs(float(adj) - fd) / abs(fd)
        rows.append(row)
        print("[fp64] " + json.dumps(row), flush=True)
    start_state = f"after {int(data['warm'])} plain substeps" if initial else "from rest"
    report = {"scene": f"differentiable puff {CELLS[0]}x{CELLS[1]}x{CELLS[2]}, {SUBSTEPS} substeps {start_state}, "
                       f"images {PIXELS}x{PIXELS} every {CHECKPOINT} substeps, {MARCH} samples per ray",
              "loss": float(data["loss"]), "gpu_forward_backward_s": float(data["seconds"]),
              "fp32_vs_fp64_image_max_abs": float(np.abs(data["images"] - base).max()),
              "image_max": float(np.abs(base).max()), "rows": rows}
    if out is not None:
        out.write_text(json.dumps(report, indent=1))


def main() -> None:
    parser = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    parser.add_argument("phase", choices=("gpu", "fp64"))
    parser.add_argument("--npz", type=Path, required=True)
    parser.add_argument("--steps", type=float, nargs="+", default=[1e-3, 1e-5])
    parser.add_argument("--json", type=Path, default=None)
    parser.add_argument("--warm", type=int, default=0, help="plain substeps before the rollout (gpu phase)")
    args = parser.parse_args()
    if args.phase == "gpu":
        gpu_phase(args.npz, args.warm)
    else:
        fp64_phase(args.npz, args.steps, args.json)


if __name__ == "__main__":
    main()
