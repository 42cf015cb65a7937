"""Benchmark: cell-substeps/s of the full simulation substep at 128 x 192 x 128 on one MI355X.

A step is one substep of the buoyant smoke puff on a 128 x 192 x 128 grid: emit (while active),
buoyancy, vorticity confinement, MacCormack advection of u, v, w, projection with multigrid V(2,2) cycles
to a 1e-5 relative residual (FP64 norm on the device), MacCormack advection of smoke (with the smoke
budget) and temperature. Inputs live in HBM before the timed region; W warmup substeps (the 12 emitting
frames by default) run first, then exactly K substeps are timed between device synchronisations.

One JSON line on stdout. Besides the contract's keys it carries:
  roofline       the dominant kernel (fused red-black sweep at 128 x 192 x 128): algorithmic bytes per
                 launch / average launch time from HIP events, against 8 TB/s HBM;
  substep_roofline  B_alg = 4 N (36 + 18.43 C) bytes per substep (C = V-cycles run) x substeps/s / 8 TB/s;
  raymarch       256 x 256 px, 192 samples per ray, light volume with 64 steps per cell;
  frame_cache    the viewer's binary16 frame-cache kernels at 128 x 192 x 128 (GB/s, HIP events);
  recipe_run     the smoke-puff recipe run end to end (80 x 100 x 80 cells, 203 substeps, 64 frames,
                 sheet assembly; again in the normal layout): stage wall times beside the reference's
                 published RTX 2080 Ti ones;
  mixed, ensemble, adjoint
                 the three additions at their own workloads (256 x 384 x 256 mixed-precision projection;
                 32-member batched ensemble; 64 x 96 x 64 differentiable rollout), as --mixed, --ensemble and
                 --adjoint print them alone (--no-additions leaves them out);
  cpu_baseline   the C port of the same operators (the oracle's source) built with -O3 -march=native
                 -fopenmp, on a bounded sample at all host threads and at one thread, with the CPU model.
"""

import argparse
import ctypes
import json
import math
import os
import sys
import time
import traceback
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBPS = 8000.0
METRIC = "cell-substeps/s at 128×192×128 (+% HBM roofline) vs plain GPU and CPU"
GRID = (128, 192, 128)
CELL_M = 0.01


def scene_table(cells=GRID, frames=96, radius_cells=12, tolerance=1e-5, cycles=2):
    """The smoke puff (plume_amd.sim.scenes.puff_table) with alpha 0.1, beta 1, confinement 0.3."""
    from plume_amd.sim.scenes import puff_table

    return puff_table(cells, frames=frames, radius_cells=radius_cells, tolerance=tolerance, cycles=cycles)


def substep_bytes(cells: int, cycles: float) -> float:
    """B_alg = 4 N (36 + 18.43 C): the fixed 36 floats per cell plus 16.43 (multigrid) + 2 (FP64 norm) per
    V-cycle."""
    return 4.0 * cells * (36.0 + 18.43 * cycles)


def sweep_bytes(cells: int) -> float:
    """One fused red-black sweep: read q and rhs, write q (3 floats per cell)."""
    return 12.0 * cells


def lib_timing(family: str):
    from plume_amd._lib import LIB

    return LIB, family.encode()


def time_kernel_family(run, family: str) -> tuple[float, int]:
    """Runs `run()` with HIP-event timing of every launch of `family`; returns (avg ms, launches)."""
    import torch

    lib, name = lib_timing(family)
    lib.plm_timing_enable(name)
    try:
        run()
        torch.cuda.synchronize()
        total = ctypes.c_double()
        count = ctypes.c_int64()
        rc = lib.plm_timing_read(ctypes.byref(total), ctypes.byref(count))
        if rc != 0:
            raise RuntimeError(lib.plm_last_error().decode())
    finally:
        lib.plm_timing_disable()
    return (total.value / count.value if count.value else float("nan")), count.value


TRAFFIC_FILE = ROOT / "profiles" / "rb_sweep_tile_128x192x128_traffic.json"


def measured_traffic():
    """HBM bytes per launch of the top-level sweep from rocprofv3 FETCH_SIZE / WRITE_SIZE runs, each scaled
    by what the counter reads for a streaming kernel of known size (the file records both runs)."""
    if not TRAFFIC_FILE.exists():
        return None
    facts = json.loads(TRAFFIC_FILE.read_text())
    return facts.get("hbm_bytes_per_launch")


def copy_probe(mib: int, repeats: int = 20) -> float:
    """A plain streaming kernel (dst = -src over `mib` MiB of FP32) `repeats` times; returns GB/s counted as
    read + write bytes. Also the known-size kernel the HBM counters are scaled by."""
    import torch

    count = mib * (1 << 20) // 4
    src = torch.ones(count, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.neg(src, out=dst)
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(repeats):
        torch.neg(src, out=dst)
    end.record()
    end.synchronize()
    seconds = start.elapsed_time(end) * 1e-3 / repeats
    return 2.0 * count * 4 / seconds / 1e9


def cpu_facts() -> dict:
    """The host CPU: lscpu's model name, sockets, cores per socket and threads per core, and the logical CPUs
    this process may run on."""
    import subprocess

    facts = {"logical_cpus_available": len(os.sched_getaffinity(0)), "os_cpu_count": os.cpu_count()}
    try:
        text = subprocess.run(["lscpu"], capture_output=True, text=True, check=True).stdout
    except (OSError, subprocess.CalledProcessError):
        return facts
    wanted = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
              "Thread(s) per core": "threads_per_core", "CPU(s)": "cpus"}
    for line in text.splitlines():
        key, _, value = line.partition(":")
        if key.strip() in wanted:
            facts[wanted[key.strip()]] = value.strip()
    return facts


def cpu_baseline(gpu_grid, recipe, derived, time_s, steps_by_threads):
    """The C port of the substep's operators (oracle/plume_oracle.c, the oracle's own source) built for this
    host with gcc -O3 -march=native -fopenmp, FP32, on the GPU's warm state: {threads: seconds for that many
    substeps} for each (threads, substeps) pair."""
    import tempfile

    import numpy as np

    from oracle import Oracle
    from oracle.build import build_native
    from oracle.sim import Stepper, make_grid

    library = build_native(Path(tempfile.mkdtemp(prefix="plume_cpu_")) / "liborc_native_f32.so")
    o = Oracle("f32", library=library)
    out = {}
    for threads, steps in steps_by_threads:
        o.set_threads(threads)
        grid = make_grid(o, derived, recipe.domain.cell_m, recipe.domain.faces, recipe.domain.ambient_k)
        for name in ("u", "v", "w", "smoke", "temperature"):
            getattr(grid, name).data[...] = getattr(gpu_grid, name).data.cpu().numpy()
        grid.pressure[...] = gpu_grid.pressure.cpu().numpy()
        stepper = Stepper(grid, recipe)
        start = time.perf_counter()
        for index in range(steps):
            stepper.step(time_s + index * derived.dt_s, derived.dt_s)
        out[threads] = (time.perf_counter() - start, steps)
        assert np.isfinite(grid.smoke.data).all()
        print(f"[bench] cpu baseline: {threads} threads, {steps} substeps in {out[threads][0]:.2f} s",
              file=sys.stderr, flush=True)
    return out


def frame_cache_bench(grid, repeats: int = 20) -> dict:
    """The viewer's frame-cache kernels on the bench's warm 128 x 192 x 128 fields, each timed over `repeats`
    launches with HIP events bound to each dispatch (the library's timing families, as for the sweep):
    narrowing a scalar field to binary16 (4 B read + 2 B written per cell), the fused cell-centre velocity to
    binary16 (three face arrays read, 6 B written per cell) and widening back (2 B read + 4 B written per
    value)."""
    import torch

    from plume_amd._lib import call, ptr, stream

    cells = grid.smoke.data.numel()
    half = torch.empty(cells, dtype=torch.float16, device="cuda")
    wide = torch.empty(cells, dtype=torch.float32, device="cuda")
    centre = torch.empty(3 * cells, dtype=torch.float16, device="cuda")
    u, v, w = (f.data for f in (grid.u, grid.v, grid.w))
    kernels = {
        "narrow_f16": (lambda: call("plm_narrow_f16", cells, ptr(grid.smoke.data), ptr(half, torch.float16),
                                    stream()), 6.0 * cells),
        "widen_f16": (lambda: call("plm_widen_f16", cells, ptr(half, torch.float16), ptr(wide), stream()),
                      6.0 * cells),
        "narrow_center_velocity": (lambda: call("plm_narrow_center_velocity", ctypes.byref(grid.info), ptr(u), ptr(v), ptr(w),
                                                ptr(centre, torch.float16), stream()),
                                   4.0 * (u.numel() + v.numel() + w.numel()) + 6.0 * cells),
    }
    out = {}
    for name, (launch, nbytes) in kernels.items():
        launch()
        torch.cuda.synchronize()
        ms, launches = time_kernel_family(lambda: [launch() for _ in range(repeats)], name)
        seconds = ms * 1e-3
        out[name] = {"us": seconds * 1e6, "launches": launches, "bytes": nbytes, "GBps": nbytes / seconds / 1e9,
                     "hbm_frac": nbytes / seconds / 1e9 / HBM_PEAK_GBPS}
    return out


PUFF_RECIPE = ROOT / "tests" / "golden" / "smoke_puff.toml"  # the reference's recipes/smoke_puff.toml
# The reference's stage timings of this run on an RTX 2080 Ti (Warp CUDA): simulate 2.30 s (render excluded),
# render 0.16 s in the shade layout (light volume, march and copy to the host, 64 frames) and 0.27 s in the
# normal layout (normal field, march and copy)
PUFF_2080TI_S = {"simulate": 2.30, "render": 0.16, "render_normal": 0.27}
NORMAL_LAYOUT_SET = "output.channels=heat,thickness,normal_x,normal_y"


def recipe_run() -> dict:
    """`plume run recipes/smoke_puff.toml` end to end on the HIP path (80 x 100 x 80 cells, 64 frames of a
    160 x 200 px view, sheet assembly and numbers): stage wall times of the second of two runs."""
    from plume_amd.pipeline import run_recipe

    run_recipe(PUFF_RECIPE, preview=False, device="cuda")
    run = run_recipe(PUFF_RECIPE, preview=False, device="cuda")
    normal = run_recipe(PUFF_RECIPE, preview=False, device="cuda", sets=(NORMAL_LAYOUT_SET,))
    seconds = {**run.seconds, "render_normal": normal.seconds["render"]}
    cells = run.derived.cells[0] * run.derived.cells[1] * run.derived.cells[2]
    substeps = run.steps.total_steps
    print(f"[bench] recipe run: {substeps} substeps, simulate {run.seconds['simulate']:.3f} s, render "
          f"{run.seconds['render']:.3f} s (normal layout {seconds['render_normal']:.3f} s), assemble "
          f"{run.seconds['assemble']:.3f} s", file=sys.stderr, flush=True)
    return {"recipe": "smoke_puff.toml (the reference's own scene)", "cells": list(run.derived.cells),
            "frames": len(run.raw_frames), "substeps": substeps, "simulate_s": run.seconds["simulate"],
            "render_s": run.seconds["render"], "render_normal_s": seconds["render_normal"],
            "assemble_s": run.seconds["assemble"], "cell_substeps_per_s": cells * substeps / run.seconds["simulate"],
            "reference_rtx2080ti_s": PUFF_2080TI_S,
            "speedup_vs_reference_rtx2080ti": {k: PUFF_2080TI_S[k] / seconds[k] for k in PUFF_2080TI_S}}


MIXED_GRID = (256, 384, 256)


def mixed_bench(args) -> dict:
    """Mixed-precision pressure projection at 256x384x256: the seeded divergent field
    (64 Gaussian vortex blobs plus a radial emitter source), solved from q = 0 to a common relative residual
    by the plain FP32 multigrid (tolerance mode) and by FP32 iterative refinement around BF16 and FP16
    multigrid with an FP64 coarsest 8x12x8 level (f16x2: the FP16 form with the simulation level swept in
    packed FP16 arithmetic), every form with the same V(sweeps, sweeps) cycle. The
    common target is max(tolerance, 2 x the FP32 floor),
    the floor being where plain FP32 V-cycles stall (below it no FP32 pressure gets; the FP64 oracle's
    solution rounded to FP32 sits at the same level). Solve time, FP64 relative residual and
    post-projection divergence per form; the errors against an FP64 oracle solution are
    tests/report_mixed.py."""
    from plume_amd.sim.device import init_device
    from plume_amd.sim.mixed import divergent_problem, fp32_floor, solve_forms

    device = init_device("cuda")
    problem = divergent_problem(MIXED_GRID, device)
    floor = fp32_floor(problem.grid)
    target = max(args.mixed_tolerance, 2.0 * floor)
    print(f"[bench] mixed: FP32 floor {floor:.3g}, target {target:.3g}", file=sys.stderr, flush=True)
    results = {}
    for name, (entry, _) in solve_forms(problem, target, sweeps=args.mixed_sweeps).items():
        results[name] = entry
        print(f"[bench] mixed {name}: {entry['seconds'] * 1e3:.2f} ms, residual {entry['relative_residual']:.3g}, "
              f"post div {entry['post_projection_divergence']:.3g}", file=sys.stderr, flush=True)
    plain = results["f32"]["cell_solves_per_s"]
    for name in ("bf16", "f16", "f16x2"):
        results[name]["speedup_vs_f32"] = results[name]["cell_solves_per_s"] / plain
    return {"metric": "pressure-solve cells/s at 256x384x256 to a common relative residual (FP32 vs BF16/FP16 "
                      "multigrid with FP32 iterative refinement; f16x2: FP16 with the simulation level swept in "
                      "packed FP16 arithmetic)", "grid": list(MIXED_GRID),
            "tolerance": args.mixed_tolerance, "fp32_floor": floor, "target": target,
            "target_rule": "max(tolerance, 2 x FP32 floor); floor = where plain FP32 V-cycles stall",
            "sweeps": args.mixed_sweeps,
            "data": "synthetic (seeded divergent field: 64 Gaussian vortex blobs + radial emitter source)",
            "forms": results}


def ensemble_bench(args) -> dict:
    """Batched ensemble: 32 smoke puffs with seeded sizes over 48^3, 64x96x64, 96x128x96
    and 128x192x128, emitter radii 4-12 cells, alpha in [0.05, 0.2] and epsilon in [0, 0.5]. The 12 emitting
    frames and one more run eagerly; then the same number of frames is timed twice from the state reached:
    batched (one captured HIP graph launch per two substeps of all 32 members, the members on 8 streams as
    concurrent graph branches, the solves of the members up to --ensemble-batch-cells cells as one batched
    multigrid) and plain (every member stepped eagerly, one after the other, as a single-grid run does)."""
    import torch

    from plume_amd.sim.device import init_device
    from plume_amd.sim.ensemble import Ensemble, ensemble_tables

    device = init_device("cuda")
    ens = Ensemble(ensemble_tables(args.ensemble_members, seed=20261204, frames=48), device,
                   batch_cells=args.ensemble_batch_cells or None)
    ens.extra_cycles = args.ensemble_extra_cycles
    ens.autotune()
    cells = ens.cells
    print(f"[bench] ensemble of {len(ens.members)}: {cells} cells", file=sys.stderr, flush=True)
    while ens.frames_done < 13:
        ens.run_frame(replay=False)
    ens.run_frame(replay=True)  # captures the graph
    torch.cuda.synchronize()
    print(f"[bench] batched multigrid tiles per depth: {ens.batch_tiles}", file=sys.stderr, flush=True)

    def timed(replay: bool) -> tuple[float, int]:
        steps = 0
        torch.cuda.synchronize()
        start = time.perf_counter()
        for _ in range(args.ensemble_frames):
            steps += ens.run_frame(replay=replay)
        torch.cuda.synchronize()
        return time.perf_counter() - start, steps

    batched_s, batched_steps = timed(True)
    print(f"[bench] batched: {batched_s:.3f} s for {batched_steps} substeps", file=sys.stderr, flush=True)
    plain_s, plain_steps = timed(False)
    print(f"[bench] plain: {plain_s:.3f} s for {plain_steps} substeps", file=sys.stderr, flush=True)
    residuals = [m.stepper.pressure_solve.last_relative_residual() for m in ens.members]
    batched = cells * batched_steps / batched_s
    plain = cells * plain_steps / plain_s
    sizes = {}
    for m in ens.members:
        key = "x".join(map(str, m.grid.cells))
        sizes[key] = sizes.get(key, 0) + 1
    return {"metric": "ensemble cell-substeps/s, 32 uneven smoke instances (batched hipGraph vs plain eager)",
            "members": len(ens.members), "cells": cells, "sizes": sizes,
            "batched": {"cell_substeps_per_s": batched, "seconds": batched_s, "substeps": batched_steps},
            "plain": {"cell_substeps_per_s": plain, "seconds": plain_s, "substeps": plain_steps},
            "speedup": batched / plain, "frames_timed": args.ensemble_frames,
            "extra_cycles": ens.extra_cycles, "frames_redone_eagerly": list(ens.redone),
            "batched_multigrid": {"members": len(ens.batched), "max_cells": args.ensemble_batch_cells or None,
                                  "tiles_per_depth": [shape for shape, _ in ens.batch_tiles]},
            "max_final_relative_residual": max(residuals),
            "data": "synthetic (seeded puff recipes; frames 13 onwards, after the emitters stop)"}


def adjoint_bench(args) -> dict:
    """Differentiable rollout: the puff at 64x96x64 from rest, 32 substeps, the (heat,
    thickness) images of 128x128 pixels after every 8 substeps (the state is stored only there), the L2 loss
    against the images of the same rollout with smoke, heat and lift sources 10% stronger; gradients with
    respect to the three per-cell source fields and alpha. Times the forward rollout alone and forward plus
    backward, and checks the gradient along each source's own direction and alpha against a central difference
    of the same FP32 rollout at h = 1e-3 (the FP64 comparison is tests/report_adjoint.py)."""
    import torch

    from plume_amd.sim.adjoint import Sources, rollout
    from plume_amd.sim.device import init_device
    from plume_amd.sim.scenes import differentiable_puff

    device = init_device("cuda")
    cells, substeps, checkpoint, alpha0 = (64, 96, 64), 32, 8, 0.1
    recipe, derived, sim, sources = differentiable_puff(device, cells, pixels=128, march_steps=192)
    sim.solve.autotune()
    with torch.no_grad():
        target = rollout(Sources(*(1.1 * t for t in sources.tensors())), torch.tensor(alpha0, dtype=torch.float64),
                         sim, substeps, checkpoint).clone()
    target64 = target.double()

    def loss_at(srcs, a) -> float:
        with torch.no_grad():
            images = rollout(Sources(*srcs), torch.tensor(a, dtype=torch.float64), sim, substeps, checkpoint)
            return float(((images.double() - target64) ** 2).sum())

    def forward_backward():
        params = [t.clone().requires_grad_(True) for t in sources.tensors()]
        alpha = torch.tensor(alpha0, dtype=torch.float64, requires_grad=True)
        images = rollout(Sources(*params), alpha, sim, substeps, checkpoint)
        ((images - target) ** 2).sum().backward()
        return params, alpha

    def timed(fn, repeats):
        fn()
        torch.cuda.synchronize()
        start = time.perf_counter()
        for _ in range(repeats):
            result = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - start) / repeats, result

    forward_s, _ = timed(lambda: loss_at(list(sources.tensors()), alpha0), args.adjoint_repeats)
    both_s, (params, alpha) = timed(forward_backward, args.adjoint_repeats)
    print(f"[bench] adjoint: forward {forward_s * 1e3:.1f} ms, forward + backward {both_s * 1e3:.1f} ms",
          file=sys.stderr, flush=True)
    h = 1e-3
    checks = {}
    base = list(sources.tensors())
    for k, name in enumerate(("smoke", "temperature", "lift", "alpha")):
        if k < 3:
            adj = float((params[k].grad.double() * base[k].double()).sum())
            plus = [t * (1 + h) if i == k else t for i, t in enumerate(base)]
            minus = [t * (1 - h) if i == k else t for i, t in enumerate(base)]
            fd = (loss_at(plus, alpha0) - loss_at(minus, alpha0)) / (2 * h)
        else:
            adj = float(alpha.grad) * alpha0
            fd = (loss_at(base, alpha0 * (1 + h)) - loss_at(base, alpha0 * (1 - h))) / (2 * h)
        checks[name] = {"adjoint": adj, "fp32_central_difference": fd, "relative_error": abs(adj - fd) / abs(fd)}
        print(f"[bench] d loss along {name}: adjoint {adj:.6e}, FD {fd:.6e}", file=sys.stderr, flush=True)
    n = cells[0] * cells[1] * cells[2]
    return {"metric": "differentiable rollout at 64x96x64: 32 substeps + 4 images of 128x128, forward and "
                      "forward+backward (gradients w.r.t. per-cell sources and alpha)",
            "cells": n, "substeps": substeps, "checkpoint": checkpoint, "pixels": 128, "march_steps": 192,
            "forward_ms": forward_s * 1e3, "forward_backward_ms": both_s * 1e3,
            "backward_over_forward": (both_s - forward_s) / forward_s,
            "forward_cell_substeps_per_s": n * substeps / forward_s,
            "gradient_cell_substeps_per_s": n * substeps / both_s,
            "gradient_check_h": h, "gradient_check": checks,
            "max_relative_error": max(c["relative_error"] for c in checks.values()),
            "data": "synthetic (seeded puff recipe, Gaussian sphere sources, from rest)"}


def main() -> int:
    parser = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=48, help="timed substeps (K)")
    parser.add_argument("--warmup", type=int, default=48, help="untimed substeps first (W); 48 = 12 frames")
    parser.add_argument("--cpu-steps", type=int, default=32, help="substeps of the CPU baseline sample (0 = skip)")
    parser.add_argument("--no-render", action="store_true")
    parser.add_argument("--no-recipe-run", action="store_true",
                        help="skip timing the smoke-puff recipe run end to end (80 x 100 x 80, 64 frames)")
    parser.add_argument("--tolerance", type=float, default=1e-5)
    parser.add_argument("--top-tile", type=int, default=-1, metavar="SHAPE",
                        help="pin the sweep tile shape of the 128x192x128 level instead of autotuning it")
    parser.add_argument("--tiles", default="", metavar="S0,S1,...",
                        help="pin the sweep tile shapes of the levels from the top down (128x192x128, 64x96x64, ...) "
                             "after autotuning, e.g. for a profiled run, whose per-dispatch timings the profiler "
                             "inflates")
    parser.add_argument("--ensemble", action="store_true",
                        help="run the 32-member batched ensemble benchmark instead and exit")
    parser.add_argument("--ensemble-members", type=int, default=32)
    parser.add_argument("--ensemble-frames", type=int, default=4)
    parser.add_argument("--ensemble-extra-cycles", type=int, default=1,
                        help="device-gated V-cycles each replayed solve runs beyond the most its member needed "
                             "lately (a frame whose solve stays open is re-run eagerly)")
    parser.add_argument("--ensemble-batch-cells", type=int, default=1_500_000,
                        help="members of at most this many cells share one batched multigrid (0: every member)")
    parser.add_argument("--mixed", action="store_true",
                        help="run the 256x384x256 mixed-precision projection benchmark instead and exit")
    parser.add_argument("--mixed-tolerance", type=float, default=1e-6)
    parser.add_argument("--mixed-sweeps", type=int, default=2,
                        help="smoothing sweeps per level and direction, the same for every form (2: the puff "
                             "recipe's V(2,2))")
    parser.add_argument("--adjoint", action="store_true",
                        help="run the 64x96x64 differentiable-rollout benchmark instead and exit")
    parser.add_argument("--adjoint-repeats", type=int, default=3)
    parser.add_argument("--no-additions", action="store_true",
                        help="leave the mixed-precision, ensemble and adjoint sub-benchmarks out of the line")
    parser.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                        help="only run the streaming-copy probe over MIB MiB (HBM counter calibration) and exit")
    args = parser.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; there is no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        return 2

    import torch

    if args.mixed:
        print(json.dumps(mixed_bench(args)), flush=True)
        return 0
    if args.ensemble:
        print(json.dumps(ensemble_bench(args)), flush=True)
        return 0
    if args.adjoint:
        print(json.dumps(adjoint_bench(args)), flush=True)
        return 0

    if args.copy_probe > 0:
        init_gpu = torch.cuda.is_available()
        if not init_gpu:
            raise SystemExit("the copy probe needs a GPU")
        print(json.dumps({"copy_probe_mib": args.copy_probe, "GB/s": copy_probe(args.copy_probe, repeats=4)}))
        return 0

    from plume_amd.recipe.derived import derive
    from plume_amd.recipe.load import fill_dependent_defaults, parse_recipe
    from plume_amd.render.camera import make_camera
    from plume_amd.render.frame_renderer import FrameRenderer, domain_view_params
    from plume_amd.sim.device import init_device
    from plume_amd.sim.grid import make_grid
    from plume_amd.sim.stepper import Stepper

    device = init_device("cuda")
    recipe = fill_dependent_defaults(parse_recipe(scene_table(tolerance=args.tolerance)), "bench")
    derived = derive(recipe)
    assert derived.cells == GRID, derived.cells
    cells = GRID[0] * GRID[1] * GRID[2]
    grid = make_grid(derived, recipe.domain.cell_m, recipe.domain.faces, recipe.domain.ambient_k, device)
    stepper = Stepper(grid, recipe)
    solve = stepper.pressure_solve
    dt = derived.dt_s
    print(f"[bench] grid {GRID}, dt {dt:.6f} s, warmup {args.warmup}, steps {args.steps}", file=sys.stderr,
          flush=True)

    tuned = solve.autotune()
    if args.top_tile >= 0:
        from plume_amd._lib import call

        call("plm_set_sweep_tile", (ctypes.c_int32 * 3)(*GRID), args.top_tile)
        tuned[GRID] = (args.top_tile, tuned.get(GRID, (None, []))[1])
    if args.tiles:
        from plume_amd._lib import call

        for level, shape in zip(solve.levels, (int(s) for s in args.tiles.split(","))):
            dims = tuple(int(d) for d in level.info.dims)
            call("plm_set_sweep_tile", (ctypes.c_int32 * 3)(*dims), shape)
            tuned[dims] = (shape, tuned.get(dims, tuned.get(level.cells, (None, [])))[1])
    print(f"[bench] sweep tiles: {tuned}", file=sys.stderr, flush=True)
    step_index = 0

    def run(n):
        nonlocal step_index
        for _ in range(n):
            stepper.step(step_index * dt, dt)
            step_index += 1

    run(args.warmup)
    torch.cuda.synchronize()
    print("[bench] warmup done", file=sys.stderr, flush=True)
    cycles_before = solve.total_cycles() if solve.tolerance > 0 else 0
    torch.cuda.synchronize()
    start = time.perf_counter()
    run(args.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - start
    cycles_after = solve.total_cycles() if solve.tolerance > 0 else 0
    mean_cycles = (cycles_after - cycles_before) / args.steps if solve.tolerance > 0 else float(solve.cycles)
    value = cells * args.steps / elapsed
    ms_per_step = 1000.0 * elapsed / args.steps
    residual = solve.last_relative_residual() if solve.tolerance > 0 else None
    from plume_amd.sim.step_cap import StepCap

    cfl_cells = StepCap(recipe).grid_speed(grid) * dt / CELL_M
    print(f"[bench] timed: {ms_per_step:.3f} ms/substep, mean cycles {mean_cycles:.2f}", file=sys.stderr,
          flush=True)

    # dominant kernel: the fused red-black sweep on the top level, timed per launch with HIP events over a
    # repeat of the timed region (same substeps, same state progression)
    sweep_ms, sweep_launches = time_kernel_family(lambda: run(min(args.steps, 12)), "rb_sweep_tile_top")
    achieved = sweep_bytes(cells) / (sweep_ms * 1e-3) / 1e9 if sweep_launches else float("nan")
    substep_gbps = substep_bytes(cells, mean_cycles) * args.steps / elapsed / 1e9

    result = {
        "metric": METRIC, "value": value, "unit": "cell-substeps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded smoke-puff scene, emitter on for frames 0-11)",
        "config": {"workload": "buoyant smoke puff substep, 128x192x128 FP32 MAC grid",
                   "grid": list(GRID), "cell_m": CELL_M, "dt_s": dt,
                   "pressure": f"multigrid V(2,2), relative residual <= {args.tolerance:g} (FP64 norm), "
                               f"mean {mean_cycles:.2f} cycles",
                   "confinement": 0.3, "buoyancy": "alpha 0.1, beta 1.0", "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": measured_traffic(),
                     "kernel": "k_rb_sweep_tile (128x192x128 level, 12 B/cell)",
                     "algorithmic_bytes_per_launch": sweep_bytes(cells),
                     "avg_launch_us": sweep_ms * 1e3, "launches": sweep_launches,
                     "copy_kernel_GBps": copy_probe(1024)},
        "substep_roofline": {"bytes_per_substep": substep_bytes(cells, mean_cycles), "achieved": substep_gbps,
                             "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": substep_gbps / HBM_PEAK_GBPS,
                             "mean_cycles": mean_cycles},
        "final_relative_residual": residual,
        "sweep_tiles": {"x".join(map(str, k)): v[0] for k, v in tuned.items()},
        "cfl_cells_per_step": cfl_cells,
    }

    if not args.no_render:
        settings = recipe.render
        view = type("View", (), {})()
        camera = make_camera(recipe.camera.elevation_deg, recipe.camera.azimuth_deg)
        # 256 x 256 px over the domain's 1.28 x 1.92 m projection, 192 samples per ray, 64-step light march
        width = height = 256
        span_x, span_y = GRID[0] * CELL_M, GRID[1] * CELL_M
        pixel_m = max(span_x, span_y) / 256
        view.camera, view.left_m, view.top_m, view.pixel_m = camera, -span_x / 2 - (256 * pixel_m - span_x) / 2, \
            span_y + (256 * pixel_m - span_y) / 2, pixel_m
        view.width_px, view.height_px = width, height
        renderer = FrameRenderer(view, settings, GRID, recipe.output.channels, device, max_steps=192, light_steps=64)
        for _ in range(2):
            renderer.prepare(grid.smoke.data, grid.info.cell_m)
            renderer.march(grid.smoke.data, grid.temperature.data, grid.info, renderer.view_params, renderer.image)
        torch.cuda.synchronize()
        frames = 10
        t0 = time.perf_counter()
        for _ in range(frames):
            renderer.prepare(grid.smoke.data, grid.info.cell_m)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(frames):
            renderer.march(grid.smoke.data, grid.temperature.data, grid.info, renderer.view_params, renderer.image)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        light_s, march_s = (t1 - t0) / frames, (t2 - t1) / frames
        result["raymarch"] = {
            "mpixel_samples_per_s": width * height * 192 / march_s / 1e6,
            "ms_per_frame_march": march_s * 1e3, "image": [height, width], "samples_per_ray": 192,
            "light_volume_ms": light_s * 1e3, "light_cell_samples_per_s": cells * 64 / light_s,
            "compulsory_hbm_frac": (12.0 * cells + 20.0 * width * height) / march_s / 1e9 / HBM_PEAK_GBPS,
        }

    result["frame_cache"] = frame_cache_bench(grid)
    if not args.no_recipe_run:
        result["recipe_run"] = recipe_run()

    if args.cpu_steps > 0:
        import dataclasses

        # all the logical CPUs the host gives this process (OMP_NUM_THREADS when it names fewer), and one
        threads = min(len(os.sched_getaffinity(0)), int(os.environ.get("OMP_NUM_THREADS", "0")) or 1 << 30)
        cpu_cycles = max(1, round(mean_cycles))  # the port runs a fixed count, as the reference does
        # a fixed cycle count, without the per-cycle residual checks of the tolerance mode
        cpu_recipe = dataclasses.replace(recipe, solver=dataclasses.replace(recipe.solver, multigrid_cycles=cpu_cycles,
                                                                            multigrid_tolerance=0.0))
        single_steps = max(2, args.cpu_steps // 8)
        timings = cpu_baseline(grid, cpu_recipe, derived, step_index * dt,
                               [(threads, args.cpu_steps), (1, single_steps)])
        many_s, _ = timings[threads]
        one_s, _ = timings[1]
        result["cpu_baseline"] = {
            "value": cells * args.cpu_steps / many_s, "unit": "cell-substeps/s", "cores": threads, "kind": "port",
            "sample": f"{args.cpu_steps} substeps ({single_steps} at 1 thread) of the same 128x192x128 scene from "
                      f"the GPU's warm state: the C port of the operators (oracle/plume_oracle.c) built with gcc -O3 "
                      f"-march=native -fopenmp, FP32, {cpu_cycles} V-cycles per substep",
            "seconds": many_s, "single_thread": {"value": cells * single_steps / one_s, "seconds": one_s,
                                                 "substeps": single_steps},
            "cpu": cpu_facts(),
            "reference_published": "Warp CPU 0.31 M cell-substeps/s on 4 cores (other hardware; context only)"}
    if not args.no_additions:
        for key, bench in (("mixed", mixed_bench), ("ensemble", ensemble_bench), ("adjoint", adjoint_bench)):
            print(f"[bench] addition: {key}", file=sys.stderr, flush=True)
            try:
                result[key] = bench(args)
            except Exception as error:  # the headline line stays; the failure is recorded in its place
                traceback.print_exc()
                result[key] = {"error": f"{type(error).__name__}: {error}"}
        if "cell_substeps_per_s" in result["ensemble"].get("batched", {}):
            result["ensemble"]["batched_vs_single_grid"] = result["ensemble"]["batched"]["cell_substeps_per_s"] / value

    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
