"""Error report of the mixed-precision pressure solve at 256 x 384 x 256 against an FP64 oracle solution.

The problem (plume_amd.sim.mixed.divergent_problem): the divergence of a seeded field of 64 Gaussian vortex
blobs plus a radial emitter source. The FP64 C oracle's multigrid solves it to a relative residual of 1e-13
on the host; its solution rounded to FP32 gives the FP32 floor (beside where plain FP32 V-cycles stall,
the bench's stand-in for it). Each form (plain FP32 multigrid, BF16 and FP16 storage with FP32 iterative
refinement) then solves to max(tolerance, 2 x floor) on the GPU, and the report gives, per form, the FP64
relative residual, the pressure error |q - q64| / |q64| and the post-projection divergence, beside the
oracle's own.

    python tests/report_mixed.py [--cells 256 384 256] [--json OUT]
"""

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE))


def main() -> None:
    parser = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    parser.add_argument("--cells", type=int, nargs=3, default=[256, 384, 256])
    parser.add_argument("--tolerance", type=float, default=1e-6)
    parser.add_argument("--oracle-cycles", type=int, default=40)
    parser.add_argument("--json", type=Path, default=None)
    args = parser.parse_args()
    os.environ.setdefault("OMP_NUM_THREADS", str(min(os.cpu_count() or 1, 16)))

    import torch

    from oracle import Oracle
    from oracle.sim import Multigrid as OracleMultigrid, compute_divergence as oracle_divergence
    from oracle.sim import make_grid as oracle_make_grid, poisson_residual
    from plume_amd.recipe.derived import derive
    from plume_amd.recipe.load import fill_dependent_defaults, parse_recipe
    from plume_amd.sim.device import init_device
    from plume_amd.sim.mixed import divergent_problem, fp32_floor, solve_forms
    from plume_amd.sim.multigrid import relative_residual
    from plume_amd.sim.scenes import puff_table

    cells = tuple(args.cells)
    device = init_device("cuda")
    problem = divergent_problem(cells, device)
    grid = problem.grid
    recipe = fill_dependent_defaults(parse_recipe(puff_table(cells)), "mixed")
    derived = derive(recipe)
    o = Oracle("f64")
    ogrid = oracle_make_grid(o, derived, recipe.domain.cell_m, recipe.domain.faces, recipe.domain.ambient_k)
    ogrid.divergence[...] = problem.divergence.cpu().numpy().astype(np.float64)
    solve64 = OracleMultigrid(ogrid, 1, 2)
    start = time.perf_counter()
    cycles64, residual64 = 0, 1.0
    for _ in range(args.oracle_cycles):
        solve64(ogrid)
        cycles64 += 1
        residual64 = poisson_residual(ogrid)
        print(f"[report] FP64 oracle cycle {cycles64}: relative residual {residual64:.3g}", flush=True)
        if residual64 <= 1e-13:
            break
    oracle_s = time.perf_counter() - start
    q64 = ogrid.pressure.copy()
    div0_norm = float(np.linalg.norm(ogrid.divergence))
    for f, v0 in zip(ogrid.velocity_fields(), problem.velocity):
        f.data[...] = v0.cpu().numpy().astype(np.float64)
    for axis, f in enumerate(ogrid.velocity_fields()):
        o.subtract_gradient(ogrid.info, f.data, q64, axis)
    post64 = float(np.linalg.norm(oracle_divergence(ogrid))) / div0_norm
    del ogrid, solve64

    rounded = torch.from_numpy(q64.astype(np.float32)).to(device)
    floor = relative_residual(grid.info, rounded, problem.divergence)
    estimate = fp32_floor(grid)
    target = max(args.tolerance, 2.0 * floor)
    print(f"[report] FP32 floor {floor:.4g} (plain FP32 stall {estimate:.4g}); target {target:.4g}", flush=True)
    forms = {}
    q64_norm = float(np.linalg.norm(q64))
    for name, (entry, q) in solve_forms(problem, target).items():
        entry["pressure_error_vs_fp64"] = float(np.linalg.norm(q.cpu().numpy().astype(np.float64) - q64)) / q64_norm
        forms[name] = entry
        print(f"[report] {name}: " + json.dumps(entry), flush=True)
    report = {"grid": list(cells), "tolerance": args.tolerance, "target": target,
              "fp64_oracle": {"v_cycles": cycles64, "relative_residual": residual64, "cpu_seconds": oracle_s,
                              "post_projection_divergence": post64, "threads": int(os.environ["OMP_NUM_THREADS"])},
              "fp32_floor": floor, "fp32_stall": estimate, "forms": forms}
    print(json.dumps(report), flush=True)
    if args.json is not None:
        args.json.write_text(json.dumps(report, indent=1))


if __name__ == "__main__":
    main()
