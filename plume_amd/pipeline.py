"""One recipe run through the simulation and render path to the sheet frames, the numbers file and the
settings log (cli/run.py:43-130 of the reference, without its PNG, JSON and contact-sheet outputs): load the
recipe (its extends chain, --set items and optionally its --preview form), simulate on the HIP path capturing a
raw frame after each frame's substeps, assemble the sheet frames, and measure them."""

import time
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from plume_amd.recipe.derived import Derived, derive
from plume_amd.recipe.load import load_recipe_with_origin
from plume_amd.recipe.schema import Recipe
from plume_amd.recipe.serialize import recipe_to_toml
from plume_amd.render.camera import sheet_domain_view
from plume_amd.render.frame_renderer import FrameRenderer
from plume_amd.report.numbers import RunFacts, measure, numbers_text
from plume_amd.sheet.assemble import AssembledFrames, assemble, seat_point
from plume_amd.sim.device import init_device
from plume_amd.sim.run import StepLog, simulate


@dataclass(frozen=True)
class RunResult:
    recipe: Recipe
    derived: Derived
    device: str
    raw_frames: list[np.ndarray]
    assembled: AssembledFrames
    numbers: str
    steps: StepLog
    seconds: dict[str, float]  # wall time of "simulate" (render excluded), "render" and "assemble"


def run_facts(recipe: Recipe, derived: Derived, backend: str, assembled: AssembledFrames, source: str) -> RunFacts:
    """The numbers file's summary facts (cli/run.py:122-125)."""
    nx, ny, nz = derived.cells
    return RunFacts(recipe=source, backend=backend, grid=f"{nx} x {ny} x {nz} cells of {recipe.domain.cell_m:.6f} m",
                    fps=derived.fps, frame_m=assembled.frame_m, heat_divisor=assembled.heat_divisor,
                    channels=recipe.output.channels, shade_divisor=assembled.shade_divisor)


def sheet_numbers(raw_frames: list[np.ndarray], recipe: Recipe, derived: Derived, backend: str,
                  source: str) -> tuple[AssembledFrames, str]:
    """The assembled sheet frames of a run's raw frames and its numbers file text."""
    view = sheet_domain_view(recipe, derived)
    assembled = assemble(raw_frames, view, seat_point(recipe), recipe.output.frame_px, recipe.render,
                         recipe.output.channels)
    return assembled, numbers_text(measure(assembled.frames, derived.fps), run_facts(recipe, derived, backend,
                                                                                      assembled, source))


def run_recipe(recipe_path: Path, preview: bool = False, device: str = "cuda", source: str | None = None,
               out: Path | None = None, sets: tuple[str, ...] = ()) -> RunResult:
    """Runs a recipe (with its extends chain and `--set` items) on the HIP path and returns its sheet frames and
    numbers; with `out`, also writes `numbers.txt` and the settings log `settings.toml` there. `source` is the
    recipe path the numbers file names (default: as given)."""
    import torch

    recipe, origin = load_recipe_with_origin(Path(recipe_path), preview=preview, sets=tuple(sets))
    derived = derive(recipe)
    device_name = init_device(device)
    view = sheet_domain_view(recipe, derived)
    renderer = FrameRenderer(view, recipe.render, derived.cells, recipe.output.channels, device_name)
    raw_frames = []
    seconds = {"simulate": 0.0, "render": 0.0}

    def capture(frame: int, grid) -> None:
        torch.cuda.synchronize()
        start = time.perf_counter()
        raw_frames.append(renderer.render(grid))
        seconds["render"] += time.perf_counter() - start

    start = time.perf_counter()
    steps = simulate(recipe, derived, device_name, capture)
    torch.cuda.synchronize()
    seconds["simulate"] = time.perf_counter() - start - seconds["render"]
    start = time.perf_counter()
    assembled, numbers = sheet_numbers(raw_frames, recipe, derived, device_name,
                                       source if source is not None else str(recipe_path))
    seconds["assemble"] = time.perf_counter() - start
    if out is not None:
        Path(out).mkdir(parents=True, exist_ok=True)
        (Path(out) / "numbers.txt").write_text(numbers, newline="\n")
        facts = {"warmup_steps_run": steps.warmup_steps, "steps_per_frame": list(steps.steps_per_frame),
                 "total_steps": steps.total_steps}
        (Path(out) / "settings.toml").write_text(recipe_to_toml(recipe, derived, device_name, origin, facts),
                                                 newline="\n")
    return RunResult(recipe, derived, device_name, raw_frames, assembled, numbers, steps, seconds)


def main(argv: list[str] | None = None) -> int:
    """`python -m plume_amd.pipeline RECIPE [--preview] [--set key=value ...] [--out DIR]`: runs the recipe on
    the HIP path, prints its numbers file and stage times, and with --out writes numbers.txt and settings.toml."""
    import argparse

    parser = argparse.ArgumentParser(prog="python -m plume_amd.pipeline", description=main.__doc__)
    parser.add_argument("recipe", type=Path)
    parser.add_argument("--preview", action="store_true")
    parser.add_argument("--set", dest="sets", action="append", default=[], metavar="KEY=VALUE")
    parser.add_argument("--out", type=Path, default=None)
    args = parser.parse_args(argv)
    result = run_recipe(args.recipe, preview=args.preview, device="cuda", out=args.out, sets=tuple(args.sets))
    print(result.numbers, end="")
    print(f"\nsubsteps {result.steps.total_steps}, simulate {result.seconds['simulate']:.3f} s, render "
          f"{result.seconds['render']:.3f} s, assemble {result.seconds['assemble']:.3f} s")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
