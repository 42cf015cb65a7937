"""One recipe run through the simulation and render path to the sheet, its JSON, the numbers file and the
settings log (cli/run.py:43-130 of the reference, without its contact sheets and channel strip): load the
recipe (its extends chain, --set items and optionally its --preview form), simulate on the HIP path capturing a
raw frame after each frame's substeps, assemble the sheet frames, measure them, and write the files."""

import sys
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
from plume_amd.sheet.outputs import SheetFacts, recipe_source, sheet_image, sheet_json, write_png
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
    files: dict[str, Path]  # what was written: "sheet", "json", "numbers", "settings" (empty without `out`)


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
               out: Path | None = None, sets: tuple[str, ...] = (), command: str | None = None) -> RunResult:
    """Runs a recipe (with its extends chain and `--set` items) on the HIP path and returns its sheet frames and
    numbers. With `out` it also writes, into `out` (into `out`/preview for a preview, as cli/run.py:82-86 does):
    the sheet `<name>.png`, its `<name>.json`, `numbers.txt` and the settings log `settings.toml`. `source` is
    the recipe path the numbers file and the JSON name (default: relative to its git repository's root);
    `command` the command line the JSON records as written_by (default: the `plume run` line of this call)."""
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
    source = source if source is not None else recipe_source(Path(recipe_path))
    assembled, numbers = sheet_numbers(raw_frames, recipe, derived, device_name, source)
    seconds["assemble"] = time.perf_counter() - start
    files: dict[str, Path] = {}
    if out is not None:
        folder = Path(out) / "preview" if preview else Path(out)
        folder.mkdir(parents=True, exist_ok=True)
        name = recipe.output.name
        files = {"sheet": folder / f"{name}.png", "json": folder / f"{name}.json",
                 "numbers": folder / "numbers.txt", "settings": folder / "settings.toml"}
        write_png(files["sheet"], sheet_image(assembled.frames, recipe.output.columns))
        if command is None:
            command = " ".join(["plume", "run", str(recipe_path), *(["--preview"] if preview else []),
                                *(f"--set {item}" for item in sets), "--out", str(out)])
        facts = SheetFacts(frames=len(assembled.frames), columns=recipe.output.columns,
                           frame_px=recipe.output.frame_px, fps=derived.fps, frame_m=assembled.frame_m,
                           seat=assembled.seat, channels=recipe.output.channels, source=source, written_by=command)
        files["json"].write_text(sheet_json(facts), newline="\n")
        files["numbers"].write_text(numbers, newline="\n")
        steps_facts = {"warmup_steps_run": steps.warmup_steps, "steps_per_frame": list(steps.steps_per_frame),
                       "total_steps": steps.total_steps}
        files["settings"].write_text(recipe_to_toml(recipe, derived, device_name, origin, steps_facts),
                                     newline="\n")
    return RunResult(recipe, derived, device_name, raw_frames, assembled, numbers, steps, seconds, files)


def main(argv: list[str] | None = None) -> int:
    """`python -m plume_amd.pipeline RECIPE [--preview] [--set key=value ...] [--out DIR]`: runs the recipe on
    the HIP path, prints its numbers file and stage times, and with --out writes the sheet PNG and JSON,
    numbers.txt and settings.toml."""
    import argparse

    parser = argparse.ArgumentParser(prog="python -m plume_amd.pipeline", description=main.__doc__)
    parser.add_argument("recipe", type=Path)
    parser.add_argument("--preview", action="store_true")
    parser.add_argument("--set", dest="sets", action="append", default=[], metavar="KEY=VALUE")
    parser.add_argument("--out", type=Path, default=None)
    args = parser.parse_args(argv)
    result = run_recipe(args.recipe, preview=args.preview, device="cuda", out=args.out, sets=tuple(args.sets),
                        command=" ".join(["plume", "run", *(argv if argv is not None else sys.argv[1:])]))
    print(result.numbers, end="")
    print(f"\nsubsteps {result.steps.total_steps}, simulate {result.seconds['simulate']:.3f} s, render "
          f"{result.seconds['render']:.3f} s, assemble {result.seconds['assemble']:.3f} s")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
