"""The recipe schema: every key a recipe may hold, with its default (recipe/schema.py:44-246 of the reference).

The simulation, render, camera, output and preview keys and their defaults are the reference's, so its
recipes (recipes/smoke_puff.toml and its test recipes) load unchanged. Two solver keys are new here:
`multigrid_tolerance` and `multigrid_max_cycles`, the converge-to-a-residual mode the reference lacks.
The contact-sheet `look` table is accepted and kept but not used by the simulation or the march.
"""

import dataclasses
from dataclasses import dataclass
from typing import Any

Vec3 = tuple[float, float, float]
ColorStop = tuple[float, str]

REQUIRED = dataclasses.MISSING
FACE_KINDS = ("closed", "open")
SEED_MAX = 2**31 - 1
SHADE_LAYOUT = "heat,thickness,shade,unused"
NORMAL_LAYOUT = "heat,thickness,normal_x,normal_y"
CHANNEL_LAYOUTS = (SHADE_LAYOUT, NORMAL_LAYOUT)


def setting(default: Any, unit: str = "", doc: str = "", choices: tuple[str, ...] = ()) -> Any:
    if isinstance(default, (list, dict)):
        raise TypeError("use tuples for sequence defaults so settings stay immutable")
    return dataclasses.field(default=default, metadata={"unit": unit, "doc": doc, "choices": choices})


def section(doc: str = "") -> Any:
    return dataclasses.field(default=None, metadata={"doc": doc, "section": True, "choices": ()})


@dataclass(frozen=True, kw_only=True)
class FaceSettings:
    bottom: str = setting("closed", "", "Floor of the domain (y minimum).", FACE_KINDS)
    top: str = setting("open", "", "Ceiling of the domain (y maximum).", FACE_KINDS)
    left: str = setting("open", "", "x minimum face.", FACE_KINDS)
    right: str = setting("open", "", "x maximum face.", FACE_KINDS)
    back: str = setting("open", "", "z minimum face.", FACE_KINDS)
    front: str = setting("open", "", "z maximum face.", FACE_KINDS)


@dataclass(frozen=True, kw_only=True)
class DomainSettings:
    size_m: Vec3 = setting(REQUIRED, "m", "Width, height and depth of the box; origin at the floor's centre.")
    cell_m: float = setting(REQUIRED, "m", "Edge length of one grid cell.")
    ambient_k: float = setting(293.0, "K", "Temperature of the surrounding air.")
    faces: FaceSettings = section("Each face is `closed` (a wall) or `open`.")


@dataclass(frozen=True, kw_only=True)
class TimeSettings:
    duration_s: float = setting(REQUIRED, "s", "Effect time covered by the flipbook.")
    frames: int = setting(64, "", "Number of flipbook frames.")
    substeps: int = setting(3, "", "Simulation steps per frame.")
    warmup_s: float = setting(0.0, "s", "Simulated time before frame 0.")


@dataclass(frozen=True, kw_only=True)
class NoiseSettings:
    scale_m: float = setting(0.1, "m", "Size of the emitter's breakup pattern.")
    strength: float = setting(0.0, "", "How strongly the pattern breaks up the emitter; 0 for none.")


@dataclass(frozen=True, kw_only=True)
class EmitterSettings:
    name: str = setting(REQUIRED, "", "Name.")
    shape: str = setting("sphere", "", "Emitter shape.", ("sphere", "box"))
    center_m: Vec3 = setting(REQUIRED, "m", "Center of the shape.")
    radius_m: float = setting(0.0, "m", "Radius of a sphere.")
    size_m: Vec3 = setting((0.0, 0.0, 0.0), "m", "Width, height and depth of a box.")
    edge_softness: float = setting(0.3, "", "Outer fraction of the shape over which emission fades to 0.")
    start_s: float = setting(0.0, "s", "Effect time the emitter starts.")
    end_s: float = setting(REQUIRED, "s", "Effect time the emitter stops.")
    smoke_per_s: float = setting(0.0, "1/s", "Smoke density added per second at full coverage.")
    temperature_k: float = setting(-1.0, "K", "Temperature the emitter holds its cells at; -1 means ambient.")
    velocity_mps: Vec3 = setting((0.0, 0.0, 0.0), "m/s", "Velocity the emitter holds its cells at.")
    velocity_mode: str = setting("fixed", "", "`fixed` or `outward`.", ("fixed", "outward"))
    noise: NoiseSettings = section("Seeded breakup of the emitter's shape.")


@dataclass(frozen=True, kw_only=True)
class TurbulenceSettings:
    scale_m: float = setting(0.3, "m", "Size of the swirls.")
    strength_mps2: float = setting(0.0, "m/s^2", "Acceleration of the swirling push; 0 for none.")
    speed_mps: float = setting(0.2, "m/s", "How fast the swirl pattern changes.")


@dataclass(frozen=True, kw_only=True)
class ForceSettings:
    gravity_mps2: float = setting(9.81, "m/s^2", "Strength of gravity.")
    heat_buoyancy: float = setting(1.0, "", "Lift from heat; 1 is ideal-gas buoyancy.")
    smoke_weight: float = setting(0.0, "", "Sink from smoke density, as a fraction of gravity per density.")
    wind_mps: Vec3 = setting((0.0, 0.0, 0.0), "m/s", "Velocity of the surrounding air (needs drag).")
    drag_per_s: float = setting(0.0, "1/s", "Rate the velocity relaxes toward the wind; 0 for none.")
    vorticity_confinement: float = setting(0.0, "1/s", "Strength of vorticity confinement; 0 for none.")
    turbulence: TurbulenceSettings = section("Seeded curl-noise swirls.")


@dataclass(frozen=True, kw_only=True)
class DissipationSettings:
    smoke_per_s: float = setting(0.0, "1/s", "Rate smoke fades.")
    cooling_per_s: float = setting(0.0, "1/s", "Rate temperature relaxes toward ambient.")


@dataclass(frozen=True, kw_only=True)
class SolverSettings:
    pressure_solver: str = setting("multigrid", "", "`multigrid` or `gauss_seidel`.", ("multigrid", "gauss_seidel"))
    pressure_iterations: int = setting(60, "", "Red-black sweeps per solve, for `gauss_seidel`.")
    multigrid_cycles: int = setting(2, "", "Multigrid V-cycles per pressure solve.")
    multigrid_sweeps: int = setting(2, "", "Smoothing sweeps before and after each level's correction.")
    multigrid_tolerance: float = setting(0.0, "", "Relative residual |rhs - Aq| / |rhs| (FP64) at which the "
                                                  "multigrid stops early; 0 runs exactly multigrid_cycles.")
    multigrid_max_cycles: int = setting(16, "", "Most V-cycles per solve when multigrid_tolerance > 0.")
    max_cells_per_step: float = setting(1.0, "cells", "CFL cap: cells the fastest gas may cross per step.")
    max_substeps: int = setting(32, "", "Most steps the CFL cap may give one frame.")


@dataclass(frozen=True, kw_only=True)
class CameraSettings:
    elevation_deg: float = setting(0.0, "deg", "How far the camera looks down.")
    azimuth_deg: float = setting(0.0, "deg", "Orbit around the vertical axis.")
    seat: str = setting("", "", "Emitter whose centre is the seat point; empty means the first.")


@dataclass(frozen=True, kw_only=True)
class RenderSettings:
    px_per_cell: float = setting(2.0, "", "Pixels per cell of the domain view.")
    opacity_per_density: float = setting(4.0, "1/m", "Opacity per unit smoke density per meter of ray.")
    glow_start_k: float = setting(600.0, "K", "Temperature where heat starts to glow.")
    smoke_hides_heat: float = setting(1.0, "", "How much smoke in front hides glow behind it, 0 to 1.")
    normal_blur_cells: float = setting(1.5, "cells", "Blur applied to smoke before taking surface normals.")
    normal_fade_thickness: float = setting(0.1, "", "Below this thickness, normals fade to face the camera.")
    light_dir: Vec3 = setting((-0.5, 0.7, 0.5), "", "Direction toward the key light in camera space.")
    shade_fill: float = setting(0.25, "", "Share of light every bit of smoke gets, 0 to 1.")
    shade_fade_thickness: float = setting(0.1, "", "Sheet shading: thin-pixel fade.")
    shade_neighborhood_px: float = setting(4.0, "px", "Sheet shading: neighbourhood sigma.")
    soften_px: float = setting(0.75, "px", "Sheet: Gaussian blur on the final frames.")
    crop_margin: float = setting(0.04, "", "Sheet: empty margin around the content.")


@dataclass(frozen=True, kw_only=True)
class OutputSettings:
    name: str = setting("", "", "Output name; empty means the recipe file's name.")
    frame_px: int = setting(256, "px", "Width and height of one frame.")
    columns: int = setting(0, "", "Frames per sheet row; 0 means the smallest square grid.")
    channels: str = setting(SHADE_LAYOUT, "", "Channel layout.", CHANNEL_LAYOUTS)


@dataclass(frozen=True, kw_only=True)
class PreviewSettings:
    longest_axis_cells: int = setting(48, "cells", "Cells along the longest axis in --preview.")
    frames: int = setting(16, "", "Frames in --preview.")
    frame_px: int = setting(128, "px", "Frame size in --preview.")


@dataclass(frozen=True, kw_only=True)
class BackgroundSettings:
    neutral: str = setting("#808080", "sRGB", "Background of contact_neutral.png.")
    grass: str = setting("#aac476", "sRGB", "Background of contact_grass.png.")


@dataclass(frozen=True, kw_only=True)
class LookSettings:
    ambient: float = setting(0.65)
    shade_depth: float = setting(0.75)
    smoke_ramp: tuple[ColorStop, ...] = setting(
        ((0.0, "#aea193"), (0.4, "#bc9c84"), (0.75, "#dca07a"), (1.0, "#f5b47e")))
    emission_ramp: tuple[ColorStop, ...] = setting(
        ((0.0, "#000000"), (0.2, "#3a0c04"), (0.5, "#c4441a"), (0.8, "#ff9a3a"), (1.0, "#fff2c2")))
    contact_px: int = setting(128, "px")
    backgrounds: BackgroundSettings = section("Contact sheet backgrounds.")


@dataclass(frozen=True, kw_only=True)
class Recipe:
    extends: str = setting("", "", "Parent recipe (resolved before the table is read).")
    seed: int = setting(0, "", f"Seed for every random choice, 0 to {SEED_MAX}.")
    domain: DomainSettings = section("The simulation box.")
    time: TimeSettings = section("Effect time and frames.")
    emitter: tuple[EmitterSettings, ...] = setting(REQUIRED, "", "Sources of smoke, heat and motion.")
    forces: ForceSettings = section("Forces on the gas.")
    dissipation: DissipationSettings = section("How smoke and heat fade.")
    solver: SolverSettings = section("Solver controls.")
    camera: CameraSettings = section("The orthographic camera.")
    render: RenderSettings = section("Turning the volume into channels.")
    output: OutputSettings = section("The sheet's size and name.")
    preview: PreviewSettings = section("Overrides used by --preview.")
    look: LookSettings = section("Contact sheet colours.")
