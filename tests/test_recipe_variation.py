"""Recipe variations (extends, --set) and the settings log, CPU (tests/recipe/test_variation.py of the
reference), plus round trips of the TOML writer."""

import math

import pytest
import tomli

from plume_amd.recipe.derived import derive
from plume_amd.recipe.errors import RecipeError
from plume_amd.recipe.load import load_recipe, load_recipe_with_origin, parse_recipe
from plume_amd.recipe.serialize import dumps, recipe_to_toml
from plume_amd.recipe.variation import apply_sets, apply_values, deep_merge

PARENT = """
seed = 4

[domain]
size_m = [2.0, 3.0, 2.0]
cell_m = 0.1
faces = { top = "open", left = "closed" }

[time]
duration_s = 2.0

[[emitter]]
name = "puff"
center_m = [0.0, 0.3, 0.0]
radius_m = 0.25
end_s = 0.25
smoke_per_s = 4.0
noise = { scale_m = 0.2, strength = 0.5 }

[dissipation]
smoke_per_s = 1.0
"""


def write(folder, name, text):
    folder.mkdir(parents=True, exist_ok=True)
    path = folder / name
    path.write_text(text)
    return path


def test_child_overrides_one_key_and_inherits_the_rest(tmp_path):
    write(tmp_path, "base.toml", PARENT)
    child = write(tmp_path, "child.toml", 'extends = "base.toml"\n[dissipation]\nsmoke_per_s = 2.5\n')
    recipe = load_recipe(child)
    assert recipe.dissipation.smoke_per_s == 2.5
    assert recipe.seed == 4 and recipe.emitter[0].smoke_per_s == 4.0
    assert recipe.extends == "" and recipe.output.name == "child"


def test_nested_tables_merge_key_by_key_and_lists_are_replaced(tmp_path):
    write(tmp_path, "base.toml", PARENT)
    child = write(tmp_path, "child.toml", 'extends = "base.toml"\n[domain.faces]\nright = "closed"\n')
    faces = load_recipe(child).domain.faces
    assert (faces.left, faces.right, faces.top) == ("closed", "closed", "open")
    other = write(tmp_path, "other.toml", 'extends = "base.toml"\n[[emitter]]\nname = "other"\n'
                                          'center_m = [0.0, 0.5, 0.0]\nradius_m = 0.1\nend_s = 0.5\n')
    recipe = load_recipe(other)
    assert [e.name for e in recipe.emitter] == ["other"] and recipe.emitter[0].smoke_per_s == 0.0
    assert deep_merge({"a": {"b": 1, "c": [1, 2]}}, {"a": {"c": [3]}}) == {"a": {"b": 1, "c": [3]}}


def test_grandparent_chain_resolves_relative_to_each_file(tmp_path):
    write(tmp_path / "shared", "base.toml", PARENT)
    write(tmp_path / "shared", "middle.toml", 'extends = "base.toml"\nseed = 9\n')
    child = write(tmp_path / "effects", "child.toml", 'extends = "../shared/middle.toml"\n[time]\nframes = 32\n')
    recipe, origin = load_recipe_with_origin(child)
    assert (recipe.seed, recipe.time.frames, recipe.time.duration_s) == (9, 32, 2.0)
    assert origin.extends_chain == ("../shared/middle.toml", "base.toml")


def test_loops_depth_and_bad_parents_fail(tmp_path):
    write(tmp_path, "a.toml", 'extends = "b.toml"\n')
    write(tmp_path, "b.toml", 'extends = "a.toml"\n')
    with pytest.raises(RecipeError, match="loop"):
        load_recipe(tmp_path / "a.toml")
    for index in range(10):
        write(tmp_path / "deep", f"r{index}.toml", f'extends = "r{index + 1}.toml"\n')
    write(tmp_path / "deep", "r10.toml", PARENT)
    with pytest.raises(RecipeError, match="deeper than 8"):
        load_recipe(tmp_path / "deep" / "r0.toml")
    assert load_recipe(tmp_path / "deep" / "r2.toml").seed == 4  # 8 parents is the limit
    with pytest.raises(RecipeError, match=".toml"):
        load_recipe(write(tmp_path, "c.toml", 'extends = "base.txt"\n'))
    with pytest.raises(RecipeError, match="missing.toml"):
        load_recipe(write(tmp_path, "d.toml", 'extends = "missing.toml"\n'))
    with pytest.raises(RecipeError, match="must be text"):
        load_recipe(write(tmp_path, "e.toml", "extends = 3\n"))


@pytest.mark.parametrize("parent", ["/etc/base.toml", "C:/recipes/base.toml", "C:base.toml",
                                    "\\\\server\\share\\base.toml"])
def test_an_absolute_parent_path_fails(tmp_path, parent):
    with pytest.raises(RecipeError, match="relative"):
        load_recipe(write(tmp_path, "child.toml", f"extends = '{parent}'\n"))


def test_set_overrides_numbers_vectors_text_and_emitters(tmp_path):
    path = write(tmp_path, "base.toml", PARENT)
    recipe = load_recipe(path, sets=("dissipation.cooling_per_s=0.5", "emitter.puff.velocity_mps=[0, 1.5, 0]",
                                     "emitter[0].radius_m = 0.4", "emitter.puff.noise.strength=0.9",
                                     "domain.faces.top=closed", "output.name=variant"))
    assert recipe.dissipation.cooling_per_s == 0.5
    assert recipe.emitter[0].velocity_mps == (0.0, 1.5, 0.0)
    assert recipe.emitter[0].radius_m == 0.4 and recipe.emitter[0].noise.strength == 0.9
    assert recipe.domain.faces.top == "closed" and recipe.output.name == "variant"
    child = write(tmp_path, "child.toml", 'extends = "base.toml"\nseed = 5\n')
    assert load_recipe(child, sets=("seed=6",)).seed == 6  # --set applies after extends
    with pytest.raises(RecipeError, match="did you mean cooling_per_s"):
        load_recipe(path, sets=("dissipation.cooling_per=0.5",))


@pytest.mark.parametrize("item, message", [
    ("seed", "key=value"),
    ("emitter.nobody.radius_m=1", "no emitter is named 'nobody'"),
    ("emitter[3].radius_m=1", r"emitter\[3\] does not exist"),
    ("emitter.puff=1", "name the emitter"),
    ("seed.x=1", "is a value, not a table"),
    ("extends=other.toml", "cannot change extends"),
])
def test_malformed_sets_fail(tmp_path, item, message):
    with pytest.raises(RecipeError, match=message):
        load_recipe(write(tmp_path, "base.toml", PARENT), sets=(item,))


def test_sets_and_edits_leave_the_parent_table_unchanged():
    table = tomli.loads(PARENT)
    changed = apply_sets(table, ("emitter.puff.noise.strength=0.1", "domain.faces.top=closed", "render.x.y=2"))
    assert table["emitter"][0]["noise"]["strength"] == 0.5 and table["domain"]["faces"]["top"] == "open"
    assert changed["emitter"][0]["noise"]["strength"] == 0.1 and changed["render"] == {"x": {"y": 2}}
    edited = apply_values(table, {"forces.drag_per_s": 0.25, "emitter[0].radius_m": 0.3})
    assert edited["forces"]["drag_per_s"] == 0.25 and edited["emitter"][0]["radius_m"] == 0.3
    assert "forces" not in table


def test_settings_log_of_a_variation_loads_back_to_the_same_recipe(tmp_path):
    write(tmp_path, "base.toml", PARENT)
    child = write(tmp_path, "child.toml", 'extends = "base.toml"\nseed = 5\n')
    recipe, origin = load_recipe_with_origin(child, sets=("forces.heat_buoyancy=2",))
    text = recipe_to_toml(recipe, derive(recipe), device="cuda:0", origin=origin,
                          run_facts={"total_steps": 12, "steps_per_frame": [3, 3]})
    table = tomli.loads(text)
    assert table["resolved"]["extends_chain"] == ["base.toml"]
    assert table["resolved"]["sets"] == ["forces.heat_buoyancy=2"]
    assert table["resolved"]["cells"] == list(derive(recipe).cells) and table["resolved"]["total_steps"] == 12
    assert parse_recipe(table) == recipe


def test_the_toml_writer_round_trips():
    table = {"title": 'quote " back \\ tab \t newline \n unicode é', "n": -3, "x": 1e-05, "big": 1.5e300,
             "flags": [True, False], "empty": [], "mixed": [1, 2.5, "a"], "dotted key": 1,
             "inline": {"a": [1, 2]}, "inf": math.inf,
             "table": {"k": 0.1, "sub": {"deep": "x"}},
             "emitter": [{"name": "a", "noise": {"s": 0.5}}, {"name": "b", "v": [0.0, 1.0]}]}
    back = tomli.loads(dumps(table))
    assert back == table
    assert math.copysign(1.0, tomli.loads(dumps({"z": -0.0}))["z"]) == -1.0
