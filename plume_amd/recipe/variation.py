"""Recipe variations (recipe/variation.py of the reference): `extends`, a recipe written over a parent recipe,
and `--set key=value` overrides. Both act on the parsed TOML tables before the strict build, so a key they
bring in is checked like a key written in the file.

- extends: the parent's path is relative to the child and names a .toml file; the child's values go over the
  parent's, tables merging key by key and anything else (lists included, so [[emitter]]) replaced whole; a
  parent may extend its own parent, up to 8 deep, and a loop is an error.
- --set: a dotted key (`forces.drag_per_s`), an emitter named by name (`emitter.puff.radius_m`) or position
  (`emitter[0].radius_m`); the value is read as a TOML value and kept as text when it is not one
  (`domain.faces.top=closed`). Missing tables along the key are created; extends cannot be set.
"""

import copy
import re
from dataclasses import dataclass
from pathlib import Path, PurePosixPath, PureWindowsPath
from typing import Any

import tomli

from plume_amd.recipe.errors import RecipeError

MAX_EXTENDS_DEPTH = 8
EMITTER_AT = re.compile(r"^emitter\[(\d+)\]$")


@dataclass(frozen=True)
class RecipeOrigin:
    """Where a resolved recipe came from, for the settings log (recipe/variation.py:20-25)."""

    extends_chain: tuple[str, ...] = ()  # each parent's path as written, nearest first
    sets: tuple[str, ...] = ()  # the --set items, in order


def read_toml(path: Path) -> dict:
    """The parsed TOML at `path`; a missing, unreadable or invalid file is a RecipeError."""
    try:
        return tomli.loads(Path(path).read_bytes().decode("utf-8"))
    except FileNotFoundError:
        raise RecipeError(f"recipe not found: {path}") from None
    except (tomli.TOMLDecodeError, UnicodeDecodeError) as error:
        raise RecipeError(f"{path} is not valid TOML: {error}") from None
    except OSError as error:
        raise RecipeError(f"cannot read {path}: {error.strerror or error}") from None


def check_extends_path(text: str) -> None:
    """Parents are named relative to their child, so a recipe never carries a machine-local path."""
    windows = PureWindowsPath(text)
    if PurePosixPath(text).is_absolute() or windows.is_absolute() or windows.drive:
        raise RecipeError(f"extends must be a path relative to the recipe, got {text!r}")
    if not text.endswith(".toml"):
        raise RecipeError(f"extends must name a .toml recipe, got {text!r}")


def deep_merge(parent: dict, child: dict) -> dict:
    """The child's values over the parent's: tables key by key, anything else replaced whole."""
    merged = dict(parent)
    for key, value in child.items():
        below = merged.get(key)
        merged[key] = deep_merge(below, value) if isinstance(value, dict) and isinstance(below, dict) else value
    return merged


def resolve_extends(table: dict, path: Path) -> tuple[dict, tuple[str, ...]]:
    """The recipe at `path` (parsed as `table`) merged over its parents, with no `extends` key left, and the
    parents' paths as written, nearest first (recipe/variation.py:41-70)."""
    path = Path(path)
    merged = dict(table)
    chain: list[str] = []
    visited = {path.resolve()}
    current = path
    while True:
        parent_text = merged.pop("extends", "")
        if not isinstance(parent_text, str):
            raise RecipeError(f"extends in {current} must be text, got {type(parent_text).__name__}")
        if not parent_text:
            return merged, tuple(chain)
        check_extends_path(parent_text)
        if len(chain) == MAX_EXTENDS_DEPTH:
            raise RecipeError(f"extends goes deeper than {MAX_EXTENDS_DEPTH} parents, starting at {path}")
        parent_path = current.parent / parent_text
        if parent_path.resolve() in visited:
            raise RecipeError(f"extends forms a loop: {current} extends {parent_text}, which is already in the "
                              f"chain")
        visited.add(parent_path.resolve())
        chain.append(parent_text)
        # the parent's own extends survives the merge (the child's was popped) and is followed next
        merged = deep_merge(read_toml(parent_path), merged)
        current = parent_path


def parse_value(text: str) -> Any:
    """A --set value: the TOML value `text` spells, else `text` itself."""
    try:
        return tomli.loads(f"value = {text}")["value"]
    except tomli.TOMLDecodeError:
        return text


def set_item(item: str) -> tuple[str, Any]:
    """The key and value of one `key=value` item."""
    key, sep, value = item.partition("=")
    key = key.strip()
    if not sep or not key:
        raise RecipeError(f"--set needs key=value, got {item!r}")
    if key == "extends":
        raise RecipeError("--set cannot change extends; parents are merged before --set applies. Write extends "
                          "in the recipe instead")
    return key, parse_value(value.strip())


def _emitter(table: dict, parts: list[str], at: int, label: str) -> tuple[dict, int]:
    """The emitter table that parts[at:] names, and the index of the first key part after it."""
    emitters = table.get("emitter")
    if not isinstance(emitters, list) or not emitters:
        raise RecipeError(f"{label}: the recipe has no [[emitter]]")
    by_index = EMITTER_AT.match(parts[at])
    if by_index:
        position, after = int(by_index.group(1)), at + 1
        if position >= len(emitters):
            raise RecipeError(f"{label}: emitter[{position}] does not exist; the recipe has {len(emitters)}")
    else:
        if len(parts) < at + 3:
            raise RecipeError(f"{label}: name the emitter, as emitter.<name>.<key> or emitter[0].<key>")
        names = [e.get("name") if isinstance(e, dict) else None for e in emitters]
        if parts[at + 1] not in names:
            raise RecipeError(f"{label}: no emitter is named {parts[at + 1]!r}; emitters are "
                              f"{', '.join(str(n) for n in names)}")
        position, after = names.index(parts[at + 1]), at + 2
    if not isinstance(emitters[position], dict):
        raise RecipeError(f"{label}: emitter[{position}] is not a table")
    return emitters[position], after


def set_value(table: dict, key: str, value: Any, label: str) -> None:
    """Sets the dotted `key` of `table` to `value` in place, creating missing tables along the way."""
    parts = key.split(".")
    target, at = table, 0
    while at < len(parts) - 1:
        if parts[at] == "emitter" or EMITTER_AT.match(parts[at]):
            target, at = _emitter(table, parts, at, label)
            continue
        child = target.setdefault(parts[at], {})
        if not isinstance(child, dict):
            raise RecipeError(f"{label}: {'.'.join(parts[:at + 1])} is a value, not a table")
        target, at = child, at + 1
    target[parts[-1]] = value


def apply_sets(table: dict, sets: tuple[str, ...]) -> dict:
    """A copy of `table` with each `key=value` item applied in order; `table` is left as it was."""
    result = copy.deepcopy(table)
    for item in sets:
        key, value = set_item(item)
        set_value(result, key, value, f"--set {item!r}")
    return result


def apply_values(table: dict, values: dict[str, Any]) -> dict:
    """A copy of `table` with each dotted key set to its value, by the --set key rules."""
    result = copy.deepcopy(table)
    for key, value in values.items():
        set_value(result, key, copy.deepcopy(value), f"edit {key}")
    return result
