"""A resolved recipe as TOML text (recipe/serialize.py of the reference): the settings log a run writes, every
value explicit, plus a [resolved] table of derived values that loading ignores.

The TOML writer here covers what a recipe holds: strings, booleans, integers, floats, inline arrays of those,
tables and arrays of tables ([[emitter]]). Loading its output gives back the same table.
"""

import dataclasses
import math
import re
from typing import Any

from plume_amd.recipe.variation import RecipeOrigin

BARE_KEY = re.compile(r"^[A-Za-z0-9_-]+$")


def _key(key: str) -> str:
    return key if BARE_KEY.match(key) else _string(key)


def _string(text: str) -> str:
    out = []
    for ch in text:
        if ch in '"\\':
            out.append("\\" + ch)
        elif ord(ch) < 0x20 or ord(ch) == 0x7F:
            out.append({"\n": "\\n", "\t": "\\t", "\r": "\\r"}.get(ch, f"\\u{ord(ch):04x}"))
        else:
            out.append(ch)
    return '"' + "".join(out) + '"'


def _value(value: Any) -> str:
    if isinstance(value, bool):
        return "true" if value else "false"
    if isinstance(value, int):
        return str(value)
    if isinstance(value, float):
        if math.isnan(value):
            return "nan"
        if math.isinf(value):
            return "inf" if value > 0 else "-inf"
        return repr(value)  # shortest round-trip form: 0.5, 1e-05, 2.0
    if isinstance(value, str):
        return _string(value)
    if isinstance(value, (list, tuple)):
        return "[" + ", ".join(_value(item) for item in value) + "]"
    if isinstance(value, dict):
        return "{" + ", ".join(f"{_key(k)} = {_value(v)}" for k, v in value.items()) + "}"
    raise TypeError(f"no TOML form for {type(value).__name__}")


def _is_table_array(value: Any) -> bool:
    return isinstance(value, (list, tuple)) and len(value) > 0 and all(isinstance(item, dict) for item in value)


def _body(table: dict, path: list[str], lines: list[str]) -> None:
    for key, value in table.items():
        if not isinstance(value, dict) and not _is_table_array(value):
            lines.append(f"{_key(key)} = {_value(value)}")
    for key, value in table.items():
        header = ".".join(_key(part) for part in path + [key])
        if isinstance(value, dict):
            lines += ["", f"[{header}]"]
            _body(value, path + [key], lines)
        elif _is_table_array(value):
            for item in value:
                lines += ["", f"[[{header}]]"]
                _body(item, path + [key], lines)


def dumps(table: dict) -> str:
    """TOML text of a table of recipe values."""
    lines: list[str] = []
    _body(table, [], lines)
    return "\n".join(lines).lstrip("\n") + "\n"


def _plain(value: Any) -> Any:
    if isinstance(value, dict):
        return {key: _plain(item) for key, item in value.items()}
    if isinstance(value, (list, tuple)):
        return [_plain(item) for item in value]
    return value


def recipe_table(recipe) -> dict:
    """The resolved recipe as a plain table (lists for vectors), every value explicit."""
    return _plain(dataclasses.asdict(recipe))


def recipe_to_toml(recipe, derived, device: str, origin: RecipeOrigin = RecipeOrigin(),
                   run_facts: dict | None = None) -> str:
    """The settings log (recipe/serialize.py:13-32): the recipe, then [resolved] with the derived values, the
    device, the parents and --set items already merged in, and `run_facts` (values only a run knows, such as
    the steps the CFL cap chose)."""
    table = recipe_table(recipe)
    table["resolved"] = {"cells": list(derived.cells), "size_m": list(derived.size_m), "fps": derived.fps,
                         "dt_s": derived.dt_s, "warmup_steps": derived.warmup_steps, "device": device,
                         "extends_chain": list(origin.extends_chain), "sets": list(origin.sets), **(run_facts or {})}
    return dumps(table)
