"""Helpers shared by tests: seeded fields, GPU/oracle grid pairs, comparisons and test recipes."""

import copy

import numpy as np

SEED_BASE = 20261203

ALL_CLOSED = ((0, 0, 0), (0, 0, 0))
DEFAULT_FACES = ((1, 0, 1), (1, 1, 1))  # floor closed, the rest open (recipe/schema.py:61-67 defaults)
ALL_OPEN = ((1, 1, 1), (1, 1, 1))


class Faces:
    """A FaceSettings-like object from open_lo / open_hi flags."""

    def __init__(self, open_lo, open_hi):
        kind = lambda flag: "open" if flag else "closed"  # noqa: E731
        self.left, self.bottom, self.back = (kind(f) for f in open_lo)
        self.right, self.top, self.front = (kind(f) for f in open_hi)


class SmallDerived:
    """A Derived-like object for grids built without a recipe (tests/support/grids.py of the reference)."""

    def __init__(self, cells, cell_m):
        size = tuple(n * cell_m for n in cells)
        self.cells = tuple(cells)
        self.size_m = size
        self.domain_min_m = (-size[0] / 2, 0.0, -size[2] / 2)


def rng(k=0):
    return np.random.default_rng(SEED_BASE + k)


def random_state(cells, k=0, ambient=None):
    """Seeded fields: face velocities U(-1, 1), smoke U(0, 1), temperature U(0, 1) (or ambient + U(0,1)*300)."""
    r = rng(k)
    nx, ny, nz = cells
    state = {
        "u": r.uniform(-1, 1, (nx + 1, ny, nz)),
        "v": r.uniform(-1, 1, (nx, ny + 1, nz)),
        "w": r.uniform(-1, 1, (nx, ny, nz + 1)),
        "smoke": r.uniform(0, 1, (nx, ny, nz)),
        "temperature": r.uniform(0, 1, (nx, ny, nz)),
    }
    if ambient is not None:
        state["temperature"] = ambient + 300.0 * state["temperature"]
    return {key: value.astype(np.float32) for key, value in state.items()}


def gpu_grid(cells, cell_m, faces, ambient=293.0, device="cuda:0"):
    from plume_amd.sim.grid import make_grid

    return make_grid(SmallDerived(cells, cell_m), cell_m, Faces(*faces), ambient, device)


def oracle_grid(o, cells, cell_m, faces, ambient=293.0):
    from oracle.sim import make_grid

    return make_grid(o, SmallDerived(cells, cell_m), cell_m, Faces(*faces), ambient)


def load_state(gpu, orc, state):
    import torch

    for name in ("u", "v", "w", "smoke", "temperature"):
        values = state[name]
        getattr(gpu, name).data.copy_(torch.from_numpy(values))
        getattr(orc, name).data[...] = values.astype(orc.o.dtype)


def to_np(t):
    return t.detach().cpu().numpy()


def compare(actual, expected, rel=1e-5, atol=0.0, what=""):
    """Asserts max |a - e| <= rel * max(|e|, tiny) + atol (the per-operator bar); returns mismatch stats."""
    actual = np.asarray(actual, dtype=np.float64)
    expected = np.asarray(expected, dtype=np.float64)
    assert actual.shape == expected.shape, f"{what}: shape {actual.shape} != {expected.shape}"
    assert np.all(np.isfinite(actual)), f"{what}: non-finite values"
    diff = np.abs(actual - expected)
    scale = max(float(np.abs(expected).max()) if expected.size else 0.0, 1e-30)
    worst = float(diff.max()) if diff.size else 0.0
    assert worst <= rel * scale + atol, f"{what}: max |diff| {worst:.3e} > {rel:g} * {scale:.3e} + {atol:g}"
    return {"max_diff": worst, "scale": scale, "differing": int(np.count_nonzero(diff))}


def bit_equal(actual, expected):
    a = np.ascontiguousarray(actual)
    e = np.ascontiguousarray(expected)
    return a.shape == e.shape and a.tobytes() == e.tobytes()


# ---------------------------------------------------------------------- recipes (tests/support/recipes.py)

BASE = {
    "seed": 1,
    "domain": {"size_m": [1.2, 1.2, 1.2], "cell_m": 0.05},
    "time": {"duration_s": 1.0, "frames": 4},
    "emitter": [{"name": "puff", "center_m": [0.0, 0.3, 0.0], "radius_m": 0.2, "end_s": 0.5, "smoke_per_s": 5.0}],
}


def recipe_table(**changes) -> dict:
    table = copy.deepcopy(BASE)
    for key, value in changes.items():
        if isinstance(value, dict) and isinstance(table.get(key), dict):
            table[key].update(value)
        else:
            table[key] = value
    return table


def tiny_recipe(**changes):
    from plume_amd.recipe.load import fill_dependent_defaults, parse_recipe

    return fill_dependent_defaults(parse_recipe(recipe_table(**changes)), "tiny")


TINY = """
seed = 3

[domain]
size_m = [1.0, 1.5, 1.0]
cell_m = 0.0625

[time]
duration_s = 1.0
frames = 4

[[emitter]]
name = "puff"
center_m = [0.0, 0.3, 0.0]
radius_m = 0.2
end_s = 0.3
smoke_per_s = 6.0
temperature_k = 500.0
noise = { scale_m = 0.1, strength = 0.5 }

[output]
frame_px = 64
"""
