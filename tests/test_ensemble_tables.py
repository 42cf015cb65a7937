"""The ensemble's member recipes: sizes, emitter radii, buoyancy and confinement drawn
from a seed within their ranges. No GPU calls."""


def test_members_draw_their_own_parameters():
    from plume_amd.sim.ensemble import ENSEMBLE_SIZES, ensemble_tables

    tabs = ensemble_tables(32, seed=20261204)
    sizes = {tuple(round(s / t["domain"]["cell_m"]) for s in t["domain"]["size_m"]) for t in tabs}
    assert sizes <= set(ENSEMBLE_SIZES) and len(sizes) >= 3
    radii = [t["emitter"][0]["radius_m"] / t["domain"]["cell_m"] for t in tabs]
    assert 4.0 <= min(radii) and max(radii) <= 12.0
    alphas = [t["forces"]["smoke_weight"] * t["forces"]["gravity_mps2"] for t in tabs]
    assert 0.05 <= min(alphas) and max(alphas) <= 0.2
    eps = [t["forces"]["vorticity_confinement"] for t in tabs]
    assert 0.0 <= min(eps) and max(eps) <= 0.5


def test_members_spread_over_streams_by_cells():
    from plume_amd.sim.ensemble import balanced_streams

    cells = [100, 90, 80, 10, 10, 10, 5]
    streams = balanced_streams(cells, 3)
    load = [sum(c for c, s in zip(cells, streams) if s == k) for k in range(3)]
    assert sorted(load) == [100, 100, 105] or max(load) - min(load) <= max(cells[3:])
    assert balanced_streams([7], 4) == [0]
