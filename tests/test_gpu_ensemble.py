"""The batched ensemble: members of different sizes advanced by one captured HIP graph per two substeps give,
frame after frame, exactly the fields each member gets when stepped alone and eagerly with the same substep
counts. Each frame replays the graph twice, with the host's CFL reductions (eager work and synchronisation)
between frames, and every frame is compared."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = ((16, 24, 16), (24, 24, 24), (32, 48, 32))


def tables():
    from plume_amd.sim.ensemble import ensemble_tables

    out = ensemble_tables(5, seed=11, frames=8, sizes=SIZES)
    for table in out:
        table["emitter"][0]["end_s"] = 0.1  # emitting in frames 0-2, quiet from frame 3 on
    return out


def alone_stepper(table, device):
    from plume_amd.recipe.derived import derive
    from plume_amd.recipe.load import fill_dependent_defaults, parse_recipe
    from plume_amd.sim.grid import make_grid
    from plume_amd.sim.stepper import Stepper

    recipe = fill_dependent_defaults(parse_recipe(table), "ensemble")
    derived = derive(recipe)
    grid = make_grid(derived, recipe.domain.cell_m, recipe.domain.faces, recipe.domain.ambient_k, device)
    return grid, Stepper(grid, recipe)


def fields(grid):
    return {name: getattr(grid, name).data.cpu().numpy() for name in ("u", "v", "w", "smoke", "temperature")} | {
        "pressure": grid.pressure.cpu().numpy()}


def test_graph_replay_matches_each_member_stepped_alone(device):
    from plume_amd.sim.ensemble import Ensemble

    tabs = tables()
    ens = Ensemble(tabs, device, streams=3)
    ens.extra_cycles = 4
    alone = [alone_stepper(t, device) for t in tabs]
    replayed_frames = 0
    for frame in range(8):
        steps = ens.run_frame(replay=True)
        _, _, replayed = ens.log[-1]
        replayed_frames += replayed
        dt = ens.frame_s / steps
        for grid, stepper in alone:
            for index in range(steps):
                stepper.step(frame * ens.frame_s + index * dt, dt)
        for member, (grid, _) in zip(ens.members, alone):
            got, expected = fields(member.grid), fields(grid)
            for name in got:
                assert np.array_equal(got[name].view(np.uint32), expected[name].view(np.uint32)), (frame, name,
                                                                                                    member.grid.cells)
    assert replayed_frames >= 4, ens.log
    assert any(member.grid.smoke.data.abs().sum().item() > 0 for member in ens.members)

