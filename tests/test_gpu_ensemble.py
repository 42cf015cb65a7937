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


@pytest.mark.parametrize("extra_cycles, batch_cells", [(None, None), (-1000, None), (None, 20000)],
                         ids=["default-budget", "starved-budget", "larger-members-beside-the-batch"])
def test_graph_replay_matches_each_member_stepped_alone(device, extra_cycles, batch_cells):
    """With the default cycle budget and with a budget of one cycle per solve (every replayed frame then ends
    with an open gate and is run again eagerly from its snapshot), replayed frames equal eager ones; every
    member's solve in one batched multigrid, or only the members of at most 20000 cells in it and the larger
    ones stepped whole on their streams beside it."""
    from plume_amd.sim.ensemble import Ensemble

    tabs = tables()
    ens = Ensemble(tabs, device, streams=3, batch_cells=batch_cells)
    if batch_cells is not None:
        assert 0 < len(ens.batched) < len(ens.members)
    if extra_cycles is not None:
        ens.extra_cycles = extra_cycles
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
    if extra_cycles is None:
        assert replayed_frames >= 4, ens.log
    else:
        # one cycle per solve leaves the gate open (at least in the first quiet frames), and every such frame is
        # undone and run eagerly; the comparison above held after each of them
        assert ens.redone and replayed_frames + len(ens.redone) >= 4, (ens.log, ens.redone)
    assert any(member.grid.smoke.data.abs().sum().item() > 0 for member in ens.members)


def test_turbulent_members_are_never_replayed(device):
    """Turbulence moves with the time (sim/turbulence.py:59-67): a graph would freeze its pattern, so an
    ensemble with turbulence steps every frame eagerly and refuses to capture."""
    from plume_amd.sim.ensemble import Ensemble

    tabs = tables()[:2]
    tabs[1]["forces"]["turbulence"] = {"scale_m": 0.05, "strength_mps2": 0.2, "speed_mps": 0.1}
    ens = Ensemble(tabs, device, streams=2)
    alone = [alone_stepper(t, device) for t in tabs]
    for frame in range(5):
        steps = ens.run_frame(replay=True)
        dt = ens.frame_s / steps
        for grid, stepper in alone:
            for index in range(steps):
                stepper.step(frame * ens.frame_s + index * dt, dt)
    assert not any(replayed for _, _, replayed in ens.log)
    for member, (grid, _) in zip(ens.members, alone):
        got, expected = fields(member.grid), fields(grid)
        for name in got:
            assert np.array_equal(got[name].view(np.uint32), expected[name].view(np.uint32)), name
    with pytest.raises(ValueError, match="turbulence"):
        ens.capture(5 * ens.frame_s, ens.frame_s / 4)

