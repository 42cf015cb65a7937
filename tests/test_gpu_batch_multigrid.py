"""The batched multigrid (plm_mg_batch): the tolerance-mode solves of grids of uneven size advanced together, each
stage of a V-cycle one launch for all of them, leave every member's pressure, cycle count and residual exactly as
solving it alone does (sim/multigrid.py:124-147 per member, the tolerance gate per member)."""

import numpy as np
import pytest
import torch

from support import ALL_CLOSED, ALL_OPEN, DEFAULT_FACES, gpu_grid, rng

pytestmark = pytest.mark.gpu

CELL_M = 0.01
# uneven members: odd extents with overhanging coarse levels, every face kind, hierarchies whose LDS tails
# start at different depths (0 for the 16 x 24 x 16 grid, whose levels all fit in LDS, up to 2)
MEMBERS = [((45, 38, 70), DEFAULT_FACES), ((64, 48, 40), ALL_CLOSED), ((24, 24, 24), ALL_OPEN),
           ((33, 50, 29), DEFAULT_FACES), ((16, 24, 16), ALL_CLOSED), ((96, 64, 64), DEFAULT_FACES)]


def member(cells, faces, seed, cap, device):
    from plume_amd.sim.multigrid import Multigrid

    grid = gpu_grid(cells, CELL_M, faces, device=device)
    r = rng(seed)
    grid.divergence.copy_(torch.from_numpy(r.uniform(-1, 1, cells).astype(np.float32)))
    grid.pressure.copy_(torch.from_numpy(r.uniform(-1e-3, 1e-3, cells).astype(np.float32)))
    return grid, Multigrid(grid, cycles=cap, sweeps=2, tolerance=1e-5, max_cycles=cap)


def state(grid, solve):
    torch.cuda.synchronize()
    return (grid.pressure.cpu().numpy().view(np.uint32).copy(), int(solve.cycles_run[0].item()),
            float(solve.norm_scratch[0].item()), float(solve.norm_scratch[1].item()))


@pytest.mark.parametrize("batch_cycles", [30, 3], ids=["enough-cycles", "short-budget"])
def test_batched_solve_equals_each_member_solved_alone(device, batch_cycles):
    """Members with caps of 30 cycles (they meet 1e-5 first) and one with a cap of 2 (it stops there with its gate
    open); with a batch budget of 3 cycles the members that need more stop with their gates open, as a solve
    capped at 3 cycles does."""
    from plume_amd.sim.multigrid import BatchedMultigrid

    caps = [30] * len(MEMBERS)
    caps[3] = 2
    batched = [member(cells, faces, 40 + i, caps[i], device) for i, (cells, faces) in enumerate(MEMBERS)]
    alone = [member(cells, faces, 40 + i, min(caps[i], batch_cycles), device) for i, (cells, faces) in enumerate(MEMBERS)]
    batch = BatchedMultigrid([s for _, s in batched], [g for g, _ in batched])
    before = [g.pressure.clone() for g, _ in batched]
    tuned = batch.autotune(repeats=1)  # touches only the members' scratch levels
    assert len(tuned) == 2 and all(0 <= shape < len(times) for shape, times in tuned), tuned
    assert all(torch.equal(b, g.pressure) for b, (g, _) in zip(before, batched))
    for _ in range(2):  # a second solve from the first's pressure: the gates and counts restart
        batch(batch_cycles)
        for grid, solve in alone:
            solve(grid)
        for i, ((gb, sb), (ga, sa)) in enumerate(zip(batched, alone)):
            pb, cb, rb, gateb = state(gb, sb)
            pa, ca, ra, gatea = state(ga, sa)
            assert cb == ca and rb == ra and gateb == gatea, (MEMBERS[i], cb, ca, rb, ra, gateb, gatea)
            assert np.array_equal(pb, pa), MEMBERS[i]
    summary = [(MEMBERS[i][0], state(g, s)[1:]) for i, (g, s) in enumerate(batched)]
    if batch_cycles == 30:  # the members with an open face meet 1e-5 (a closed box's random rhs has no solution)
        assert all(state(*batched[i])[3] == 0.0 for i in (0, 2, 5)), summary
    assert state(*batched[3])[1] == 2, summary
    batch.close()


def test_a_batch_refuses_what_it_cannot_run(device):
    from plume_amd._lib import PlumeHipError
    from plume_amd.sim.multigrid import BatchedMultigrid, Multigrid

    grid = gpu_grid((24, 24, 24), CELL_M, ALL_OPEN, device=device)
    with pytest.raises(ValueError, match="tolerance"):
        BatchedMultigrid([Multigrid(grid, cycles=2, sweeps=2)], [grid])
    with pytest.raises(PlumeHipError, match="even"):
        BatchedMultigrid([Multigrid(grid, cycles=2, sweeps=3, tolerance=1e-5, max_cycles=4)], [grid])
