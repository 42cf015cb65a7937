"""The frame loop's block schedule and step times (host logic of plume_amd.sim.run, no GPU), pinned by the
reference's frame-loop tests (tests/sim/test_step_cap.py:70-84: a 7-step warmup runs in blocks of 3, 3 and 1)
and by the block clock's definition (sim/run.py:62-73: block start = base steps before it × dt)."""

import numpy as np

from plume_amd.sim.run import StepLog, _schedule, _step_times


def test_warmup_runs_in_blocks_of_a_frames_steps_then_one_block_per_frame():
    blocks = list(_schedule(7, 3, 4))
    assert [(b.label, b.base) for b in blocks if b.frame is None] == [("warmup", 3), ("warmup", 3), ("warmup", 1)]
    assert [(b.label, b.base, b.frame) for b in blocks if b.frame is not None] == [
        (f"frame {i}", 3, i) for i in range(4)]
    assert [b.base for b in _schedule(6, 3, 1)] == [3, 3, 3]
    assert [b.base for b in _schedule(0, 4, 2)] == [4, 4]


def test_unsplit_blocks_keep_the_uncapped_clock():
    dt = 1.0 / 12.0
    times = _step_times(5, 3, 3, dt)
    assert times == [(5 * dt, dt), (6 * dt, dt), (7 * dt, dt)]


def test_split_blocks_share_the_block_evenly_and_end_where_it_ends():
    dt = 1.0 / 24.0
    times = _step_times(8, 4, 7, dt)
    assert len(times) == 7
    starts = np.array([t for t, _ in times])
    lengths = np.array([length for _, length in times])
    assert starts[0] == 8 * dt
    assert np.all(lengths == 4 * dt / 7)
    np.testing.assert_allclose(np.diff(starts), 4 * dt / 7, rtol=1e-12)
    assert abs(starts[-1] + lengths[-1] - 12 * dt) < 1e-15
    # the next block starts from its own base count, not from this sum
    assert _step_times(12, 4, 4, dt)[0][0] == 12 * dt


def test_step_log_counts_added_steps():
    log = StepLog(warmup_steps=8, steps_per_frame=(3, 5, 3), base_steps_per_frame=3, base_warmup_steps=7)
    assert log.total_steps == 19
    assert log.added_steps == 19 - (7 + 9)
