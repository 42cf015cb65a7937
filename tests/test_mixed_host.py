"""Host logic of the mixed-precision refinement (no GPU): the stall threshold and the power-of-two scale."""

import math

from plume_amd.sim.mixed import STALL, _pow2_scale, stall_threshold


def test_stall_threshold_keeps_half_for_strong_cycles_and_loosens_for_weak_ones():
    # a V(2,2)-like first round (ratio ~0.1-0.3) keeps the default bar
    for ratio in (0.05, 0.1, 0.27, 1.0 / 3.0):
        assert stall_threshold(ratio) == STALL
    # a weak first round (V(1,1) on BF16 cut the residual to 0.65) must not read as stalled
    assert stall_threshold(0.65) > 0.65
    assert math.isclose(stall_threshold(0.6), 0.9)
    # capped: a round that no longer reduces the residual always stalls
    assert stall_threshold(0.8) == 0.95 and stall_threshold(2.0) == 0.95
    ratios = [0.01 * k for k in range(1, 150)]
    bars = [stall_threshold(r) for r in ratios]
    assert all(a <= b for a, b in zip(bars, bars[1:]))


def test_pow2_scale_brings_the_rms_near_one_exactly():
    for rms in (1e-12, 3.7e-6, 0.01, 1.0, 42.0, 6.5e4):
        scale = _pow2_scale(rms)
        assert math.frexp(scale)[0] == 0.5  # a power of two
        assert 2.0 ** -0.5 <= rms * scale <= 2.0 ** 0.5
