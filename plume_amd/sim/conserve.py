"""Keeping clamped MacCormack from adding smoke (sim/conserve.py of the reference).

After each smoke advection the MacCormack total is held between the plain semi-Lagrangian total and the total
before the step, by scaling the whole field by one number. The reference copies three line-sum planes to the
host every substep for this; here the line sums, their FP64 totals, the factor and the conditional scale all
run on the device (plm_smoke_budget), so the substep has no host round trip.
"""

import torch

from plume_amd._lib import call, int3, ptr, stream


def held_total(before: float, semi_lagrangian: float, maccormack: float) -> float:
    """MacCormack's total held to the scheme's own bounds (sim/conserve.py:33-35)."""
    return max(semi_lagrangian, min(maccormack, before))


class SmokeBudget:
    """Scratch for one grid's smoke totals (sim/conserve.py:38-43)."""

    def __init__(self, grid):
        self.shape = tuple(grid.cells)
        self.lines = torch.zeros((3, grid.cells[1], grid.cells[2]), dtype=torch.float32, device=grid.device)
        # before, semi-Lagrangian, MacCormack, factor (FP64, on the device)
        self.totals = torch.zeros(4, dtype=torch.float64, device=grid.device)

    def hold(self, grid) -> None:
        """Call after advecting smoke and before swapping: rescales `grid.smoke.next` into its bounds
        (sim/conserve.py:45-52)."""
        smoke = grid.smoke
        call("plm_smoke_budget", int3(self.shape), ptr(smoke.data), ptr(smoke.hat), ptr(smoke.next), ptr(self.lines),
             ptr(self.totals, torch.float64), stream())

    @property
    def last_factor(self) -> float:
        """The last hold's factor, target / MacCormack total (synchronises)."""
        return float(self.totals[3].item())

    @property
    def last_totals(self) -> tuple[float, float, float]:
        """The last hold's before, semi-Lagrangian and MacCormack totals (synchronises)."""
        before, semi_lagrangian, maccormack, _ = self.totals.tolist()
        return before, semi_lagrangian, maccormack
