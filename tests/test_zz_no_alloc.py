"""plm_smoke_budget_noalloc, the smoke budget for callers that cannot allocate: through the C ABI, its results
equal plm_smoke_budget's bit for bit on two calls in a row (the second on a smaller grid, so stale line sums
of the first call would show), and it rejects grids beyond its static scratch."""

import numpy as np
import pytest
import torch

from plume_amd._lib import LIB, PlumeHipError, call, int3, ptr, stream

pytestmark = pytest.mark.gpu


def _fields(shape, seed, device):
    rng = np.random.default_rng(seed)
    data = rng.uniform(0.0, 1.0, shape).astype(np.float32)
    hat = (data * rng.uniform(0.9, 1.1, shape)).astype(np.float32)
    # MacCormack overshoots the semi-Lagrangian total, so the budget scales `next` (factor != 1)
    nxt = (hat * np.float32(1.02) + rng.uniform(0.0, 0.01, shape)).astype(np.float32)
    return [torch.from_numpy(a).to(device) for a in (data, hat, nxt)]


def _allocating(shape, data, hat, nxt):
    lines = torch.zeros((3, shape[1], shape[2]), dtype=torch.float32, device=data.device)
    totals = torch.zeros(4, dtype=torch.float64, device=data.device)
    call("plm_smoke_budget", int3(shape), ptr(data), ptr(hat), ptr(nxt), ptr(lines), ptr(totals, torch.float64),
         stream())
    return totals


def _no_alloc(shape, data, hat, nxt, totals=None):
    call("plm_smoke_budget_noalloc", int3(shape), ptr(data), ptr(hat), ptr(nxt),
         ptr(totals, torch.float64) if totals is not None else None, stream())


def test_no_alloc_budget_equals_the_allocating_one_on_two_calls(device):
    for seed, shape in ((11, (40, 56, 48)), (12, (17, 23, 9))):
        data, hat, nxt = _fields(shape, seed, device)
        expected_next = nxt.clone()
        expected_totals = _allocating(shape, data, hat, expected_next)
        totals = torch.full((4,), -1.0, dtype=torch.float64, device=device)
        _no_alloc(shape, data, hat, nxt, totals)
        torch.cuda.synchronize()
        assert float(expected_totals[3]) != 1.0, "the case must exercise the scale"
        assert torch.equal(nxt, expected_next), shape
        assert torch.equal(totals, expected_totals), (shape, totals.tolist(), expected_totals.tolist())


def test_no_alloc_budget_without_a_totals_output(device):
    shape = (24, 32, 16)
    data, hat, nxt = _fields(shape, 13, device)
    expected_next = nxt.clone()
    _allocating(shape, data, hat, expected_next)
    _no_alloc(shape, data, hat, nxt)
    torch.cuda.synchronize()
    assert torch.equal(nxt, expected_next)


def test_no_alloc_budget_rejects_grids_beyond_its_scratch(device):
    limit = 262144  # PLM_BUDGET_NOALLOC_MAX_LINES
    shape = (1, 513, 512)
    assert shape[1] * shape[2] > limit
    data, hat, nxt = (torch.zeros(shape, dtype=torch.float32, device=device) for _ in range(3))
    rc = LIB.plm_smoke_budget_noalloc(int3(shape), ptr(data), ptr(hat), ptr(nxt), None, stream())
    assert rc == -1
    assert b"exceeds" in LIB.plm_last_error()
    with pytest.raises(PlumeHipError, match="static scratch"):
        _no_alloc(shape, data, hat, nxt)
    # the largest accepted plane still runs, and matches the allocating form
    shape = (2, 512, 512)
    data, hat, nxt = _fields(shape, 14, device)
    expected_next = nxt.clone()
    expected_totals = _allocating(shape, data, hat, expected_next)
    totals = torch.zeros(4, dtype=torch.float64, device=device)
    _no_alloc(shape, data, hat, nxt, totals)
    torch.cuda.synchronize()
    assert torch.equal(nxt, expected_next) and torch.equal(totals, expected_totals)
