"""The mixed-precision pressure solve (BF16 / FP16 multigrid storage, FP32 iterative refinement, FP64 norms and
coarsest level) against the FP64 oracle: it reaches the target relative residual, and its pressure and
post-projection divergence match the FP64 solution as closely as the plain FP32 solve at the same residual.

The target is max(1e-6, 2 x the FP32 floor), the floor being the relative residual of the FP64 oracle's
solution rounded to FP32: no solve that keeps the pressure in FP32 gets below it (1.8e-6 on this grid)."""

import numpy as np
import pytest

from support import ALL_CLOSED, ALL_OPEN, DEFAULT_FACES, gpu_grid, oracle_grid

pytestmark = pytest.mark.gpu

CELLS = (48, 64, 40)
TOLERANCE = 1e-6


def to_np(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("storage", ["bf16", "f16"])
def test_storage_conversion_rounds_to_nearest_even(device, storage):
    import torch

    from plume_amd._lib import call, ptr, stream
    from plume_amd.sim.mixed import DTYPES, STORAGE

    g = torch.Generator().manual_seed(5)
    x = (torch.randn(100_003, generator=g) * 3.0).to(device)
    out = torch.empty(x.shape, dtype=DTYPES[storage], device=device)
    for scale in (1.0, 0.25, 8.0):
        call("plm_to_storage", x.numel(), ptr(x), scale, STORAGE[storage], out.data_ptr(), stream())
        expected = (x * scale).to(DTYPES[storage])
        assert torch.equal(out.view(torch.int16), expected.view(torch.int16)), scale
    y = torch.randn(x.shape, generator=g).to(device)
    expected = y + out.float() * 0.125
    call("plm_add_storage", y.numel(), ptr(y), out.data_ptr(), STORAGE[storage], 0.125, stream())
    assert torch.equal(y, expected)


@pytest.mark.parametrize("storage", ["f32", "bf16", "f16"])
@pytest.mark.parametrize("level", [0, 1])
def test_fused_refinement_step_equals_the_unfused_sequence(device, oracle64, storage, level):
    """plm_refine_residual gives, bit for bit, x + e * inv_scale (plm_add_storage), the FP32 residual of that x
    (plm_residual), its scaled copy in the storage type (plm_to_storage) and the FP64 norms
    (plm_residual_norms): on the unit-coupling simulation level and on a coarser level (couplings other than
    1, ghost gaps). Without a correction it writes no x; x_out overlapping x is refused."""
    import ctypes

    import torch

    from plume_amd._lib import LIB, PlumeHipError, call, ptr, stream
    from plume_amd.sim.mixed import DTYPES, STORAGE
    from plume_amd.sim.multigrid import build_levels

    gpu, _ = divergent_pair(device, oracle64)
    info = build_levels(gpu, 8)[level].info
    cells = tuple(info.dims)
    code, dtype = STORAGE[storage], DTYPES[storage]
    bits = torch.int32 if storage == "f32" else torch.int16
    g = torch.Generator().manual_seed(11 + level)
    x = torch.randn(cells, generator=g).to(device)
    b = (torch.randn(cells, generator=g) * 100.0).to(device)
    e = (torch.randn(cells, generator=g) * 4.0).to(dtype).to(device)
    scratch = torch.zeros(LIB.plm_norm_scratch_doubles(ctypes.byref(info)), dtype=torch.float64, device=device)
    norms, expected_norms = (torch.zeros(2, dtype=torch.float64, device=device) for _ in range(2))
    x_out = torch.full_like(x, float("nan"))
    r_out = torch.empty(cells, dtype=dtype, device=device)
    call("plm_refine_residual", ctypes.byref(info), ptr(x), e.data_ptr(), 0.25, code, ptr(b), ptr(x_out),
         r_out.data_ptr(), 8.0, ptr(scratch, torch.float64), ptr(norms, torch.float64), stream())

    y = x.clone()
    call("plm_add_storage", y.numel(), ptr(y), e.data_ptr(), code, 0.25, stream())
    r = torch.empty_like(x)
    call("plm_residual", ctypes.byref(info), ptr(y), ptr(b), ptr(r), stream())
    stored = torch.empty_like(r_out)
    call("plm_to_storage", r.numel(), ptr(r), 8.0, code, stored.data_ptr(), stream())
    call("plm_residual_norms", ctypes.byref(info), ptr(y), ptr(b), ptr(scratch, torch.float64),
         ptr(expected_norms, torch.float64), stream())
    assert torch.equal(x_out.view(torch.int32), y.view(torch.int32))
    assert torch.equal(r_out.view(bits), stored.view(bits))
    assert torch.equal(norms, expected_norms)
    # the norms themselves against FP64 sums of the same FP32 residual
    r64, b64 = r.double(), b.double()
    want = torch.stack([(r64 * r64).sum(), (b64 * b64).sum()])
    assert torch.allclose(norms, want, rtol=1e-12, atol=0.0)

    # no correction: the residual of x itself, nothing written to x
    call("plm_refine_residual", ctypes.byref(info), ptr(y), None, 0.0, code, ptr(b), None, r_out.data_ptr(), 8.0,
         ptr(scratch, torch.float64), ptr(norms, torch.float64), stream())
    assert torch.equal(r_out.view(bits), stored.view(bits))
    assert torch.equal(norms, expected_norms)
    with pytest.raises(PlumeHipError, match="overlap"):
        call("plm_refine_residual", ctypes.byref(info), ptr(y), e.data_ptr(), 1.0, code, ptr(b), ptr(y), None,
             1.0, ptr(scratch, torch.float64), ptr(norms, torch.float64), stream())


PACKED_CASES = [((48, 64, 40), DEFAULT_FACES), ((36, 36, 128), ALL_CLOSED), ((26, 34, 68), ALL_OPEN),
                ((30, 20, 64), DEFAULT_FACES)]


@pytest.mark.parametrize("cells, faces", PACKED_CASES, ids=["48x64x40", "36x36x128-closed", "26x34x68-open",
                                                            "30x20x64"])
@pytest.mark.parametrize("sweeps, zero_start", [(1, 0), (2, 0), (2, 1)], ids=["one", "two", "two-from-zero"])
@pytest.mark.parametrize("scale", [1.0, 1e-4], ids=["unit", "small"])
@pytest.mark.parametrize("level", [0, 1])
def test_packed_fp16_sweeps_match_the_fp64_sweeps(device, oracle64, cells, faces, sweeps, zero_start, scale, level):
    """plm_rb_sweeps_lp in packed FP16 arithmetic (PLM_STORAGE_F16_PACKED) against the FP64 oracle's
    red_black_sweeps (sim/pressure.py:80-92, relaxation 1) from the same FP16 values: within 16 roundings to FP16
    (11 significant bits, and 2^-25 absolute where the values are subnormal) of the largest value, as is the
    FP32-arithmetic FP16 sweep. A wrong neighbour, parity, coupling or cell_m² is off by the values themselves.
    The grids hold inner tiles (all eight rows and columns around them in the level, 64 whole cells along k)
    and boundary tiles, k extents of 40 to 128 cells (partial and whole 64-cell tiles) and every face kind; the
    "small" scale puts cell_m² rhs / t1 among FP16's subnormals, as in the correction solve. Level 1 is the
    grid's first coarser level (flat, open faces coupling by 1 / (0.5 + ghost_gap), 2 cm cells; 13 x 17 x 34
    sweeps unpacked, its k extent not a multiple of 4)."""
    import ctypes

    import torch

    from oracle.sim import level_infos
    from plume_amd._lib import call, stream
    from plume_amd.sim.mixed import STORAGE
    from plume_amd.sim.multigrid import build_levels

    cm = 0.01
    gpu = gpu_grid(cells, cm, faces, device=device)
    orc = oracle_grid(oracle64, cells, cm, faces)
    info = build_levels(gpu, 8)[level].info
    cells, orc_info = level_infos(oracle64, cells, cm, orc.origin_m, orc.open_lo, orc.open_hi)[level]
    assert tuple(info.dims) == tuple(cells)
    r = np.random.default_rng(20261203 + 41)
    q0 = (r.uniform(-1, 1, cells) * scale).astype(np.float16)
    rhs = (r.uniform(-1, 1, cells) * (2.0 * scale / cm**2)).astype(np.float16)  # cell_m² rhs up to 2 |q|
    expected = np.zeros(cells) if zero_start else q0.astype(np.float64)
    oracle64.rb_sweeps(orc_info, expected, rhs.astype(np.float64), sweeps, 1.0)
    # cell_m² rhs up to 2 |q| on level 0; (2 cell_m)² rhs up to 8 |q| on level 1
    bound = 16.0 * (2.0 ** -11 * float(np.abs(expected).max()) + 2.0 ** -25)
    results = {}
    rhs_d = torch.from_numpy(rhs).to(device)
    for storage in ("f16", "f16x2"):
        q = torch.from_numpy(q0).to(device)
        scratch = torch.full_like(q, float("nan"))
        call("plm_rb_sweeps_lp", ctypes.byref(info), STORAGE[storage], q.data_ptr(), rhs_d.data_ptr(),
             scratch.data_ptr(), sweeps, zero_start, stream())
        got = q.cpu().numpy().astype(np.float64)
        assert np.all(np.isfinite(got)), storage
        results[storage] = float(np.abs(got - expected).max())
    print(cells, faces, sweeps, zero_start, scale, level, "max |diff| vs FP64:", results, "bound", bound)
    for storage, worst in results.items():
        assert worst <= bound, (storage, worst, bound)


@pytest.mark.parametrize("cells, faces", PACKED_CASES, ids=["48x64x40", "36x36x128-closed", "26x34x68-open",
                                                            "30x20x64"])
@pytest.mark.parametrize("level", [0, 1])
def test_packed_transfers_equal_the_fp16_kernels(device, oracle64, cells, faces, level):
    """The packed level's residual + restrict and prolongation (PLM_STORAGE_F16_PACKED: dword rows, each coarse
    k's interpolation shared by the fine cells that read it) give the FP16 kernels' values bit for bit, and
    those match the FP64 oracle's residual_kernel + restrict_kernel and prolongation_kernel
    (sim/multigrid.py:36-80) within FP16 rounding. Level 1 of 30 x 20 x 64 has an odd i extent (15); level 1
    of 26 x 34 x 68 is not packed (k extent 34) and takes the FP16 kernels either way."""
    import ctypes

    import torch

    from oracle.sim import level_infos
    from plume_amd._lib import call, stream
    from plume_amd.sim.mixed import STORAGE
    from plume_amd.sim.multigrid import build_levels

    cm = 0.01
    gpu = gpu_grid(cells, cm, faces, device=device)
    orc = oracle_grid(oracle64, cells, cm, faces)
    info = build_levels(gpu, 8)[level].info
    fine, orc_info = level_infos(oracle64, cells, cm, orc.origin_m, orc.open_lo, orc.open_hi)[level]
    coarse = tuple((n + 1) // 2 for n in fine)
    r = np.random.default_rng(20261203 + 43)
    h2 = (cm * 2**level) ** 2  # the level's cell_m²: q of that size keeps A q / cell_m² near 1
    q = (r.uniform(-1, 1, fine) * h2).astype(np.float16)
    rhs = (r.uniform(-1, 1, fine) * 10.0).astype(np.float16)
    correction = (r.uniform(-1, 1, coarse) * h2).astype(np.float16)
    restricted, prolonged = {}, {}
    q_d, rhs_d, correction_d = (torch.from_numpy(a).to(device) for a in (q, rhs, correction))
    for storage in ("f16", "f16x2"):
        out = torch.full(coarse, float("nan"), dtype=torch.float16, device=device)
        call("plm_residual_restrict_lp", ctypes.byref(info), STORAGE[storage], q_d.data_ptr(), rhs_d.data_ptr(),
             out.data_ptr(), stream())
        restricted[storage] = out.cpu().numpy()
        target = q_d.clone()
        call("plm_prolong_add_lp", ctypes.byref(info), STORAGE[storage], correction_d.data_ptr(), target.data_ptr(),
             stream())
        prolonged[storage] = target.cpu().numpy()
    assert np.array_equal(restricted["f16"].view(np.uint16), restricted["f16x2"].view(np.uint16))
    assert np.array_equal(prolonged["f16"].view(np.uint16), prolonged["f16x2"].view(np.uint16))
    residual = np.zeros(fine)
    oracle64.residual(orc_info, q.astype(np.float64), rhs.astype(np.float64), residual)
    expected = np.zeros(coarse)
    oracle64.restrict(residual, expected)
    got = restricted["f16x2"].astype(np.float64)
    assert np.abs(got - expected).max() <= 2.0 ** -9 * np.abs(expected).max(), np.abs(got - expected).max()
    expected_q = q.astype(np.float64)
    oracle64.prolong_add(correction.astype(np.float64), expected_q)
    assert np.abs(prolonged["f16x2"] - expected_q).max() <= 2.0 ** -10 * np.abs(expected_q).max()


def divergent_pair(device, oracle64):
    from oracle.sim import compute_divergence as oracle_divergence
    from plume_amd.sim.mixed import divergent_velocity
    from plume_amd.sim.pressure import compute_divergence, zero_closed_faces

    gpu = gpu_grid(CELLS, 0.01, DEFAULT_FACES, device=device)
    divergent_velocity(gpu, seed=3, blobs=16, emitter_radius_cells=4.0)
    zero_closed_faces(gpu)
    compute_divergence(gpu)
    orc = oracle_grid(oracle64, CELLS, 0.01, DEFAULT_FACES)
    for f, g in zip(orc.velocity_fields(), gpu.velocity_fields()):
        f.data[...] = to_np(g.data).astype(np.float64)
    oracle_divergence(orc)
    orc.divergence[...] = to_np(gpu.divergence).astype(np.float64)  # the same FP32 right-hand side
    return gpu, orc


def fp64_solution(orc):
    from oracle.sim import Multigrid, poisson_residual

    solve = Multigrid(orc, 1, 2)
    for _ in range(60):
        solve(orc)
        if poisson_residual(orc) < 1e-13:
            break
    return orc.pressure.copy()


def post_projection_divergence(gpu, velocity0):
    import torch

    from plume_amd.sim.pressure import compute_divergence, subtract_gradient

    rhs = gpu.divergence.clone()
    for f, v0 in zip(gpu.velocity_fields(), velocity0):
        f.data.copy_(v0)
    subtract_gradient(gpu)
    compute_divergence(gpu)
    ratio = float(torch.linalg.vector_norm(gpu.divergence.double()) / torch.linalg.vector_norm(rhs.double()))
    gpu.divergence.copy_(rhs)
    return ratio


def fp32_floor(gpu, q64) -> float:
    import torch

    from plume_amd.sim.multigrid import relative_residual

    gpu.pressure.copy_(torch.from_numpy(q64.astype(np.float32)))
    return relative_residual(gpu.info, gpu.pressure, gpu.divergence)


def test_mixed_precision_solves_match_the_fp64_solution(device, oracle64):
    from plume_amd.sim.mixed import MixedPrecisionMultigrid
    from plume_amd.sim.multigrid import Multigrid, relative_residual

    from plume_amd.sim.mixed import fp32_floor as stall_floor

    gpu, orc = divergent_pair(device, oracle64)
    q64 = fp64_solution(orc)
    floor = fp32_floor(gpu, q64)
    assert 1e-7 < floor < 1e-5, floor
    # where plain FP32 V-cycles stall (the bench's floor) sits at the FP64 solution's FP32 floor
    stall = stall_floor(gpu)
    print("FP32 floor", floor, "plain FP32 stall", stall)
    assert 0.5 * floor < stall < 2.0 * floor, (floor, stall)
    target = max(TOLERANCE, 2.0 * floor)
    velocity0 = [f.data.clone() for f in gpu.velocity_fields()]
    errors, divergences = {}, {}
    for storage in ("f32-plain", "f32", "bf16", "f16", "f16x2"):
        gpu.pressure.zero_()
        if storage == "f32-plain":
            solve = Multigrid(gpu, 2, 2, tolerance=target, max_cycles=60)
        else:
            solve = MixedPrecisionMultigrid(gpu, storage, tolerance=target)
        solve(gpu)
        rel = relative_residual(gpu.info, gpu.pressure, gpu.divergence)
        assert rel <= target, (storage, rel, target)
        if storage != "f32-plain":
            assert solve.log.residuals[-1] == pytest.approx(rel, rel=1e-6)
            assert solve.log.rounds <= 10 and not solve.log.stalled, (storage, solve.log.residuals)
        q = to_np(gpu.pressure).astype(np.float64)
        errors[storage] = np.linalg.norm(q - q64) / np.linalg.norm(q64)
        divergences[storage] = post_projection_divergence(gpu, velocity0)
    print("pressure error vs FP64:", errors, "post-projection divergence:", divergences)
    for storage, err in errors.items():
        # a 1e-6 residual bounds the error through the operator's conditioning; every form lands within a
        # small factor of the plain FP32 solve at the same residual
        assert err <= max(1e-4, 10.0 * errors["f32-plain"]), (storage, errors)
        assert divergences[storage] <= max(1e-4, 10.0 * divergences["f32-plain"]), (storage, divergences)


def test_refinement_continues_from_a_warm_start(device, oracle64):
    """A second solve from the converged pressure does no refinement round."""
    from plume_amd.sim.mixed import MixedPrecisionMultigrid

    gpu, _ = divergent_pair(device, oracle64)
    solve = MixedPrecisionMultigrid(gpu, "bf16", tolerance=1e-5)
    solve(gpu)
    first = solve.log.rounds
    solve(gpu)
    assert first >= 1 and solve.log.rounds == 0


def test_refinement_stops_at_the_fp32_floor(device, oracle64):
    """Asked for less than the FP32 pressure can represent, the solve stops once a round no longer halves the
    residual, at the floor's level, instead of running every round."""
    from plume_amd.sim.mixed import MixedPrecisionMultigrid

    gpu, orc = divergent_pair(device, oracle64)
    floor = fp32_floor(gpu, fp64_solution(orc))
    gpu.pressure.zero_()
    solve = MixedPrecisionMultigrid(gpu, "bf16", tolerance=1e-9, max_rounds=40)
    solve(gpu)
    assert solve.log.stalled and solve.log.rounds < 15, solve.log.residuals
    assert solve.log.residuals[-1] < 3.0 * floor
