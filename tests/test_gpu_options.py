"""Option semantics on the GPU.

* `schurCache` is honoured only where the reference honours it (joint-focal.ts:102-108: the
  shared-focal route, at most 16 cameras, M (6C + 1) <= 262,144); on the fixed route and outside
  the guard it is validated and ignored (bundle.ts:825-826), so the solve returns the matrix-free
  bits. Inside the guard the bundle takes the direct route (`directSolve`'s bits).
* The mixed form with a refinement tolerance looser than the PCG tolerance: a step is accepted
  when its FP64 residual meets the tolerance its solution was certified at.
* A batch keeps each bundle's own PCG budget (maxCgIterations), so default options on windows of
  different camera counts give each window's single solve, bit for bit.
"""
import numpy as np
import pytest

from oracle import oracle as O
from typecv_amd import bundle as B
from typecv_amd import synthetic as S

pytestmark = pytest.mark.gpu


def test_schur_cache_is_ignored_on_the_fixed_route(device):
    print("\n[options] schurCache, fixed route", flush=True)
    model = S.bundle_fixture()
    plain = B.bundle_adjust_gpu(device, model, {"iterations": 20})
    cached = B.bundle_adjust_gpu(device, model, {"iterations": 20, "schurCache": True})
    assert cached["execution"]["report"]["direct_solves"] == 0
    for a, b in zip(plain["reconstruction"]["views"], cached["reconstruction"]["views"]):
        assert np.array_equal(a["pose"]["R"], b["pose"]["R"])
        assert np.array_equal(a["pose"]["t"], b["pose"]["t"])
    for a, b in zip(plain["reconstruction"]["landmarks"], cached["reconstruction"]["landmarks"]):
        assert np.array_equal(a["xyz"], b["xyz"])
    assert plain["diagnostics"] == cached["diagnostics"]


def test_schur_cache_outside_the_guard_returns_the_matrix_free_bits(device):
    print("\n[options] schurCache, 20-camera shared scene", flush=True)
    packed, _ = S.trajectory_scene(cameras=20, points=3000, seed=2)
    kw = dict(iterations=10, max_cg=241, shared_focal=True, huber=1.0)
    plain, _ = B.solve_packed(device, B.make_control(packed.C, packed.P, packed.M, **kw), packed)
    cached, report = B.solve_packed(
        device, B.make_control(packed.C, packed.P, packed.M, schur_cache=True, **kw), packed)
    assert report["direct_solves"] == 0
    assert np.array_equal(plain, cached)


def test_schur_cache_inside_the_guard_takes_the_direct_route(device):
    print("\n[options] schurCache, 8-camera shared scene inside the guard", flush=True)
    packed, _ = S.trajectory_scene(cameras=8, points=300, seed=3, track=(3, 8))
    assert packed.M * (6 * packed.C + 1) <= 262144, packed.M
    kw = dict(iterations=10, max_cg=97, shared_focal=True, huber=1.0)
    cached, report = B.solve_packed(
        device, B.make_control(packed.C, packed.P, packed.M, schur_cache=True, **kw), packed)
    direct, _ = B.solve_packed(
        device, B.make_control(packed.C, packed.P, packed.M, direct=True, **kw), packed)
    assert report["direct_solves"] > 0
    assert np.array_equal(cached, direct)
    # the same option on the fixed route of the same scene: ignored
    fixed_kw = dict(iterations=10, max_cg=96, huber=1.0)
    plain, _ = B.solve_packed(device, B.make_control(packed.C, packed.P, packed.M, **fixed_kw),
                              packed)
    ignored, report = B.solve_packed(
        device, B.make_control(packed.C, packed.P, packed.M, schur_cache=True, **fixed_kw), packed)
    assert report["direct_solves"] == 0
    assert np.array_equal(plain, ignored)


def test_mixed_refinement_looser_than_the_pcg_tolerance(device):
    """refinementTolerance 1e-6 with cgRelativeTolerance 1e-10: PCG on the FP32 operator runs to
    1e-10 of its own recurrence; the FP64 refinement certifies each solution at 1e-6, and those
    steps are accepted (the gate is the certified tolerance). Final cost within 1e-6 relative of the
    FP64 oracle at the same options."""
    print("\n[options] mixed, refinement 1e-6, PCG 1e-10", flush=True)
    packed, _ = S.trajectory_scene(cameras=20, points=3000, seed=2)
    kw = dict(iterations=15, max_cg=241, shared_focal=True, huber=1.0, tolerance=1e-10)
    mixed = B.make_control(packed.C, packed.P, packed.M, precision="mixed",
                           refinement_tolerance=1e-6, **kw)
    got, report = B.solve_packed(device, mixed, packed)
    expected, _ = O.solve(B.make_control(packed.C, packed.P, packed.M, **kw), packed)
    base = packed.C * 12 + packed.P * 3
    print(f"  accepted {got[base + 10]:.0f} rejected {got[base + 11]:.0f} failed "
          f"{got[base + 17]:.0f}; cost {got[base + 6]:.12g} oracle {expected[base + 6]:.12g}",
          flush=True)
    assert got[base + 10] > 0
    assert got[base + 17] == 0
    assert abs(got[base + 6] - expected[base + 6]) <= 1e-6 * expected[base + 6]


def test_batch_keeps_each_bundle_budget(device):
    """Windows of 5 to 9 cameras with their default budgets (fixed route min(96, max(24, 12C)):
    60 to 96) in one batch equal their single solves bit for bit."""
    print("\n[options] per-bundle PCG budgets in a batch", flush=True)
    bundles = S.local_bundles(count=6, seed=21, cameras=(5, 9), points=(60, 300))
    controls = []
    for packed, _ in bundles:
        controls.append(B.make_control(packed.C, packed.P, packed.M, iterations=8,
                                       max_cg=min(96, max(24, 12 * packed.C)), huber=1.0,
                                       tolerance=1e-10))
    assert len({c.max_cg for c in controls}) > 1
    outs, report, _ = B.solve_batch(device, controls, [p for p, _ in bundles])
    for (packed, _), control, got in zip(bundles, controls, outs):
        alone, _ = B.solve_packed(device, control, packed)
        assert np.array_equal(got, alone), packed.C
