"""GPU parity: the HIP engine against the CPU oracle on the same seeded inputs.

Bars (stated here): the assembled linear system (RHS, preconditioned vector,
matrix-free Schur product) within 1e-10 relative to the oracle's; a whole solve's final cost within
1e-9 relative and every pose/point coordinate x within 1e-7·max(1, |x|); integer results
(accepted/rejected steps, failures, input-error words) equal.

The engine sums in a different (fixed) order than the oracle's serial CPU-tier loops — wave
butterflies and per-task partials instead of one running sum — so the two agree to rounding, which
hundreds of PCG iterations amplify by the conditioning of the reduced camera system; hence the
coordinate bound scales with the coordinate (scenes here reach |x| ~ 35).
"""
import numpy as np
import pytest

from typecv_amd import bundle as B
from typecv_amd import synthetic as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FIXED = lambda C: min(96, max(24, 12 * C))  # noqa: E731
SHARED = lambda C: min(256, max(32, 12 * C + 1))  # noqa: E731


def packed_problem(model, options):
    shared = bool(options.get("refineSharedFocal"))
    problem = B.bundle_problem(model, options, SHARED if shared else FIXED)
    packed = B.pack_bundle_scene(problem["input"], problem["byId"], problem["active"])
    control = B.make_control(
        packed.C, packed.P, packed.M, directional=packed.P - problem["finitePoints"],
        anchor=problem["byId"][problem["anchorId"]], scale_view=problem["byId"][problem["scaleId"]],
        normalize_scale=problem["fixed"] == 1, shared_focal=shared,
        iterations=problem["iterations"], max_cg=problem["maxCg"], huber=problem["huber"],
        tolerance=problem["tolerance"], precision=options.get("precision", "fp64"))
    return packed, control


def cases():
    yield "bundle_fixture", S.bundle_fixture(), {"iterations": 20}
    yield "bundle_fixture_outlier", S.bundle_fixture(outlier=True), {"iterations": 20, "huberPx": 1}
    yield "joint_focal", S.joint_focal_fixture()["input"], {
        "iterations": 40, "refineSharedFocal": True, "maxCgIterations": 256,
        "cgRelativeTolerance": 1e-10}
    yield "ring_64", S.ring_bundle_fixture(views=64, points=960), {"iterations": 8}


CASES = list(cases())


def packed_cases():
    ring, _ = S.ring_scene(seed=3, cameras=10, points=400)
    yield "ring_scene", ring, B.make_control(ring.C, ring.P, ring.M, iterations=20, max_cg=50)
    traj, _ = S.trajectory_scene(cameras=40, points=3000, seed=5)
    yield "trajectory_shared", traj, B.make_control(
        traj.C, traj.P, traj.M, iterations=15, max_cg=256, shared_focal=True, huber=1.0)
    # over 1024 observations per camera: every camera list is split into two tasks
    dense, _ = S.trajectory_scene(cameras=20, points=3000, seed=2)
    yield "trajectory_split_cameras", dense, B.make_control(
        dense.C, dense.P, dense.M, iterations=30, max_cg=256, shared_focal=True, huber=1.0)


PACKED = list(packed_cases())


@pytest.mark.parametrize("name,model,options", CASES, ids=[c[0] for c in CASES])
def test_schur_probe_matches_oracle(device, name, model, options):
    print(f"\n[probe] {name}", flush=True)
    packed, control = packed_problem(model, options)
    m = control.cameras * 6 + control.shared_focal
    rng = np.random.default_rng(7)
    p = rng.normal(size=m)
    for lam in (1e-3, 1.0):
        rhs, sp, z, status = O.schur_probe(control, packed, lam, 0.01, p)
        grhs, gsp, gz, gstatus, _ = B.schur_probe(device, control, packed, lam, 0.01, p)
        assert gstatus == status
        for a, b in ((grhs, rhs), (gsp, sp), (gz, z)):
            assert np.max(np.abs(a - b)) <= 1e-10 * max(1.0, np.max(np.abs(b)))


def assert_solve_parity(got, expected, base):
    cost = expected[base + 6]
    assert abs(got[base + 6] - cost) <= 1e-9 * max(1.0, abs(cost)), (got[base + 6], cost)
    scale = np.maximum(1.0, np.abs(expected[:base]))
    assert np.max(np.abs(got[:base] - expected[:base]) / scale) <= 1e-7
    for word in (10, 11, 14, 17):
        assert got[base + word] == expected[base + word], word


@pytest.mark.parametrize("name,model,options", CASES, ids=[c[0] for c in CASES])
def test_solve_matches_oracle(device, name, model, options):
    print(f"\n[solve] {name}", flush=True)
    packed, control = packed_problem(model, options)
    expected, _ = O.solve(control, packed)
    got, report = B.solve_packed(device, control, packed)
    base = packed.C * 12 + packed.P * 3
    assert_solve_parity(got, expected, base)
    if control.shared_focal:
        assert abs(got[base + 26] - expected[base + 26]) <= 1e-9


@pytest.mark.parametrize("name,packed,control", PACKED, ids=[c[0] for c in PACKED])
def test_packed_solve_matches_oracle(device, name, packed, control):
    print(f"\n[solve] {name} C={packed.C} P={packed.P} M={packed.M}", flush=True)
    expected, _ = O.solve(control, packed)
    got, report = B.solve_packed(device, control, packed)
    base = packed.C * 12 + packed.P * 3
    assert expected[base + 6] < expected[base + 8]
    assert_solve_parity(got, expected, base)


def test_solve_is_deterministic(device):
    print("\n[determinism]", flush=True)
    packed, control = PACKED[1][1], PACKED[1][2]
    a, _ = B.solve_packed(device, control, packed)
    b, _ = B.solve_packed(device, control, packed)
    assert np.array_equal(a, b)


def test_cfg1_ring_matches_the_dense_cpu_solver(device):
    """The GPU engine on the 10-camera ring scene (2,000 points, ~20k observations, 20 LM
    iterations, PCG <= 50) against the reference's dense CPU solver restated in the oracle
    (bundle-adjustment.ts:80-291): final cost within 1e-7 relative (focal.test.mjs:38)."""
    print("\n[parity] cfg1 ring vs dense CPU solver", flush=True)
    ring, _ = S.ring_scene(seed=1)
    control = B.make_control(ring.C, ring.P, ring.M, iterations=20, max_cg=50)
    got, _ = B.solve_packed(device, control, ring)
    dense = O.dense_solve(control, ring)
    base = ring.C * 12 + ring.P * 3
    assert abs(got[base + 6] - dense[base + 6]) <= 1e-7 * dense[base + 6]
    assert np.max(np.abs(got[:base] - dense[:base]) / np.maximum(1.0, np.abs(dense[:base]))) <= 1e-7
