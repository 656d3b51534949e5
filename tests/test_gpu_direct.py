"""The direct route (TCV_BA_FLAG_DIRECT, the option `directSolve`): the reduced
camera system of each bundle of at most 2,048 cameras is assembled densely, factored by Cholesky
(one workgroup per bundle up to 128 cameras, a sequence of launches above) and certified by the
matrix-free residual.

Bars (stated here):
* against the oracle's matrix-free solve with PCG tightened to 1e-12 and a budget that reaches it
  (the same LM schedule, a near-exact linear solve per iteration): final cost within 1e-9 relative, every pose/point
  coordinate x within 1e-6·max(1, |x|), accepted/rejected counts equal;
* against the reference's dense CPU solver restated in the oracle (bundle-adjustment.ts:80-291) on
  the 10-camera ring: final cost within 1e-7 relative (focal.test.mjs:38);
* the mixed form against the FP64 oracle: final cost within 1e-6 relative;
* a batch equals its bundles solved alone, bit for bit, with bundles on both factorisations and
  one too large for the direct route, which keeps PCG.
"""
import numpy as np
import pytest

from oracle import oracle as O
from typecv_amd import bundle as B
from typecv_amd import synthetic as S

from test_gpu_parity import CASES, packed_problem

pytestmark = pytest.mark.gpu


def variant(control, **kw):
    base = dict(directional=control.directional_points, anchor=control.anchor,
                scale_view=control.scale_view, normalize_scale=bool(control.normalize_scale),
                shared_focal=bool(control.shared_focal), iterations=control.iterations,
                max_cg=control.max_cg, huber=control.huber,
                tolerance=control.relative_tolerance)
    base.update(kw)
    return B.make_control(control.cameras, control.points, control.observations, **base)


def problems():
    for name, model, options in CASES:
        if name == "ring_64":
            continue
        packed, control = packed_problem(model, options)
        yield name, packed, control
    ring, _ = S.ring_scene(seed=3, cameras=10, points=400)
    yield "ring_scene", ring, B.make_control(ring.C, ring.P, ring.M, iterations=20, max_cg=50)
    traj, _ = S.trajectory_scene(cameras=20, points=3000, seed=2)
    yield "trajectory_shared", traj, B.make_control(
        traj.C, traj.P, traj.M, iterations=20, max_cg=256, shared_focal=True, huber=1.0)
    # more points than one pass of the row workgroups' lookup table (8,192)
    wide, _ = S.trajectory_scene(cameras=12, points=20000, seed=7)
    yield "trajectory_wide", wide, B.make_control(
        wide.C, wide.P, wide.M, iterations=10, max_cg=256, shared_focal=True, huber=1.0)
    # above one workgroup's factorisation: the multi-launch route (1,800 + 1 unknowns, 29 panels)
    large, _ = S.trajectory_scene(cameras=300, points=12000, seed=5)
    yield "trajectory_large", large, B.make_control(
        large.C, large.P, large.M, iterations=8, max_cg=256, shared_focal=True, huber=1.0)


PROBLEMS = list(problems())
BY_NAME = {p[0]: p for p in PROBLEMS}
# the oracle's PCG budget per LM iteration: enough to reach 1e-12 on every problem here (the
# 300-camera one takes about 1,000 iterations; at 256 some of its steps stop short and are rejected)
TIGHT_CG = 8192


def compare(got, expected, base, cost_bar, coord_bar):
    cost = expected[base + 6]
    rel = abs(got[base + 6] - cost) / max(1.0, abs(cost))
    coord = np.max(np.abs(got[:base] - expected[:base]) / np.maximum(1.0, np.abs(expected[:base])))
    print(f"  final cost {got[base + 6]:.12g} vs {cost:.12g} (rel {rel:.2e}); "
          f"max coordinate error {coord:.2e}", flush=True)
    assert rel <= cost_bar
    assert coord <= coord_bar


@pytest.mark.parametrize("name,packed,control", PROBLEMS, ids=[p[0] for p in PROBLEMS])
def test_direct_matches_tight_pcg_oracle(device, name, packed, control):
    print(f"\n[direct] {name} C={packed.C} P={packed.P} M={packed.M}", flush=True)
    tight = variant(control, tolerance=1e-12, max_cg=TIGHT_CG)
    expected, stats = O.solve(tight, packed)
    got, report = B.solve_packed(device, variant(control, direct=True), packed)
    base = packed.C * 12 + packed.P * 3
    print(f"  direct solves {report['direct_solves']}, PCG {report['cg_iterations']} "
          f"(oracle PCG {stats['cg_iterations']:.0f})", flush=True)
    assert report["direct_solves"] > 0
    compare(got, expected, base, 1e-9, 1e-6)
    for word in (10, 11, 14):
        assert got[base + word] == expected[base + word], word
    if control.shared_focal:
        assert abs(got[base + 26] - expected[base + 26]) <= 1e-8


@pytest.mark.parametrize("name", ["trajectory_shared", "trajectory_large"])
def test_direct_needs_no_pcg_when_the_factor_certifies(device, name):
    _, packed, control = BY_NAME[name]
    got, report = B.solve_packed(device, variant(control, direct=True), packed)
    base = packed.C * 12 + packed.P * 3
    steps = got[base + 10] + got[base + 11]
    print(f"\n[direct] certified {report['direct_solves']} of {steps:.0f} LM steps, "
          f"PCG {report['cg_iterations']}", flush=True)
    assert report["direct_solves"] == steps
    assert report["cg_iterations"] == 0


def test_direct_ring_matches_the_dense_cpu_solver(device):
    print("\n[direct] cfg1 ring vs dense CPU solver", flush=True)
    ring, _ = S.ring_scene(seed=1)
    control = B.make_control(ring.C, ring.P, ring.M, iterations=20, max_cg=50, direct=True)
    got, report = B.solve_packed(device, control, ring)
    dense = O.dense_solve(control, ring)
    base = ring.C * 12 + ring.P * 3
    assert report["direct_solves"] > 0
    compare(got, dense, base, 1e-7, 1e-7)


MIXED = [BY_NAME[n] for n in ("ring_scene", "trajectory_shared", "trajectory_wide",
                                "trajectory_large")]


@pytest.mark.parametrize("name,packed,control", MIXED, ids=[p[0] for p in MIXED])
def test_direct_mixed_against_fp64_oracle(device, name, packed, control):
    """The factor of the FP32 cross blocks' system, certified and corrected in FP64: the final cost
    within the north star's 1e-6 of the FP64 oracle. Coordinates are printed, not bounded: a step
    is only as exact as its certified residual, and the far points of the wide scene move by the
    conditioning times it (1.5e-4 relative at a certified 1e-10, against 4e-9 for the FP64
    factor)."""
    print(f"\n[direct mixed] {name}", flush=True)
    expected, _ = O.solve(variant(control, tolerance=1e-12, max_cg=TIGHT_CG), packed)
    for refine in (0.0, 1e-10):
        got, report = B.solve_packed(device, variant(control, precision="mixed", direct=True,
                                                     refinement_tolerance=refine), packed)
        base = packed.C * 12 + packed.P * 3
        print(f"  refine {refine:g}: direct {report['direct_solves']}, "
              f"PCG {report['cg_iterations']}", flush=True)
        compare(got, expected, base, 1e-6, np.inf)


def test_direct_is_deterministic(device):
    print("\n[direct] determinism", flush=True)
    _, packed, control = BY_NAME["trajectory_large"]
    c = variant(control, direct=True)
    a, _ = B.solve_packed(device, c, packed)
    b, _ = B.solve_packed(device, c, packed)
    assert np.array_equal(a, b)


def test_direct_batch_equals_single_solves(device):
    print("\n[direct] batch of one-workgroup, multi-launch and PCG bundles", flush=True)
    scenes = [S.trajectory_scene(cameras=c, points=p, seed=s, track=(3, min(c, 12)))[0]
              for c, p, s in ((6, 800, 1), (130, 2000, 2), (25, 4000, 3), (2050, 6000, 4))]
    controls = [B.make_control(s.C, s.P, s.M, iterations=8, max_cg=256, shared_focal=True,
                               huber=1.0, direct=True) for s in scenes]
    outs, report, _ = B.solve_batch(device, controls, scenes)
    singles = 0
    for scene, control, out in zip(scenes, controls, outs):
        alone, rep = B.solve_packed(device, control, scene)
        singles += rep["direct_solves"]
        assert np.array_equal(out, alone)
    assert report["direct_solves"] == singles
    assert singles > 0


@pytest.mark.parametrize("case", ["joint_one_fixed", "joint_two_fixed", "directional"])
def test_direct_polish_and_backward_match_pcg(device, case):
    """The exact-Hessian solves of the Newton polish and of the adjoint backward through the direct
    route (the gauge regularisation enters the dense system through S_cc and V^-1): the polished
    solution and the pixel cotangent equal the PCG route's within 1e-8 relative."""
    import test_gpu_adjoint as T
    print(f"\n[direct] polish and backward ({case})", flush=True)
    if case == "directional":
        packed, control = T.directional_problem()
    else:
        packed, control = T.joint_problem(fixed_two=case == "joint_two_fixed")
    at, lf = T.solved(device, packed, control)
    direct = variant(control, direct=True)
    pcg_out, _ = B.polish(device, control, at, lf, iterations=4, tolerance=1e-13)
    dir_out, prep = B.polish(device, direct, at, lf, iterations=4, tolerance=1e-13)
    base = packed.C * 12 + packed.P * 3
    scale = np.maximum(1.0, np.abs(pcg_out[:base]))
    print(f"  polish: direct solves {prep['direct_solves']}, max coordinate difference "
          f"{np.max(np.abs(dir_out[:base] - pcg_out[:base]) / scale):.2e}", flush=True)
    assert prep["direct_solves"] > 0
    assert np.max(np.abs(dir_out[:base] - pcg_out[:base]) / scale) <= 1e-8
    pol = B.at_solution(at, pcg_out)
    lf = B.output_log_focal(control, pcg_out)
    sh = 1 if control.shared_focal else 0
    cot = np.random.default_rng(5).normal(size=12 * packed.C + 3 * packed.P + sh)
    # 3e-12: at 1e-12 the certificate of one of these systems sits at its FP64 floor, misses after
    # both corrections and hands over to PCG (measured: the pixel cotangents then agree to 8e-13)
    want, _, _, _ = B.backward(device, control, pol, lf, cot, tolerance=3e-12)
    got, _, _, brep = B.backward(device, direct, pol, lf, cot, tolerance=3e-12)
    rel = np.max(np.abs(got - want)) / np.max(np.abs(want))
    print(f"  backward: direct solves {brep['direct_solves']}, relative difference {rel:.2e}",
          flush=True)
    assert brep["direct_solves"] == 1
    assert rel <= 1e-8


def test_large_direct_polish_matches_pcg(device):
    """The exact-Hessian system through the multi-launch factorisation (a 140-camera bundle, 841
    unknowns; stored exact cross blocks for the certificate): the polished solution against the
    PCG route's, both solved to 1e-9 (PCG with 4,000 iterations), within 1e-8 relative. The
    certificate is the true residual relative to the right-hand side, the gradient, which is small
    at a converged solution; its FP64 floor on these bundles lies near 1e-10, past which the direct
    route hands its solves to PCG. (The adjoint of these trajectory scenes is not positive
    definite at the polished point, so the backward through the large route is not compared here;
    the small-bundle test covers the backward through the direct route.)"""
    print("\n[direct] large-bundle polish", flush=True)
    # no outliers and a Huber radius far above the noise
    scene, _ = S.trajectory_scene(cameras=140, points=3000, seed=9, track=(5, 12), outliers=0.0,
                                  noise_px=0.3)
    control = B.make_control(scene.C, scene.P, scene.M, iterations=30, max_cg=256,
                             shared_focal=True, huber=8.0, tolerance=1e-10)
    out, _ = B.solve_packed(device, control, scene)
    at, lf = B.at_solution(scene, out), B.output_log_focal(control, out)
    direct = variant(control, direct=True)
    pcg_out, _ = B.polish(device, control, at, lf, iterations=4, tolerance=1e-9)
    dir_out, prep = B.polish(device, direct, at, lf, iterations=4, tolerance=1e-9)
    base = scene.C * 12 + scene.P * 3
    scale = np.maximum(1.0, np.abs(pcg_out[:base]))
    diff = np.max(np.abs(dir_out[:base] - pcg_out[:base]) / scale)
    print(f"  polish: direct solves {prep['direct_solves']}, PCG {prep['cg_iterations']}, max "
          f"coordinate difference {diff:.2e}", flush=True)
    assert prep["direct_solves"] > 0
    assert diff <= 1e-8


def test_large_direct_correction_from_the_factor(device):
    """A 160-camera bundle (961 unknowns) at a 1e-12 certificate, near its FP64 floor: first
    certificates miss and corrections x += S^-1 r from the stored multi-launch factor (the forward
    and back substitutions of a residual) follow; a step whose two corrections still miss hands its
    x over to PCG, which finishes it in a few iterations. The result matches the oracle with PCG at
    1e-12 (cost 1e-9 relative, coordinates 1e-6)."""
    print("\n[direct] correction from the large factor", flush=True)
    scene, _ = S.trajectory_scene(cameras=160, points=6000, seed=4)
    control = B.make_control(scene.C, scene.P, scene.M, iterations=8, max_cg=256,
                             shared_focal=True, huber=1.0, tolerance=1e-12, direct=True)
    got, report = B.solve_packed(device, control, scene)
    base = scene.C * 12 + scene.P * 3
    steps = got[base + 10] + got[base + 11]
    print(f"  direct solves {report['direct_solves']} of {steps:.0f}, corrections "
          f"{report['direct_corrections']}, PCG {report['cg_iterations']}", flush=True)
    assert report["direct_corrections"] >= 1
    assert report["direct_solves"] >= steps - 1
    assert report["cg_iterations"] <= 16
    expected, _ = O.solve(variant(control, direct=False, max_cg=TIGHT_CG), scene)
    compare(got, expected, base, 1e-9, 1e-6)
