"""Differentiation on the shared-focal trajectory workload (FP64 forward, 30 LM iterations).

* The Jacobian kernels at full size (500 cameras, 200,000 points, 1.89M observations): 2,000
  sampled observations against central differences (h = 1e-6, the reference's geometry tests' step,
  tests/geometry.test.ts:46-72) of the oracle's projection, pose chart, point and log focal.
* The Gauss-Newton adjoint at full size, at the forward's own output. The exact mode is undefined
  there (the 30-iteration forward is not a minimum: point blocks of the exact Hessian are
  indefinite, and the damped exact-Hessian polish does not reach one within 200 steps); the
  Gauss-Newton system sum_i J_i^T D_i J_i is. Checked on the host from the Jacobian output: the
  adjoint solves H_GN lambda = v (v a point cotangent orthogonal to the scale gauge, so the gauge
  map's adjoint leaves it unchanged), and the pixel and input-focal cotangents are D_i J_i lambda
  and -sum_i (D_i J_i lambda) . Jf_i; PCG and the direct route agree to 1e-4.
* The exact adjoint against finite differences at 200 cameras / 80,000 points (the workload's 400
  points per camera), the largest size of those tried (40 to 500 cameras) where the damped polish
  reaches a stationary point (gradient 1e3 -> 1e-3), so that the implicit derivative exists and
  re-solves of perturbed inputs stay in its linear regime: 64 single-observation pixel
  perturbations (h = 1e-3 px) and 8 cameras' input focal (h = 1e-6 in log f), both against central
  differences of Newton re-polishes. The polish holds 15 points fixed there (tracks run off to
  1e8 times the scene depth, their Gauss-Newton blocks singular to rounding); the cotangent is zero
  on them. The Gauss-Newton mode's error against the same differences is reported and bounded.
"""
import dataclasses

import numpy as np
import pytest

from typecv_amd import _native
from typecv_amd import bundle as B
from typecv_amd import synthetic as S
import derivatives as DV

pytestmark = pytest.mark.gpu


def forward(device, cameras, points):
    packed, info = S.trajectory_scene(cameras=cameras, points=points, seed=2)
    control = B.make_control(packed.C, packed.P, packed.M, shared_focal=True, huber=1.0,
                             iterations=30, max_cg=min(256, max(32, 12 * packed.C + 1)))
    out, rep = B.solve_packed(device, control, packed)
    assert rep["lm_iterations"] == 30
    return packed, control, B.at_solution(packed, out), B.output_log_focal(control, out)


@pytest.fixture(scope="module")
def full(device):
    print("\n[differentiation] 500-camera forward", flush=True)
    packed, control, at, lf = forward(device, 500, 200_000)
    J = B.jacobian(device, control, at, lf)
    return packed, control, at, lf, J


def test_jacobian_at_full_size(full):
    packed, control, at, lf, J = full
    assert packed.M > 1_800_000
    sample = np.random.default_rng(2).choice(packed.M, size=2000, replace=False)
    worst = DV.jacobian_vs_differences(control, at, packed.index, J, lf, sample)
    print(f"  jacobian vs central differences: max rel err {worst:.2e}", flush=True)
    assert worst <= 1e-7


def test_gauss_newton_backward_at_full_size(device, full):
    packed, control, at, lf, J = full
    C, P, M = packed.C, packed.P, packed.M
    cot, chart = DV.point_cotangent(control, at.scene, np.random.default_rng(5))
    with pytest.raises(_native.EngineError, match="not positive definite"):
        B.backward(device, control, at, lf, cot, tolerance=1e-10, max_cg=20000, mode="exact")
    pix, foc, lam, rep = B.backward(device, control, at, lf, cot, tolerance=1e-10, max_cg=20000,
                                    mode="gaussNewton")
    rows = np.ones(len(lam), dtype=bool)
    rows[6 * control.anchor: 6 * control.anchor + 6] = False
    res = np.linalg.norm((DV.gn_product(control, packed.index, J, lam) - chart)[rows])
    res /= np.linalg.norm(chart)
    pe = np.max(np.abs(DV.pixel_cotangent(control, packed.index, J, lam) - pix)) / np.max(np.abs(pix))
    fe = np.max(np.abs(DV.focal_cotangent(control, packed.index, J, lam, exact=False) - foc))
    fe /= np.max(np.abs(foc))
    print(f"  GN backward: {rep['cg_iterations']} PCG iterations, {rep['solve_ms']:.0f} ms; "
          f"|H_GN lambda - v| / |v| {res:.2e}, pixel {pe:.2e}, focal {fe:.2e}, "
          f"max |lambda| {np.max(np.abs(lam)):.2e}", flush=True)
    # the system's conditioning (|lambda| reaches 1e8 for a unit cotangent) bounds the residual
    assert res <= 1e-5
    assert pe <= 1e-10 and fe <= 1e-10
    direct = B.make_control(C, P, M, shared_focal=True, huber=1.0, iterations=30,
                            max_cg=control.max_cg, direct=True)
    dpix, dfoc, _, drep = B.backward(device, direct, at, lf, cot, tolerance=1e-10, max_cg=20000,
                                     mode="gaussNewton")
    # the direct route's certificate may miss on a system this ill-conditioned: it then hands over
    # to PCG from the direct solution
    dp = np.max(np.abs(dpix - pix)) / np.max(np.abs(pix))
    df = np.max(np.abs(dfoc - foc)) / np.max(np.abs(foc))
    print(f"  direct route: {drep['solve_ms']:.0f} ms, {drep['direct_solves']} certified direct "
          f"solves, {drep['cg_iterations']} PCG iterations; pixel {dp:.2e}, focal {df:.2e}",
          flush=True)
    assert dp <= 1e-4 and df <= 1e-4


def held_points(control, packed, J) -> np.ndarray:
    """Points whose Gauss-Newton block sum w Jp^T Jp is singular to rounding (smallest eigenvalue
    below 1e-11 of the largest): the points the exact-Hessian entries hold fixed."""
    M, P = control.observations, control.points
    _, _, Jp, _, w, ok = DV.parts(J)
    pts = np.asarray(packed.index[M: 2 * M], dtype=np.int64)
    V = np.zeros((P, 3, 3))
    np.add.at(V, pts[ok], w[ok, None, None] * np.einsum("mai,maj->mij", Jp[ok], Jp[ok]))
    e = np.linalg.eigvalsh(V)
    return np.nonzero(e[:, 0] < 1e-11 * e[:, 2])[0]


def test_exact_backward_against_differences_at_200_cameras(device):
    print("\n[differentiation] 200-camera forward and polish", flush=True)
    packed, control, at, lf = forward(device, 200, 80_000)
    C, P, M = packed.C, packed.P, packed.M
    out, prep = B.polish(device, control, at, lf, iterations=4, tolerance=1e-12,
                         lm_iterations=200, max_cg=20000)
    assert prep["refinement_steps"] == 4
    pol = B.at_solution(at, out)
    lf = B.output_log_focal(control, out)
    J = B.jacobian(device, control, pol, lf)
    held = held_points(control, packed, J)
    print(f"  polish: {prep['lm_iterations']} damped, 4 Newton steps, {prep['held_points']} points "
          f"held", flush=True)
    assert prep["held_points"] == len(held) > 0
    rng = np.random.default_rng(21)
    n_out = 12 * C + 3 * P + 1
    cot = rng.normal(size=n_out)
    for p in held:
        cot[12 * C + 3 * p: 12 * C + 3 * p + 3] = 0
    pix, foc, lam, rep = B.backward(device, control, pol, lf, cot, tolerance=1e-12, max_cg=20000)
    assert rep["held_points"] == len(held)
    for p in held:
        assert np.all(lam[6 * C + 3 * p: 6 * C + 3 * p + 3] == 0)
    gpix, gfoc, _, _ = B.backward(device, control, pol, lf, cot, tolerance=1e-12, max_cg=20000,
                                  mode="gaussNewton")

    def fd_of(scene, h):
        outs = []
        for sign in (1, -1):
            moved = dataclasses.replace(pol, scene=scene(sign * h))
            o, _ = B.polish(device, control, moved, lf, iterations=2, tolerance=1e-12,
                            max_cg=20000)
            outs.append(o)
        d = cot[:n_out - 1] @ ((outs[0][:n_out - 1] - outs[1][:n_out - 1]) / (2 * h))
        return d + cot[-1] * (B.output_log_focal(control, outs[0])
                              - B.output_log_focal(control, outs[1])) / (2 * h)

    radius = np.hypot(J[:, 0], J[:, 1])
    base = 18 * C + 3 * P
    errs, gn_errs = [], []
    for k in range(64):
        i = int(rng.integers(0, M))
        d = rng.normal(size=2)
        d /= np.linalg.norm(d)
        hp = 1e-3

        def scene(step, i=i, d=d):
            s = pol.scene.copy()
            s[base + 2 * i: base + 2 * i + 2] += step * d
            return s

        fd = fd_of(scene, hp)
        if abs(radius[i] - control.huber) < 2 * hp:
            continue  # the kink of the robust cost: a one-sided derivative
        errs.append(abs(fd - float(pix[i] @ d)) / max(abs(fd), 1e-2 * np.max(np.abs(pix))))
        gn_errs.append(abs(fd - float(gpix[i] @ d)) / max(abs(fd), 1e-2 * np.max(np.abs(pix))))
    models = 12 * C + 3 * P
    fds, cams = [], np.linspace(1, C - 1, 8).astype(int)
    for c in cams:
        def scene(step, c=c):
            s = pol.scene.copy()
            s[models + 6 * c: models + 6 * c + 2] *= np.exp(step)
            return s

        fds.append(fd_of(scene, 1e-6))
    fds = np.array(fds)
    fe = np.max(np.abs(foc[cams] - fds)) / np.max(np.abs(fds))
    ge = np.max(np.abs(gfoc[cams] - fds)) / np.max(np.abs(fds))
    print(f"  exact: pixel max rel err {max(errs):.2e} over {len(errs)} perturbations, focal "
          f"{fe:.2e}; Gauss-Newton: pixel median {np.median(gn_errs):.2e}, max {max(gn_errs):.2e}, "
          f"focal {ge:.2e}", flush=True)
    assert len(errs) >= 60
    assert max(errs) <= 2e-4
    assert fe <= 1e-3
    # the Gauss-Newton mode drops the residual-weighted curvature: percent-level here
    assert np.median(gn_errs) <= 0.05 and ge <= 0.1
