"""Pins of the dense derivative references (oracle/adjoint.py), on the CPU.

* The oracle's chart gradient against central differences of its own half robust cost along random
  chart directions (movePose for cameras), with Huber outliers and the shared focal row.
* The FD Hessian against second differences of the cost, and symmetric.
* Dense Newton reaches a stationary point with a positive definite Hessian (two fixed cameras: no
  gauge freedom).
* Implicit differentiation with the dense Hessian against finite differences of the re-solved
  problem under pixel perturbations — the definition of the derivative the engine's backward
  computes.
"""
import numpy as np
import pytest

from oracle import adjoint as A
from oracle import oracle as O
from typecv_amd import synthetic as S
from problems import packed_problem


def _problem(noise=0.6, outliers=False, shared=True, fixed_two=True):
    model = S.joint_focal_fixture(views=5, points=30, focal_scale=0.97, noise=noise)["input"]
    if outliers:
        # two observations far off, beyond the Huber radius
        model["landmarks"][3]["observations"][1]["pixel"][0] += 25.0
        model["landmarks"][7]["observations"][2]["pixel"][1] -= 18.0
    options = {"iterations": 30, "refineSharedFocal": shared, "huberPx": 2.0,
               "maxCgIterations": 256, "cgRelativeTolerance": 1e-10}
    if fixed_two:
        options["variableViewIds"] = [v["id"] for v in model["views"][2:]]
    return packed_problem(model, options)


def _solved(packed, control):
    out, _ = O.solve(control, packed)
    n = packed.C * 12 + packed.P * 3
    scene = packed.scene.copy()
    scene[:n] = out[:n]
    lf = float(out[n + 26]) if control.shared_focal else 0.0
    return scene, lf


def _direction(control, index, rng, scale=1.0):
    C, P, sh, n = A.sizes(control)
    d = np.zeros(n)
    rows = A.variable_rows(control, index)
    d[rows] = rng.normal(size=len(rows)) * scale
    return d


@pytest.mark.parametrize("outliers", [False, True])
def test_gradient_matches_cost_differences(outliers):
    packed, control = _problem(outliers=outliers)
    scene, lf = _solved(packed, control)
    rng = np.random.default_rng(3)
    g, _, skipped = O.gradient(control, scene, packed.index, lf)
    assert not skipped
    for _ in range(4):
        d = _direction(control, packed.index, rng, 1e-2)
        h = 1e-5
        sp, fp = A.move(control, scene, lf, h * d)
        sm, fm = A.move(control, scene, lf, -h * d)
        ep = O.gradient(control, sp, packed.index, fp)[1]
        em = O.gradient(control, sm, packed.index, fm)[1]
        fd = (ep - em) / (2 * h)
        assert abs(fd - g @ d) <= 1e-6 * max(1.0, abs(g @ d)) + 1e-7, (fd, g @ d)


def test_fd_hessian_matches_second_differences():
    packed, control = _problem(outliers=True)
    scene, lf = _solved(packed, control)
    rows = A.variable_rows(control, packed.index)
    H = A.fd_hessian(control, scene, packed.index, lf)
    rng = np.random.default_rng(5)
    for _ in range(3):
        d = _direction(control, packed.index, rng, 1e-2)
        h = 1e-3
        e0 = O.gradient(control, scene, packed.index, lf)[1]
        sp, fp = A.move(control, scene, lf, h * d)
        sm, fm = A.move(control, scene, lf, -h * d)
        ep = O.gradient(control, sp, packed.index, fp)[1]
        em = O.gradient(control, sm, packed.index, fm)[1]
        second = (ep - 2 * e0 + em) / h**2
        quad = d[rows] @ H @ d[rows]
        assert abs(second - quad) <= 1e-4 * abs(quad), (second, quad)


def test_dense_newton_reaches_a_minimum():
    packed, control = _problem(outliers=True)
    scene, lf = _solved(packed, control)
    rows = A.variable_rows(control, packed.index)
    g0 = np.linalg.norm(O.gradient(control, scene, packed.index, lf)[0][rows])
    scene, lf = A.newton(control, scene, packed.index, lf, steps=3)
    g1 = np.linalg.norm(O.gradient(control, scene, packed.index, lf)[0][rows])
    assert g1 <= 1e-7 * max(g0, 1.0), (g0, g1)
    H = A.fd_hessian(control, scene, packed.index, lf)
    assert np.linalg.eigvalsh(H).min() > 0


def test_dense_vjp_matches_finite_differences_of_the_solution():
    packed, control = _problem(outliers=True)
    scene, lf = _solved(packed, control)
    scene, lf = A.newton(control, scene, packed.index, lf, steps=3)
    C, P, sh, n = A.sizes(control)
    rng = np.random.default_rng(11)
    # a cotangent on the points and the focal scale (chart = output coordinates there)
    v = np.zeros(n)
    v[6 * C:] = rng.normal(size=n - 6 * C)
    base = packed.__class__(packed.C, packed.P, packed.M, scene, packed.index)
    grad, _ = A.dense_vjp(control, base, scene, lf, v, control.huber)
    M = packed.M
    h = 1e-4
    for trial in range(3):
        du = np.zeros(2 * M)
        du[rng.integers(0, 2 * M, size=4)] = rng.normal(size=4)
        pix = 18 * C + 3 * P
        sp = scene.copy()
        sp[pix:] += h * du
        sm = scene.copy()
        sm[pix:] -= h * du
        xp, fp = A.newton(control, sp, packed.index, lf, steps=3)
        xm, fm = A.newton(control, sm, packed.index, lf, steps=3)
        dX = (xp[12 * C: 12 * C + 3 * P] - xm[12 * C: 12 * C + 3 * P]) / (2 * h)
        fd = v[6 * C: 6 * C + 3 * P] @ dX + (v[-1] * (fp - fm) / (2 * h) if sh else 0.0)
        assert abs(fd - grad @ du) <= 1e-5 * max(abs(fd), 1e-3), (trial, fd, grad @ du)
