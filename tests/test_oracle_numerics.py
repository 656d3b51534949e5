"""Pins of the CPU oracle (oracle/ba_oracle.c) that do not come from the oracle itself.

* projection derivative and the pose (left update), point and log-focal Jacobians against central
  differences, as the reference's own tests do (tests/geometry.test.ts:19-72: h = 1e-5 / tol 1e-6
  for the projection, h = 1e-6 / tol 1e-5 for pose and point);
* solve3 / solve6 against numpy.linalg.solve on random SPD blocks, and their refusals;
* movePose against the matrix exponential of the skew matrix; focalScale against exp;
  radialLimit against the zero of d(r·D(r))/dr; robustCost against the Huber closed form;
* the matrix-free Schur product, reduced RHS and block-Jacobi preconditioner against a dense
  Hessian assembled here in numpy from the oracle's per-observation Jacobians (brute force:
  S = U - W V^-1 W^T, b = g_c - W V^-1 g_p), on the fixed and shared-focal routes;
* PCG against a dense solve of the same system (true residual within the stopping tolerance).
"""
import math

import numpy as np
import pytest
from scipy.linalg import expm

from oracle import oracle as O
from typecv_amd import bundle as B
from typecv_amd import synthetic as S

FIXED = lambda C: min(96, max(24, 12 * C))  # noqa: E731
SHARED = lambda C: min(256, max(32, 12 * C + 1))  # noqa: E731


def problem(model, options):
    shared = bool(options.get("refineSharedFocal"))
    prob = B.bundle_problem(model, options, SHARED if shared else FIXED)
    packed = B.pack_bundle_scene(prob["input"], prob["byId"], prob["active"])
    control = B.make_control(
        packed.C, packed.P, packed.M, directional=packed.P - prob["finitePoints"],
        anchor=prob["byId"][prob["anchorId"]], scale_view=prob["byId"][prob["scaleId"]],
        normalize_scale=prob["fixed"] == 1, shared_focal=shared, iterations=prob["iterations"],
        max_cg=prob["maxCg"], huber=prob["huber"], tolerance=prob["tolerance"])
    return packed, control


RADIAL = np.array([900.0, 920.0, 480.0, 270.0, -0.05, 0.005])


def test_projection_derivative_central_differences():
    rng = np.random.default_rng(1)
    for _ in range(200):
        q = np.array([rng.uniform(-1, 1), rng.uniform(-0.8, 0.8), rng.uniform(2, 6)])
        ok, px, du, dv = O.project(RADIAL, q)
        assert ok
        h = 1e-5
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            _, a, _, _ = O.project(RADIAL, q + e)
            _, b, _, _ = O.project(RADIAL, q - e)
            fd = (a - b) / (2 * h)
            assert abs(fd[0] - du[k]) < 1e-6 * max(1, abs(du[k]))
            assert abs(fd[1] - dv[k]) < 1e-6 * max(1, abs(dv[k]))


def test_projection_refusals_and_closed_form():
    # behind the camera, beyond the injective radial branch, or a plain pinhole value
    assert not O.project(RADIAL, [0.0, 0.0, -1.0])[0]
    limit = O.radial_limit(RADIAL)
    x = math.sqrt(limit) * 1.01
    assert not O.project(RADIAL, [x, 0.0, 1.0])[0]
    ok, px, du, dv = O.project([800, 800, 320, 240, 0, 0], [0.5, -0.25, 2.0])
    assert ok and np.allclose(px, [800 * 0.25 + 320, 800 * -0.125 + 240], rtol=0, atol=1e-12)
    assert np.allclose(du, [400, 0, -800 * 0.5 / 4], atol=1e-12)


def test_radial_limit_is_the_first_critical_radius():
    for k1, k2 in ((-0.05, 0.005), (-0.3, 0.02), (-0.2, 0.0), (0.1, -0.01)):
        m = np.array([1, 1, 0, 0, k1, k2], dtype=float)
        lim = O.radial_limit(m)
        if lim < 1e30:
            # d(r D(r))/dr = 1 + 3 k1 r^2 + 5 k2 r^4 vanishes at r^2 = limit
            assert abs(1 + 3 * k1 * lim + 5 * k2 * lim * lim) < 1e-9
    assert O.radial_limit(np.array([1, 1, 0, 0, 0.1, 0.0])) == 1e30


def test_robust_cost_is_huber():
    for e in (0.0, 0.3, 1.0, 1.5, 10.0):
        for h in (0.5, 1.0, 2.0):
            expect = e * e if e <= h else 2 * h * e - h * h
            assert O.robust_cost(e, h) == pytest.approx(expect, rel=1e-15, abs=0)


def random_spd(rng, n):
    a = rng.normal(size=(n, n))
    return a @ a.T + n * np.eye(n) * rng.uniform(0.1, 2)


def test_solve3_solve6_match_numpy():
    rng = np.random.default_rng(2)
    for n, solve in ((3, O.solve3), (6, O.solve6)):
        for _ in range(100):
            A = random_spd(rng, n)
            scale = np.diag(10.0 ** rng.uniform(-3, 3, n))
            A = scale @ A @ scale  # badly scaled: the equilibration is what makes this work
            b = rng.normal(size=n)
            ok, x = solve(A, b)
            assert ok
            ref = np.linalg.solve(A, b)
            assert np.max(np.abs(x - ref)) <= 1e-9 * np.max(np.abs(ref))
        # refusals: a non-positive diagonal, and a (numerically) singular block
        A = random_spd(rng, n)
        A[1, 1] = -1.0
        assert not solve(A, np.ones(n))[0]
        v = rng.normal(size=n)
        assert not solve(np.outer(v, v), np.ones(n))[0]


def test_move_pose_is_the_left_so3_exponential():
    rng = np.random.default_rng(3)
    for scale in (1e-9, 1e-4, 0.3, 2.0):
        R0 = expm(skew(rng.normal(size=3)))
        p = np.concatenate([R0.reshape(9), rng.normal(size=3)])
        d = np.concatenate([rng.normal(size=3) * scale, rng.normal(size=3)])
        out = O.move_pose(p, d)
        dR = expm(skew(d[:3]))
        assert np.allclose(out[:9].reshape(3, 3), dR @ R0, atol=1e-14)
        assert np.allclose(out[9:], dR @ p[9:] + d[3:], atol=1e-13)


def skew(w):
    return np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])


def test_focal_scale_matches_exp():
    for x in np.linspace(-20, 20, 401):
        assert O.focal_scale(x) == pytest.approx(math.exp(x), rel=1e-13)


def observation_pixel(pose, model, X):
    q = pose[:9].reshape(3, 3) @ X + pose[9:]
    ok, px, _, _ = O.project(model, q)
    assert ok
    return px


@pytest.mark.parametrize("shared", [False, True])
def test_jacobians_central_differences(shared):
    if shared:
        model = S.joint_focal_fixture(views=4, points=12)["input"]
        options = {"refineSharedFocal": True}
    else:
        model = S.bundle_fixture(count=12)
        options = {}
    packed, control = problem(model, options)
    m = control.cameras * 6 + control.shared_focal
    log_f = 0.05 if shared else 0.0
    _, _, _, _, blocks = O.schur_probe(control, packed, 1e-3, log_f, np.zeros(m), blocks=True)
    J = blocks["J"]
    poses, models, points = packed.poses(), packed.models().copy(), packed.points()
    if shared:
        models[:, :2] *= math.exp(log_f)
    obs_cam = packed.section("observationCamera")
    obs_pt = packed.section("observationPoint")
    active = packed.section("cameraActive")
    for i in range(packed.M):
        c, p = int(obs_cam[i]), int(obs_pt[i])
        pose, k, X = poses[c], models[c], points[p]
        if active[c]:
            h = 1e-6
            for j in range(6):
                d = np.zeros(6)
                d[j] = h
                a = observation_pixel(O.move_pose(pose, d), k, X)
                b = observation_pixel(O.move_pose(pose, -d), k, X)
                fd = (a - b) / (2 * h)
                assert abs(fd[0] - J[i, j]) < 1e-5 * max(1, abs(J[i, j]))
                assert abs(fd[1] - J[i, 6 + j]) < 1e-5 * max(1, abs(J[i, 6 + j]))
        else:
            assert np.all(J[i, :12] == 0)
        h = 1e-6
        for j in range(3):
            e = np.zeros(3)
            e[j] = h
            fd = (observation_pixel(pose, k, X + e) - observation_pixel(pose, k, X - e)) / (2 * h)
            assert abs(fd[0] - J[i, 12 + j]) < 1e-5 * max(1, abs(J[i, 12 + j]))
            assert abs(fd[1] - J[i, 15 + j]) < 1e-5 * max(1, abs(J[i, 15 + j]))
        if shared:
            jf = blocks["focal"][6 * packed.C + 6 + 2 * i: 6 * packed.C + 8 + 2 * i]
            up, dn = k.copy(), k.copy()
            up[:2] *= math.exp(h)
            dn[:2] *= math.exp(-h)
            fd = (observation_pixel(pose, up, X) - observation_pixel(pose, dn, X)) / (2 * h)
            assert np.all(np.abs(fd - jf) < 1e-5 * np.maximum(1, np.abs(jf)))


def dense_system(packed, control, blocks, lam):
    """The damped Hessian of the variable parameters assembled from per-observation Jacobians,
    its Schur complement on the camera (+ focal) block and the reduced RHS, in numpy."""
    C, P, M = packed.C, packed.P, packed.M
    shared = bool(control.shared_focal)
    active = packed.section("cameraActive").astype(bool)
    cams = [c for c in range(C) if active[c]]
    col = {c: 6 * k for k, c in enumerate(cams)}
    na = 6 * len(cams) + (1 if shared else 0)
    n = na + 3 * P
    H = np.zeros((n, n))
    g = np.zeros(n)
    J = blocks["J"]
    obs_cam = packed.section("observationCamera")
    obs_pt = packed.section("observationPoint")
    for i in range(M):
        c, p = int(obs_cam[i]), int(obs_pt[i])
        Ji = np.zeros((2, n))
        if active[c]:
            Ji[0, col[c]: col[c] + 6] = J[i, 0:6]
            Ji[1, col[c]: col[c] + 6] = J[i, 6:12]
        Ji[0, na + 3 * p: na + 3 * p + 3] = J[i, 12:15]
        Ji[1, na + 3 * p: na + 3 * p + 3] = J[i, 15:18]
        if shared:
            Ji[:, na - 1] = blocks["focal"][6 * C + 6 + 2 * i: 6 * C + 8 + 2 * i]
        w, r = J[i, 20], J[i, 18:20]
        H += w * Ji.T @ Ji
        g -= w * Ji.T @ r
    d = np.diag(H).copy()
    H[np.diag_indices(n)] += lam * np.maximum(d, 1e-10)
    U, W, V = H[:na, :na], H[:na, na:], H[na:, na:]
    Vi = np.linalg.inv(V)
    Sd = U - W @ Vi @ W.T
    b = g[:na] - W @ Vi @ g[na:]
    rows = [6 * c + j for c in cams for j in range(6)] + ([6 * C] if shared else [])
    return Sd, b, np.array(rows), cams


CASES = [
    ("fixed", lambda: S.bundle_fixture(count=16), {}),
    ("fixed_huber_outlier", lambda: S.bundle_fixture(count=16, outlier=True), {"huberPx": 1.0}),
    ("local_window", lambda: S.bundle_fixture(count=16, views=5), {"variableViewIds": [30, 40]}),
    ("shared_focal", lambda: S.joint_focal_fixture(views=5, points=16)["input"],
     {"refineSharedFocal": True}),
]


@pytest.mark.parametrize("name,make,options", CASES, ids=[c[0] for c in CASES])
def test_schur_product_matches_dense_brute_force(name, make, options):
    packed, control = problem(make(), options)
    m = control.cameras * 6 + control.shared_focal
    rng = np.random.default_rng(4)
    for lam in (1e-3, 0.7):
        p = rng.normal(size=m)
        rhs, sp, z, status, blocks = O.schur_probe(control, packed, lam, 0.02, p, blocks=True)
        assert status == 0
        Sd, b, rows, cams = dense_system(packed, control, blocks, lam)
        ref = Sd @ p[rows]
        assert np.max(np.abs(sp[rows] - ref)) <= 1e-10 * np.max(np.abs(ref))
        assert np.max(np.abs(rhs[rows] - b)) <= 1e-10 * max(1, np.max(np.abs(b)))
        fixed = np.setdiff1d(np.arange(6 * control.cameras), rows)
        assert np.array_equal(sp[fixed], p[fixed])  # fixed camera rows: identity block
        assert np.all(rhs[fixed] == 0)
        for k, c in enumerate(cams):  # block-Jacobi preconditioner: inverse of S_cc
            blk = Sd[6 * k: 6 * k + 6, 6 * k: 6 * k + 6]
            zc = np.linalg.solve(blk, p[6 * c: 6 * c + 6])
            assert np.max(np.abs(z[6 * c: 6 * c + 6] - zc)) <= 1e-9 * np.max(np.abs(zc))
        if control.shared_focal:
            assert z[m - 1] == pytest.approx(p[m - 1] / Sd[-1, -1], rel=1e-9)


@pytest.mark.parametrize("name,make,options", CASES, ids=[c[0] for c in CASES])
def test_pcg_solves_the_dense_system(name, make, options):
    packed, control = problem(make(), options)
    control.max_cg = 256
    control.relative_tolerance = 1e-10
    lam = 1e-2
    x, st = O.pcg(control, packed, lam, 0.02)
    assert st["converged"] == 1 and st["failed"] == 0
    m = control.cameras * 6 + control.shared_focal
    _, _, _, _, blocks = O.schur_probe(control, packed, lam, 0.02, np.zeros(m), blocks=True)
    Sd, b, rows, _ = dense_system(packed, control, blocks, lam)
    # the recurrence residual met the tolerance; the true residual is within a few times it
    assert np.linalg.norm(Sd @ x[rows] - b) <= 1e-8 * np.linalg.norm(b)
    assert np.max(np.abs(x[rows] - np.linalg.solve(Sd, b))) <= 1e-6 * np.max(np.abs(x[rows]))
