"""GPU derivative entries (tcv_ba_jacobian, tcv_ba_polish, tcv_ba_backward) against the oracle's
linearisation, its chart gradient, the dense FD-Hessian implicit derivative (oracle/adjoint.py,
pinned on the CPU by test_oracle_adjoint.py) and finite differences of re-solved problems.

Bars: Jacobian rows equal the oracle's to 1e-12 relative and central differences of the oracle's
residuals (h = 1e-6, the step of the reference's geometry tests) to 1e-5 relative; the polished
scene's oracle gradient at most 1e-8 of the forward output's, the anchor pose unchanged and the
gauge baseline kept; the backward's pixel cotangent within 1e-6 (relative to its largest entry) of
the dense implicit derivative, and within 1e-5 relative of central differences (h = 1e-4 px) of
re-polished solutions, gauge maps and directional points included; the input-focal cotangent within
1e-6 of the dense -lambda . dg/d(log f_c) (differences of the oracle gradient) and 1e-5 of central
differences (h = 1e-6 in log f_c) of re-polished solutions, on the shared and the fixed route.
"""
import dataclasses

import numpy as np
import pytest

from oracle import adjoint as A
from oracle import oracle as O
from typecv_amd import bundle as B
from typecv_amd import synthetic as S
from problems import packed_problem

pytestmark = pytest.mark.gpu


def joint_problem(fixed_two: bool, outliers: bool = True, views: int = 5, points: int = 30):
    model = S.joint_focal_fixture(views=views, points=points, focal_scale=0.97, noise=0.6)["input"]
    if outliers:
        model["landmarks"][3]["observations"][1]["pixel"][0] += 25.0
        model["landmarks"][7]["observations"][2]["pixel"][1] -= 18.0
    options = {"iterations": 30, "refineSharedFocal": True, "huberPx": 2.0,
               "maxCgIterations": 256, "cgRelativeTolerance": 1e-10}
    if fixed_two:
        options["variableViewIds"] = [v["id"] for v in model["views"][2:]]
    return packed_problem(model, options)


def directional_problem():
    m = S.bundle_fixture(count=20)
    R0 = np.asarray(m["views"][0]["pose"]["R"]).reshape(3, 3)
    d = R0.T @ np.array([0.05, -0.02, 1.0])
    obs = []
    for j, v in enumerate(m["views"]):
        R = np.asarray(v["pose"]["R"]).reshape(3, 3)
        q = R @ d
        k = v["camera"]
        x, y = q[0] / q[2], q[1] / q[2]
        r2 = x * x + y * y
        f = 1 + k["k1"] * r2 + k["k2"] * r2 * r2
        obs.append({"viewId": v["id"], "featureId": 700,
                    "pixel": [k["fx"] * x * f + k["cx"] + 0.2 * (-1) ** j,
                              k["fy"] * y * f + k["cy"] - 0.1 * j]})
    extra = [{"trackId": 999, "direction": d * 3, "observations": obs}]
    return packed_problem(m, {"iterations": 20, "huberPx": 1.0, "directionalConstraints": extra,
                              "cgRelativeTolerance": 1e-10})


def solved(device, packed, control):
    out, _ = B.solve_packed(device, control, packed)
    return B.at_solution(packed, out), B.output_log_focal(control, out)


def polished(device, packed, control, log_focal):
    out, rep = B.polish(device, control, packed, log_focal, iterations=4, tolerance=1e-13)
    return B.at_solution(packed, out), B.output_log_focal(control, out), out, rep


def test_jacobian_matches_oracle_and_central_differences(device):
    print("\n[adjoint] jacobian", flush=True)
    packed, control = joint_problem(fixed_two=False)
    lf = 0.02
    J = B.jacobian(device, control, packed, lf)
    C, P, M = packed.C, packed.P, packed.M
    _, _, _, _, blocks = O.schur_probe(control, packed, 0.0, lf, np.zeros(6 * C + 1), blocks=True)
    b, focal = blocks["J"], blocks["focal"]
    jf = focal[6 * C + 6:].reshape(M, 2)
    ref = np.concatenate([b[:, 18:20], b[:, 0:12], b[:, 12:18], jf, b[:, 20:21]], axis=1)
    assert np.max(np.abs(J - ref) / np.maximum(1.0, np.abs(ref))) <= 1e-12
    # central differences of the oracle's residuals along chart parameters
    cams = packed.index[:M].astype(np.int64)
    pts = packed.index[M: 2 * M].astype(np.int64)
    n = 6 * C + 3 * P + 1
    rng = np.random.default_rng(2)
    params = list(range(6, 18)) + list(6 * C + rng.choice(3 * P, 12, replace=False)) + [n - 1]
    h = 1e-6
    for k in params:
        e = np.zeros(n)
        e[k] = h
        sp, fp = A.move(control, packed.scene, lf, e)
        sm, fm = A.move(control, packed.scene, lf, -e)
        rp = O.schur_probe(control, dataclasses.replace(packed, scene=sp), 0.0, fp,
                           np.zeros(6 * C + 1), blocks=True)[4]["J"][:, 18:20]
        rm = O.schur_probe(control, dataclasses.replace(packed, scene=sm), 0.0, fm,
                           np.zeros(6 * C + 1), blocks=True)[4]["J"][:, 18:20]
        fd = (rp - rm) / (2 * h)
        if k < 6 * C:
            c, j = divmod(k, 6)
            rows = cams == c
            an = J[rows][:, [2 + j, 8 + j]]
        elif k < n - 1:
            p, j = divmod(k - 6 * C, 3)
            rows = pts == p
            an = J[rows][:, [14 + j, 17 + j]]
        else:
            rows = np.ones(M, dtype=bool)
            an = J[:, 20:22]
        assert np.max(np.abs(fd[rows] - an)) <= 1e-5 * max(1.0, np.max(np.abs(an))), k


@pytest.mark.parametrize("fixed_two", [True, False])
def test_polish_reaches_a_stationary_point(device, fixed_two):
    print(f"\n[adjoint] polish fixed_two={fixed_two}", flush=True)
    packed, control = joint_problem(fixed_two=fixed_two, views=6, points=60)
    at, lf = solved(device, packed, control)
    rows = A.variable_rows(control, packed.index)
    g0 = np.linalg.norm(O.gradient(control, at.scene, packed.index, lf)[0][rows])
    norms = []
    for steps in (1, 2):
        o, _ = B.polish(device, control, at, lf, iterations=steps, tolerance=1e-13)
        s = B.at_solution(at, o)
        norms.append(np.linalg.norm(
            O.gradient(control, s.scene, packed.index, B.output_log_focal(control, o))[0][rows]))
    pol, lf1, out, rep = polished(device, at, control, lf)
    g1 = np.linalg.norm(O.gradient(control, pol.scene, packed.index, lf1)[0][rows])
    print(f"gradient {g0:.3e} -> {norms[0]:.3e} -> {norms[1]:.3e} -> {g1:.3e}", flush=True)
    # Newton on the exact Hessian converges quadratically (a Hessian off by a relative e would
    # leave a fraction ~e after the first step), down to the rounding floor of the gradient itself:
    # the engine and the oracle round the ~500 px projections differently, ~1e-13 px per residual
    assert norms[0] <= 1e-2 * g0, (g0, norms)
    # floor: |J| ~ 1e3 px/unit times ~1e-13 px over ~100 residuals per row, ~1e-9 in norm here
    floor = 2e-8
    assert norms[1] <= max(1e-2 * norms[0], floor), (g0, norms)
    assert g1 <= max(1e-6 * g0, floor), (g0, g1)
    base = packed.C * 12 + packed.P * 3
    assert out[base + 10] >= 1  # Newton steps taken
    anchor = control.anchor
    assert np.array_equal(pol.scene[12 * anchor: 12 * anchor + 12],
                          at.scene[12 * anchor: 12 * anchor + 12])
    if control.normalize_scale:
        def center(s, k):
            R = s[12 * k: 12 * k + 9].reshape(3, 3)
            return -R.T @ s[12 * k + 9: 12 * k + 12]
        b0 = np.linalg.norm(center(at.scene, control.scale_view) - center(at.scene, anchor))
        b1 = np.linalg.norm(center(pol.scene, control.scale_view) - center(pol.scene, anchor))
        assert abs(b1 - b0) <= 1e-12 * b0


def test_backward_matches_dense_implicit_derivative(device):
    print("\n[adjoint] backward vs dense", flush=True)
    packed, control = joint_problem(fixed_two=True)
    at, lf = solved(device, packed, control)
    pol, lf, _, _ = polished(device, at, control, lf)
    C, P = packed.C, packed.P
    rng = np.random.default_rng(9)
    # a cotangent on the points and the focal scale: the chart and the outputs coincide there
    chart = np.zeros(6 * C + 3 * P + 1)
    chart[6 * C:] = rng.normal(size=3 * P + 1)
    cot = np.zeros(12 * C + 3 * P + 1)
    cot[12 * C:] = chart[6 * C:]
    pixels, focal, adjoint, rep = B.backward(device, control, pol, lf, cot, tolerance=1e-12)
    ref, lam = A.dense_vjp(control, pol, pol.scene, lf, chart, control.huber)
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(pixels.reshape(-1) - ref)) <= 1e-6 * scale
    rows = A.variable_rows(control, packed.index)
    assert np.max(np.abs(adjoint[rows] - lam[rows])) <= 1e-6 * np.max(np.abs(lam))
    # the input-focal cotangent against -lambda . dg/d(eps_c) by differences of the oracle gradient
    ref_focal = A.dense_focal_vjp(control, pol.scene, packed.index, lf, lam)
    print(f"  focal: max |engine - dense| {np.max(np.abs(focal - ref_focal)):.2e} of "
          f"{np.max(np.abs(ref_focal)):.2e}", flush=True)
    assert np.max(np.abs(focal - ref_focal)) <= 1e-6 * np.max(np.abs(ref_focal))


@pytest.mark.parametrize("case", ["joint_focal_one_fixed", "directional"])
def test_backward_matches_differences_of_polished_solutions(device, case):
    print(f"\n[adjoint] backward vs FD ({case})", flush=True)
    if case == "directional":
        packed, control = directional_problem()
    else:
        packed, control = joint_problem(fixed_two=False)
    at, lf = solved(device, packed, control)
    pol, lf, _, _ = polished(device, at, control, lf)
    C, P, M = packed.C, packed.P, packed.M
    sh = 1 if control.shared_focal else 0
    n_out = 12 * C + 3 * P + sh
    rng = np.random.default_rng(21)
    cot = rng.normal(size=n_out)
    pixels, focal, _, rep = B.backward(device, control, pol, lf, cot, tolerance=1e-12)
    grad = pixels.reshape(-1)
    h = 1e-4
    pix = 18 * C + 3 * P
    for trial in range(5):
        du = np.zeros(2 * M)
        du[rng.integers(0, 2 * M, size=4)] = rng.normal(size=4)
        outs = []
        for sign in (1, -1):
            scene = pol.scene.copy()
            scene[pix:] += sign * h * du
            moved = dataclasses.replace(pol, scene=scene)
            out, _ = B.polish(device, control, moved, lf, iterations=4, tolerance=1e-13)
            outs.append(out)
        diff = (outs[0][:n_out - sh] - outs[1][:n_out - sh]) / (2 * h)
        fd = cot[:n_out - sh] @ diff
        if sh:
            fd += cot[-1] * (B.output_log_focal(control, outs[0])
                             - B.output_log_focal(control, outs[1])) / (2 * h)
        an = grad @ du
        assert abs(fd - an) <= 1e-5 * max(abs(fd), 1e-2 * np.max(np.abs(grad))), (trial, fd, an)
    # the input-focal cotangent: each camera's fx, fy scaled by e^(+-h)
    h = 1e-6
    base = 12 * C + 3 * P
    fds = []
    for c in range(C):
        outs = []
        for sign in (1, -1):
            scene = pol.scene.copy()
            scene[base + 6 * c: base + 6 * c + 2] *= np.exp(sign * h)
            out, _ = B.polish(device, control, dataclasses.replace(pol, scene=scene), lf,
                              iterations=4, tolerance=1e-13)
            outs.append(out)
        fd = cot[:n_out - sh] @ ((outs[0][:n_out - sh] - outs[1][:n_out - sh]) / (2 * h))
        if sh:
            fd += cot[-1] * (B.output_log_focal(control, outs[0])
                             - B.output_log_focal(control, outs[1])) / (2 * h)
        fds.append(fd)
    fds = np.array(fds)
    print(f"  focal: max |adjoint - fd| {np.max(np.abs(focal - fds)):.2e} of "
          f"{np.max(np.abs(fds)):.2e}", flush=True)
    assert np.max(np.abs(focal - fds)) <= 1e-5 * np.max(np.abs(fds))
