"""The operator `bundle_adjust_gpu` on the GPU engine: the reference tests' invariants
(tests/bundle-adjustment.test.ts:31-161, tests/typeshade/cpu.test.mjs:486-607,
tests/typeshade/focal.test.mjs:13-94), the result record, ownership and error conventions."""
import copy
import math

import numpy as np
import pytest

from oracle import oracle as O
from typecv_amd import _native
from typecv_amd import bundle as B
from typecv_amd import synthetic as S

pytestmark = pytest.mark.gpu


def coords(result):
    views = result["reconstruction"]["views"]
    return (np.concatenate([np.concatenate([v["pose"]["R"], v["pose"]["t"]]) for v in views]),
            np.array([p["xyz"] for p in result["reconstruction"]["landmarks"]]))


def test_fixture_converges_keeps_gauge_and_input(device):
    print("\n[operator] fixture", flush=True)
    model = S.bundle_fixture()
    snapshot = copy.deepcopy(model)
    result = B.bundle_adjust_gpu(device, model, {"iterations": 20})
    d = result["diagnostics"]
    assert d["finalCost"] < 0.05 * d["initialCost"] and d["finalRmsePx"] < 0.3
    assert d["observations"] == 160 and result["execution"]["backend"] == "hip"
    views = result["reconstruction"]["views"]
    assert np.array_equal(views[0]["pose"]["R"], model["views"][0]["pose"]["R"])
    assert np.array_equal(views[0]["pose"]["t"], model["views"][0]["pose"]["t"])
    c0 = -views[0]["pose"]["R"].reshape(3, 3).T @ views[0]["pose"]["t"]
    c1 = -views[1]["pose"]["R"].reshape(3, 3).T @ views[1]["pose"]["t"]
    assert abs(np.linalg.norm(c0 - c1) - d["baseline"]) < 1e-10
    for a, b in zip(model["landmarks"], snapshot["landmarks"]):  # input never mutated
        assert np.array_equal(a["xyz"], b["xyz"])


def test_zero_iterations_and_forced_failure_return_the_input(device):
    print("\n[operator] zero iterations / forced failure", flush=True)
    model = S.bundle_fixture()
    zero = B.bundle_adjust_gpu(device, model, {"iterations": 0})
    assert zero["diagnostics"]["finalCost"] == zero["diagnostics"]["initialCost"]
    failed = B.bundle_adjust_gpu(device, model, {"iterations": 1, "maxCgIterations": 1,
                                                 "cgRelativeTolerance": 1e-12})
    assert failed["diagnostics"]["acceptedSteps"] == 0
    assert failed["linearSolverFailures"] == 1
    for r in (zero, failed):
        for v, w in zip(r["reconstruction"]["views"], model["views"]):
            assert np.array_equal(v["pose"]["R"], w["pose"]["R"])
        for p, q in zip(r["reconstruction"]["landmarks"], model["landmarks"]):
            assert np.array_equal(p["xyz"], q["xyz"])


def test_device_detected_input_errors(device):
    print("\n[operator] input errors", flush=True)
    model = S.bundle_fixture()
    model["views"][1]["pose"] = {k: np.array(v) for k, v in model["views"][0]["pose"].items()}
    with pytest.raises(B.RangeError, match="nonzero baseline"):
        B.bundle_adjust_gpu(device, model, {})
    model = S.bundle_fixture()
    model["landmarks"][3]["xyz"] = np.array([0.0, 0.0, -50.0])
    with pytest.raises(B.RangeError, match="in front"):
        B.bundle_adjust_gpu(device, model, {})


def test_joint_focal_recovers_ground_truth(device):
    print("\n[operator] joint focal", flush=True)
    fixture = S.joint_focal_fixture()
    model = fixture["input"]
    result = B.bundle_adjust_gpu(device, model, {"iterations": 40, "refineSharedFocal": True,
                                                 "maxCgIterations": 256,
                                                 "cgRelativeTolerance": 1e-10})
    d = result["diagnostics"]
    assert d["finalCost"] < 1e-7 and d["finalRmsePx"] < 1e-4 and d["initialCost"] > 50000
    views = result["reconstruction"]["views"]
    assert abs(views[0]["camera"]["fx"] - 800) < 0.001
    scale = views[0]["camera"]["fx"] / model["views"][0]["camera"]["fx"]
    for v, w in zip(views, model["views"]):
        assert abs(v["camera"]["fx"] / w["camera"]["fx"] - scale) < 1e-12
        assert v["camera"]["k1"] == w["camera"]["k1"]


def test_bad_index_is_refused_before_any_solve(device):
    print("\n[engine] corrupted index", flush=True)
    packed, _ = S.ring_scene(seed=4, cameras=10, points=100)
    control = B.make_control(packed.C, packed.P, packed.M, iterations=2, max_cg=10)
    bad = copy.deepcopy(packed)
    lists = bad.section("cameraObservations")
    lists[[3, 4]] = lists[[4, 3]]  # breaks the ascending order of one camera list
    with pytest.raises(_native.EngineError, match="index"):
        B.solve_packed(device, control, bad)
    bad = copy.deepcopy(packed)
    bad.index[0] = 10_000  # a camera id past C
    with pytest.raises(_native.EngineError, match="index"):
        B.solve_packed(device, control, bad)
    bad_control = B.make_control(packed.C, packed.P, packed.M, iterations=2, max_cg=300)
    with pytest.raises(_native.EngineError, match="scene/options"):
        B.solve_packed(device, bad_control, packed)
    out, _ = B.solve_packed(device, control, packed)  # the context still works
    assert np.isfinite(out).all()


def test_release_frees_every_device_buffer(device):
    print("\n[engine] release", flush=True)
    B.bundle_adjust_gpu(device, S.bundle_fixture(), {"iterations": 2})
    assert device.statistics["liveBytes"] > 0
    device.drop()
    assert device.statistics["liveBytes"] == 0
    B.bundle_adjust_gpu(device, S.bundle_fixture(), {"iterations": 2})


def test_directional_constraints_solve(device):
    print("\n[operator] directional constraints", flush=True)
    m = S.bundle_fixture(count=20)
    R0 = np.asarray(m["views"][0]["pose"]["R"]).reshape(3, 3)
    d = R0.T @ np.array([0.05, -0.02, 1.0])
    obs = []
    for v in m["views"]:
        R = np.asarray(v["pose"]["R"]).reshape(3, 3)
        q = R @ d
        k = v["camera"]
        x, y = q[0] / q[2], q[1] / q[2]
        r2 = x * x + y * y
        f = 1 + k["k1"] * r2 + k["k2"] * r2 * r2
        obs.append({"viewId": v["id"], "featureId": 700, "pixel": [k["fx"] * x * f + k["cx"] + 0.2,
                                                                   k["fy"] * y * f + k["cy"] - 0.1]})
    extra = [{"trackId": 999, "direction": d * 3, "observations": obs}]
    result = B.bundle_adjust_gpu(device, m, {"iterations": 10, "directionalConstraints": extra})
    assert len(result["reconstruction"]["landmarks"]) == 20
    out = result["directionalConstraints"][0]["direction"]
    assert abs(np.linalg.norm(out) - 1) < 1e-12
    assert result["diagnostics"]["directionalObservations"] == 4
    assert result["diagnostics"]["pointParameters"] == 62
    # the same problem on the oracle
    prob = B.bundle_problem(m, {"iterations": 10, "directionalConstraints": extra},
                            lambda C: min(96, max(24, 12 * C)))
    packed = B.pack_bundle_scene(prob["input"], prob["byId"], prob["active"])
    control = B.make_control(packed.C, packed.P, packed.M, directional=1, iterations=10,
                             max_cg=prob["maxCg"])
    expected, _ = O.solve(control, packed)
    base = packed.C * 12 + packed.P * 3
    assert abs(result["diagnostics"]["finalCost"] - expected[base + 6]) <= 1e-9 * max(1, expected[base + 6])
    assert np.max(np.abs(out - expected[base - 3: base])) < 1e-7
