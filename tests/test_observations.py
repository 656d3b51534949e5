"""Fixed-scene observation evaluation (`evaluateObservationsGpu`): the oracle's restatement of
the CPU tier's kernels (oracle/ba_oracle.c, tcv_oracle_observations) pinned to closed forms, and
the host request (packing and checks) against the reference's rules. The reference's own tests
compare two device implementations with each other on these cases
(tests/typeshade/indexed-observations.test.mjs); they hold no golden numbers, so the pins here are
the cases' construction: every pixel is the exact projection plus a known offset."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from typecv_amd import observations as OB
from typecv_amd.bundle import RangeError

from obs_cases import observation_cases, observation_model


def oracle_result(model, max_error=3.0, options=None, last_only=False):
    req = OB.observation_request(model, max_error, options or {}, last_only)
    flags, errors, summary = O.observations(
        req["cameras"], req["points_xyz"], req["pixels"], req["indices"][:2 * req["count"]],
        req["offsets"], max_error, req["depth"]["maxRelativeDepthStdDev"],
        req["depth"]["pixelNoiseStdDevPx"])
    return req, flags, errors, summary


def test_errors_are_the_constructed_offsets():
    """radial-tail-129: error = ((p + i) % 4) * 0.03 for every observation (the pixel is the exact
    projection plus that x offset), all 129 inliers, mean = the offsets' mean."""
    model = observation_model()
    req, flags, errors, summary = oracle_result(model)
    offsets = np.array([((p + i) % 4) * 0.03 for p in range(43) for i in range(3)])
    assert np.all(flags == 1)
    assert np.max(np.abs(errors - offsets)) < 1e-9
    assert summary[1] == 129
    assert abs(summary[0] / summary[1] - offsets.mean()) < 1e-10
    result = OB.unpack_observations(summary, flags, req["lengths"], req["count"], False)
    assert result["inlierCount"] == 129 and len(result["masks"]) == 43


def test_threshold_outlier_and_behind_camera():
    """mixed case: +40 px on one observation exceeds 3 px; a landmark at z = -10 is behind every
    camera (projection rejected); a 0.05 px threshold keeps only the offsets 0 and 0.03."""
    cases = {c["name"]: c for c in observation_cases()}
    req, flags, errors, summary = oracle_result(cases["mixed-inliers-behind-camera"]["model"])
    masks = OB.unpack_observations(summary, flags, req["lengths"], req["count"], False)["masks"]
    assert masks[1].tolist() == [1, 1, 0]
    assert masks[2].tolist() == [0, 0, 0]
    assert summary[1] == 15 - 1 - 3
    _, flags, errors, summary = oracle_result(observation_model(points=5), max_error=0.05)
    expected = [1 if ((p + i) % 4) * 0.03 <= 0.05 else 0 for p in range(5) for i in range(3)]
    assert flags.tolist() == expected


def _depth_relative_std(model, track, noise=1.0, h=1e-6):
    """Independent reference for stableDepth: the information A = sum J^T J of the track's
    projections with J = d(pixel)/dX by central differences, then noise sqrt(axis^T A^-1 axis) / z
    per observation (axis = the camera's optical axis, R's third row)."""
    views = {v["id"]: v for v in model["views"]}
    p = model["landmarks"][track]
    X = np.array(p["xyz"], dtype=float)

    def pix(view, Y):
        R = np.array(view["pose"]["R"], dtype=float).reshape(3, 3)
        q = R @ Y + np.array(view["pose"]["t"], dtype=float)
        k = view["camera"]
        x, y = q[0] / q[2], q[1] / q[2]
        r2 = x * x + y * y
        d = 1 + k["k1"] * r2 + k["k2"] * r2 * r2
        return np.array([k["fx"] * x * d + k["cx"], k["fy"] * y * d + k["cy"]]), q

    A = np.zeros((3, 3))
    for o in p["observations"]:
        v = views[o["viewId"]]
        J = np.zeros((2, 3))
        for c in range(3):
            e = np.zeros(3)
            e[c] = h * max(1.0, abs(X[c]))
            J[:, c] = (pix(v, X + e)[0] - pix(v, X - e)[0]) / (2 * e[c])
        A += J.T @ J
    rel = []
    for o in p["observations"]:
        v = views[o["viewId"]]
        axis = np.array(v["pose"]["R"], dtype=float)[6:9]
        rel.append(noise * math.sqrt(axis @ np.linalg.solve(A, axis)) / pix(v, X)[1][2])
    return max(rel)


@pytest.mark.parametrize("name", ["radial-tail-129-guarded", "far-depth-rejected",
                                  "long-track-65", "different-calibration-and-rotated-poses"])
def test_depth_guard_matches_the_information_bound(name):
    """The guarded cases: a track is rejected exactly when its largest relative depth standard
    deviation (from a finite-difference information matrix) exceeds the limit."""
    case = {c["name"]: c for c in observation_cases()}[name]
    model, options = case["model"], case["options"]
    req, flags, errors, summary = oracle_result(model, options=options)
    rejected = OB.unpack_observations(summary, flags, req["lengths"], req["count"],
                                      True)["depthRejected"]
    limit = options["maxRelativeDepthStdDev"]
    for h in range(len(model["landmarks"])):
        rel = _depth_relative_std(model, h)
        assert abs(rel - limit) > 1e-6 * limit  # no case on the boundary
        assert bool(rejected[h]) == (rel > limit), (h, rel)
    if name == "far-depth-rejected":
        assert rejected.tolist() == [1, 1, 1] and summary[1] == 0
    else:
        assert not rejected.any()


def test_unused_view_is_not_checked_against_gpu_bounds():
    case = {c["name"]: c for c in observation_cases()}["unused-view-outside-gpu-bounds"]
    req, flags, _, summary = oracle_result(case["model"])
    assert req["views"] == 3 and summary[1] == 6
    bad = observation_model(points=2)
    bad["views"][0]["camera"]["fx"] = 1e8
    with pytest.raises(RangeError, match="supported GPU geometry bounds"):
        OB.observation_request(bad, 3.0, {}, False)


def test_request_checks_and_packing():
    model = observation_model(points=4)
    for bad in (0.0, 1e-13, 2e6, math.inf, math.nan):
        with pytest.raises(RangeError, match="Invalid GPU observation threshold"):
            OB.observation_request(model, bad, {}, False)
    for options in ({"maxRelativeDepthStdDev": 0}, {"maxRelativeDepthStdDev": -1},
                    {"pixelNoiseStdDevPx": 0}, {"maxRelativeDepthStdDev": math.inf}):
        with pytest.raises(RangeError, match="depth sensitivity"):
            OB.observation_request(model, 3.0, options, False)
    # shuffled: camera records in first-observation order, indices (record, landmark)
    case = {c["name"]: c for c in observation_cases()}["shuffled-view-point-observation-indices"]
    req = OB.observation_request(case["model"], 3.0, case["options"], False)
    first = case["model"]["landmarks"][0]["observations"]
    order = [o["viewId"] for o in first]
    views = {v["id"]: v for v in case["model"]["views"]}
    for rec, vid in enumerate(order):
        assert req["cameras"][rec * 18 + 9] == views[vid]["pose"]["t"][0]
    assert req["indices"][:8].tolist() == [0, 0, 1, 0, 2, 0, 3, 0]
    assert req["offsets"].tolist() == [0, 4, 8, 12, 16, 20]


def test_continuation_packs_last_observations():
    """evaluateContinuedObservationsGpu without the guard: one observation per landmark (its last),
    the records of the views those observations use, in first-use order."""
    model = observation_model(points=5)
    req = OB.observation_request(model, 3.0, {}, True)
    assert req["count"] == 5 and req["views"] == 1 and req["lengths"] == [1] * 5
    _, flags, errors, summary = oracle_result(model, last_only=True)
    offsets = [((p + 2) % 4) * 0.03 for p in range(5)]
    assert np.max(np.abs(errors - offsets)) < 1e-9 and summary[1] == 5
