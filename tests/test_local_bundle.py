"""Local and partitioned bundle problems (typecv_amd/local_bundle.py), the reference's
tests/local-bundle.test.ts and tests/bundle-partitions.test.ts on the restated tests/scene.ts
generator, with the CPU oracle as the solver where the reference calls its CPU `bundleAdjust`."""
import copy

import numpy as np
import pytest

from oracle import oracle as O
from typecv_amd import bundle as B
from typecv_amd import local_bundle as L
from typecv_amd import synthetic as S


def local_scene():
    """localScene, tests/local-bundle.test.ts:10-29."""
    model, rng = S.reference_scene(count=48, views=7, noise=0.1)
    variable = {40, 50}
    views = [{**v, "pose": S.update_pose(v["pose"], [0.001, -0.002, 0.001, 0.02, -0.01, 0.01])}
             if v["id"] in variable else v for v in model["views"]]
    landmarks = []
    for i, p in enumerate(model["landmarks"]):
        keep = [20, 30, 40, 50] if i < 24 else [10, 20, 60, 70]
        xyz = p["xyz"] + np.array([(rng() - 0.5) * 0.08 for _ in range(3)]) if i < 24 else p["xyz"]
        landmarks.append({**p, "xyz": xyz,
                          "observations": [o for o in p["observations"] if o["viewId"] in keep]})
    return {"views": views, "landmarks": landmarks}


def oracle_solve(model, options):
    """The operator's packing of (model, options), solved by the oracle; the refined model."""
    problem = B.bundle_problem(model, options, lambda C: min(96, max(24, 12 * C)))
    packed = B.pack_bundle_scene(problem["input"], problem["byId"], problem["active"])
    control = B.make_control(
        packed.C, packed.P, packed.M, anchor=problem["byId"][problem["anchorId"]],
        scale_view=problem["byId"][problem["scaleId"]], normalize_scale=problem["fixed"] == 1,
        iterations=problem["iterations"], max_cg=problem["maxCg"], huber=problem["huber"],
        tolerance=problem["tolerance"])
    out, _ = O.solve(control, packed)
    base = packed.C * 12 + packed.P * 3
    views = [{**v, "pose": {"R": out[12 * c: 12 * c + 9], "t": out[12 * c + 9: 12 * c + 12]}}
             if problem["active"][c] else v for c, v in enumerate(problem["input"]["views"])]
    landmarks = [{**p, "xyz": out[12 * packed.C + 3 * i: 12 * packed.C + 3 * i + 3]}
                 for i, p in enumerate(problem["input"]["landmarks"])]
    return {"views": views, "landmarks": landmarks}, out[base + 8], out[base + 6], packed.M


def huber_cost(model, huber=2.0):
    views = {v["id"]: v for v in model["views"]}
    total = 0.0
    for p in model["landmarks"]:
        for o in p["observations"]:
            v = views[o["viewId"]]
            R = np.asarray(v["pose"]["R"]).reshape(3, 3)
            q = R @ np.asarray(p["xyz"]) + np.asarray(v["pose"]["t"])
            k = v["camera"]
            e = np.hypot(k["fx"] * q[0] / q[2] + k["cx"] - o["pixel"][0],
                         k["fy"] * q[1] / q[2] + k["cy"] - o["pixel"][1])
            total += e * e if e <= huber else 2 * huber * e - huber * huber
    return total


def test_extraction_keeps_every_selected_observation_and_connected_cameras():
    model = local_scene()
    snapshot = copy.deepcopy(model)
    local = L.extract_local_bundle(model, [40, 50])
    assert local is not None
    B.validate_reconstruction(local["reconstruction"])
    assert [v["id"] for v in local["reconstruction"]["views"]] == [20, 30, 40, 50]
    assert len(local["reconstruction"]["landmarks"]) == 24
    assert local["anchorViewId"] == 20 and local["scaleViewId"] == 30
    assert local["anchorViewId"] not in local["variableViewIds"]
    for p in local["reconstruction"]["landmarks"]:
        assert p is model["landmarks"][p["trackId"]]
        assert len(p["observations"]) == 4
    _, _, _, M = oracle_solve(local["reconstruction"], {
        "iterations": 0, "variableViewIds": local["variableViewIds"],
        "anchorViewId": local["anchorViewId"], "scaleViewId": local["scaleViewId"]})
    assert M == 96
    np.testing.assert_equal(model, snapshot)


def test_local_solve_improves_and_merge_preserves_unrelated_geometry():
    model = local_scene()
    local = L.extract_local_bundle(model, [40, 50])
    refined, initial, final, _ = oracle_solve(local["reconstruction"], {
        "iterations": 20, "huberPx": 2, "variableViewIds": local["variableViewIds"],
        "anchorViewId": local["anchorViewId"], "scaleViewId": local["scaleViewId"]})
    assert final < initial
    merged = L.merge_local_bundle(model, local, refined)
    assert len(merged["views"]) == len(model["views"])
    assert len(merged["landmarks"]) == len(model["landmarks"])
    for old, new in zip(model["views"], merged["views"]):
        assert new["camera"] is old["camera"]
        if old["id"] not in (40, 50):
            assert new is old
    for i, (old, new) in enumerate(zip(model["landmarks"], merged["landmarks"])):
        assert new["observations"] is old["observations"]
        if i >= 24:
            assert new is old
    B.validate_reconstruction(merged)
    assert huber_cost(merged) <= huber_cost(model)


def test_merge_takes_only_selected_coordinates_and_variable_poses():
    model = local_scene()
    local = L.extract_local_bundle(model, [40, 50])
    identity = {"R": np.eye(3).reshape(9), "t": np.zeros(3)}
    changed = {
        "views": [{**v, "camera": {**v["camera"], "fx": v["camera"]["fx"] * 2}, "pose": identity}
                  for v in model["views"]],
        "landmarks": [{**p, "xyz": np.array([1.0, 2.0, 3.0]), "observations": []}
                      for p in model["landmarks"]],
    }
    merged = L.merge_local_bundle(model, local, changed)
    for old, new in zip(model["views"], merged["views"]):
        assert new["camera"] is old["camera"]
        assert new["pose"] is (identity if old["id"] in (40, 50) else old["pose"])
    for i, (old, new) in enumerate(zip(model["landmarks"], merged["landmarks"])):
        assert new["observations"] is old["observations"]
        np.testing.assert_array_equal(new["xyz"], [1, 2, 3] if i < 24 else old["xyz"])


def test_unsafe_gauges_disconnected_subsets_and_unconstrained_cameras_fall_back():
    model = local_scene()
    assert L.extract_local_bundle(model, []) is None
    assert L.extract_local_bundle(model, [40, 40]) is None
    assert L.extract_local_bundle(model, [40, 999]) is None
    ids = [v["id"] for v in model["views"]]
    assert L.extract_local_bundle(model, ids[1:]) is None
    assert L.extract_local_bundle(model, ids) is None
    coincident = {**model, "views": [{**v, "pose": model["views"][1]["pose"]} if v["id"] == 30
                                     else v for v in model["views"]]}
    assert L.extract_local_bundle(coincident, [40, 50]) is None
    disconnected = {**model, "landmarks": [
        {**p, "observations": [o for o in p["observations"]
                               if (i < 12 and o["viewId"] in (20, 40))
                               or (12 <= i < 24 and o["viewId"] in (30, 50)) or i >= 24]}
        for i, p in enumerate(model["landmarks"])]}
    B.validate_reconstruction(disconnected)
    assert L.extract_local_bundle(disconnected, [40, 50]) is None


def disconnected_map():
    """disconnected(), tests/bundle-partitions.test.ts:13-33."""
    model, _ = S.reference_scene(count=48, views=7, noise=0.15)
    return {"views": model["views"], "landmarks": [
        {**p, "xyz": p["xyz"] + 0.015,
         "observations": [o for o in p["observations"]
                          if o["viewId"] in ((10, 20, 30) if i < 24 else (40, 50, 60))]}
        for i, p in enumerate(model["landmarks"])]}


def test_connected_map_keeps_the_original_problem_and_options():
    model, _ = S.reference_scene(count=24, views=3)
    options = {"iterations": 2, "refineSharedFocal": True}
    plan = L.plan_connected_bundles(model, options)
    assert "partition" not in plan
    assert len(plan["problems"]) == 1
    assert plan["problems"][0]["reconstruction"] is model
    assert plan["problems"][0]["options"] is options


def test_disconnected_map_refines_within_carried_gauges():
    model = disconnected_map()
    snapshot = copy.deepcopy(model)
    with pytest.raises(B.RangeError, match="one connected observation graph"):
        B.bundle_problem(model, {}, lambda C: 24)
    plan = L.plan_connected_bundles(model, {"iterations": 8, "anchorViewId": 10,
                                            "scaleViewId": 20, "refineSharedFocal": True})
    part = plan["partition"]
    assert part["components"] == [[10, 20, 30], [40, 50, 60], [70]]
    assert part["deferredViewIds"] == [70]
    assert part["sharedFocalDeferred"] is True
    assert len(plan["problems"]) == 2
    results = []
    for p in plan["problems"]:
        o = p["options"]
        assert o["refineSharedFocal"] is False
        assert o["anchorViewId"] not in o["variableViewIds"]
        assert o["scaleViewId"] not in o["variableViewIds"]
        refined, initial, final, _ = oracle_solve(p["reconstruction"], o)
        assert final < initial
        results.append(refined)
    merged = L.merge_bundle_refinements(model, results)
    assert len(merged["views"]) == 7 and len(merged["landmarks"]) == 48
    assert merged["views"][6] is model["views"][6]
    by_id = {v["id"]: v for v in merged["views"]}
    for i in part["fixedViewIds"]:
        old = next(v for v in model["views"] if v["id"] == i)
        np.testing.assert_array_equal(by_id[i]["pose"]["R"], old["pose"]["R"])
        np.testing.assert_array_equal(by_id[i]["pose"]["t"], old["pose"]["t"])
    assert [v["camera"] for v in merged["views"]] == [v["camera"] for v in model["views"]]
    assert all(a["observations"] == b["observations"]
               for a, b in zip(merged["landmarks"], model["landmarks"]))
    np.testing.assert_equal(model, snapshot)


def test_partitions_without_two_fixed_boundary_poses_or_baseline_are_deferred():
    model = disconnected_map()
    local = L.plan_connected_bundles(model, {"variableViewIds": [30, 40, 50, 60]})
    assert len(local["problems"]) == 1
    assert local["partition"]["deferredViewIds"] == [40, 50, 60, 70]
    collapsed = {**model, "views": [{**v, "pose": model["views"][3]["pose"]}
                                    if v["id"] in (40, 50, 60) else v for v in model["views"]]}
    plan = L.plan_connected_bundles(collapsed, {})
    assert len(plan["problems"]) == 1
    assert plan["partition"]["deferredViewIds"] == [40, 50, 60, 70]
