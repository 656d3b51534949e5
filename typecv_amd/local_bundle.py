"""Local and partitioned bundle problems within a map (src/sfm/bundle/local-bundle.ts).

The mapper (src/sfm/mapper/mapper.ts:873-908) refines a window of recent views: it extracts a local
problem that keeps every observation of each landmark the window sees and two fixed boundary
cameras as the gauge, solves it, and merges only the window's poses and the selected coordinates
back. Disconnected observation graphs are split into connected problems first. These are the same
operations on the reference's object model; `solve_local_bundles` solves many extracted problems in
one batched engine call (the block-diagonal batch of include/typecv_ba.h).
"""
from __future__ import annotations

import math
from typing import Mapping, Sequence

import numpy as np

from . import bundle as B


def _center(pose: Mapping) -> np.ndarray:
    """`cameraCenter`, src/geometry/pose.ts:34-39: -R^T t."""
    R = np.asarray(pose["R"], dtype=np.float64).reshape(3, 3)
    return -(R.T @ np.asarray(pose["t"], dtype=np.float64))


def _widest_pair(centers: Sequence[np.ndarray], floor: float):
    """The first pair (i < j, row-major order) of largest distance above `floor`, or None."""
    best, pair = floor, None
    for i in range(len(centers)):
        for j in range(i + 1, len(centers)):
            d = float(np.linalg.norm(centers[i] - centers[j]))
            if d > best:
                best, pair = d, (i, j)
    return pair, best


def extract_local_bundle(model: Mapping, variable_view_ids: Sequence[int]):
    """`extractLocalBundle`, local-bundle.ts:130-190: the landmarks observed by a variable view,
    with every one of their observations, and the views they touch; the gauge is the widest pair of
    fixed (boundary) views, the anchor first. None when the window is empty or repeats a view, a
    variable view observes nothing, fewer than two boundary views remain, the widest boundary
    baseline is not above 1e-10, or the local graph is disconnected."""
    variable = set(variable_view_ids)
    if not variable or len(variable) != len(variable_view_ids):
        return None
    landmarks = [p for p in model["landmarks"]
                 if any(o["viewId"] in variable for o in p["observations"])]
    neighbors: dict[int, list[int]] = {}
    for p in landmarks:
        first = p["observations"][0]["viewId"]
        for o in p["observations"]:
            neighbors.setdefault(o["viewId"], [])
            if o["viewId"] != first:
                neighbors[o["viewId"]].append(first)
                neighbors[first].append(o["viewId"])
    if any(i not in neighbors for i in variable_view_ids):
        return None
    views = [v for v in model["views"] if v["id"] in neighbors]
    boundary = [v for v in views if v["id"] not in variable]
    if len(views) != len(neighbors) or len(boundary) < 2:
        return None
    pair, baseline = _widest_pair([_center(v["pose"]) for v in boundary], 0.0)
    if pair is None or not baseline > 1e-10:
        return None
    anchor = boundary[pair[0]]["id"]
    connected, pending = {anchor}, [anchor]
    k = 0
    while k < len(pending):
        for i in neighbors[pending[k]]:
            if i not in connected:
                connected.add(i)
                pending.append(i)
        k += 1
    if len(connected) != len(views):
        return None
    return {"reconstruction": {"views": views, "landmarks": landmarks},
            "variableViewIds": list(variable_view_ids), "anchorViewId": anchor,
            "scaleViewId": boundary[pair[1]]["id"]}


def merge_local_bundle(model: Mapping, problem: Mapping, refined: Mapping) -> dict:
    """`mergeLocalBundle`, local-bundle.ts:193-208: only the variable views' poses and the selected
    landmarks' coordinates come from `refined`; cameras, observations and everything else stay."""
    variable = set(problem["variableViewIds"])
    selected = {p["trackId"] for p in problem["reconstruction"]["landmarks"]}
    views = {v["id"]: v for v in refined["views"]}
    points = {p["trackId"]: p for p in refined["landmarks"] if p["trackId"] in selected}
    return {
        "views": [{**v, "pose": views[v["id"]]["pose"]} if v["id"] in variable else v
                  for v in model["views"]],
        "landmarks": [{**p, "xyz": points[p["trackId"]]["xyz"]} if p["trackId"] in points else p
                      for p in model["landmarks"]],
    }


def plan_connected_bundles(model: Mapping, options: Mapping) -> dict:
    """`planConnectedBundles`, local-bundle.ts:17-110: split the views into connected components of
    the observation graph. A connected map keeps the original problem and options (the same
    objects). Otherwise each component with landmarks gets its own problem gauged by its widest
    pair of fixed boundary views (distance above 1e-10), with those two and any view outside the
    window fixed, shared focal refinement off; components without landmarks or without such a pair
    are deferred."""
    parent = {v["id"]: v["id"] for v in model["views"]}

    def root(i):
        r = i
        while parent[r] != r:
            r = parent[r]
        while i != r:
            nxt = parent[i]
            parent[i] = r
            i = nxt
        return r

    for p in model["landmarks"]:
        first = root(p["observations"][0]["viewId"])
        for o in p["observations"]:
            parent[root(o["viewId"])] = first
    groups: dict[int, list[int]] = {}
    for v in model["views"]:
        groups.setdefault(root(v["id"]), []).append(v["id"])
    if len(groups) == 1:
        return {"problems": [{"reconstruction": model, "options": options}]}
    points: dict[int, list] = {}
    for p in model["landmarks"]:
        points.setdefault(root(p["observations"][0]["viewId"]), []).append(p)
    by_id = {v["id"]: v for v in model["views"]}
    window = options.get("variableViewIds")
    variable = None if window is None else set(window)
    problems, fixed, deferred = [], [], []
    for r, ids in groups.items():
        landmarks = points.get(r, [])
        boundary = [i for i in ids if variable is None or i not in variable]
        pair, _ = _widest_pair([_center(by_id[i]["pose"]) for i in boundary], 1e-10)
        if not landmarks or pair is None:
            deferred.extend(ids)
            continue
        anchor, scale = boundary[pair[0]], boundary[pair[1]]
        fixed.extend(i for i in ids
                     if i in (anchor, scale) or (variable is not None and i not in variable))
        problems.append({
            "reconstruction": {"views": [by_id[i] for i in ids], "landmarks": landmarks},
            "options": {**options, "anchorViewId": anchor, "scaleViewId": scale,
                        "variableViewIds": [i for i in ids if i not in (anchor, scale)
                                            and (variable is None or i in variable)],
                        "refineSharedFocal": False, "focalGroups": None},
        })
    return {"problems": problems,
            "partition": {"components": list(groups.values()), "fixedViewIds": fixed,
                          "deferredViewIds": deferred,
                          "sharedFocalDeferred": bool(options.get("refineSharedFocal", False))}}


def merge_bundle_refinements(model: Mapping, refined: Sequence[Mapping]) -> dict:
    """`mergeBundleRefinements`, local-bundle.ts:112-121: solved views and landmarks replace theirs;
    isolated or deferred cameras and unrelated points stay."""
    views = {v["id"]: v for m in refined for v in m["views"]}
    points = {p["trackId"]: p for m in refined for p in m["landmarks"]}
    return {"views": [views.get(v["id"], v) for v in model["views"]],
            "landmarks": [points.get(p["trackId"], p) for p in model["landmarks"]]}


def local_problem_packed(problem: Mapping, options: Mapping | None = None):
    """The packed arrays and control of an extracted local problem, as `bundle_adjust_gpu` builds
    them for {variableViewIds, anchorViewId, scaleViewId} plus `options` (fixed intrinsics)."""
    opts = {**(options or {}), "variableViewIds": problem["variableViewIds"],
            "anchorViewId": problem["anchorViewId"], "scaleViewId": problem["scaleViewId"]}
    prob = B.bundle_problem(problem["reconstruction"], opts, lambda C: min(96, max(24, 12 * C)))
    packed = B.pack_bundle_scene(prob["input"], prob["byId"], prob["active"])
    control = B.make_control(
        packed.C, packed.P, packed.M, directional=packed.P - prob["finitePoints"],
        anchor=prob["byId"][prob["anchorId"]], scale_view=prob["byId"][prob["scaleId"]],
        normalize_scale=prob["fixed"] == 1, iterations=prob["iterations"],
        max_cg=prob["maxCg"], huber=prob["huber"], tolerance=prob["tolerance"],
        precision=opts.get("precision", "fp64"), stall_tolerance=B.bundle_stall_tolerance(opts))
    return prob, packed, control


def solve_local_bundles(device: "B.HipDevice", problems: Sequence[Mapping],
                        options: Mapping | None = None) -> list:
    """Solve extracted local problems (fixed intrinsics) together in one batched engine call. The
    bundles share their options; their PCG budget is the largest default of the batch. Returns, for
    each problem, its refined reconstruction (views and landmarks in the problem's order) and
    diagnostics; `merge_local_bundle` takes each back into its map."""
    packs = [local_problem_packed(p, options) for p in problems]
    controls = [c for _, _, c in packs]
    budget = max(c.max_cg for c in controls)
    for c in controls:
        c.max_cg = budget
    outs, report, finish = B.solve_batch(device, controls, [pk for _, pk, _ in packs])
    results = []
    for (prob, packed, control), values, ms in zip(packs, outs, finish):
        base = packed.C * 12 + packed.P * 3
        B.check_bundle_output(values, base)
        changed = values[base + B.BUNDLE_STATE["accepted"]] > 0
        model = prob["input"]
        views = [{**v, "pose": {"R": values[c * 12: c * 12 + 9].copy(),
                                "t": values[c * 12 + 9: c * 12 + 12].copy()}}
                 if changed and prob["active"][c] else v for c, v in enumerate(model["views"])]
        landmarks = [{**p, "xyz": values[packed.C * 12 + 3 * i: packed.C * 12 + 3 * i + 3].copy()}
                     if changed else p for i, p in enumerate(model["landmarks"])]
        results.append({
            "reconstruction": {"views": views, "landmarks": landmarks},
            "diagnostics": {
                "initialCost": float(values[base + B.BUNDLE_STATE["initialCost"]]),
                "finalCost": float(values[base + B.BUNDLE_STATE["cost"]]),
                "acceptedSteps": float(values[base + B.BUNDLE_STATE["accepted"]]),
                "rejectedSteps": float(values[base + B.BUNDLE_STATE["rejected"]]),
                "observations": packed.M, "loopMs": float(ms)},
        })
    return results
