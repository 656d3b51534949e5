"""Fixed-scene observation evaluation on the HIP engine: the reference's `evaluateObservationsGpu`
and `evaluateContinuedObservationsGpu` (src/geometry/gpu/observations.ts:34-127), the mapper's
`'observations'` job (src/sfm/gpu/mapper.ts:451-462).

The request checks and the packing follow `observationRequest` (observations.ts:132-197): the
referenced views get camera records in the order they are first observed, observations stay in
landmark order, and a continuation packs only each landmark's last observation (unless the depth
guard asks for the full evaluation). The device projects, thresholds, filters and sums
(`tcv_obs_evaluate`); the masks come back in landmark order.
"""
from __future__ import annotations

import ctypes
import math
import time
from typing import Mapping

import numpy as np

from . import _native
from .bundle import HipDevice, RangeError, validate_gpu_camera, validate_reconstruction
from .packing import finite_gpu_array


def depth_sensitivity_controls(options: Mapping) -> dict:
    """`depthSensitivityControls`, src/geometry/multiview.ts:26-36."""
    limit = options.get("maxRelativeDepthStdDev")
    noise = options.get("pixelNoiseStdDevPx", 1)
    if noise is None:
        noise = 1

    def finite(x):
        return isinstance(x, (int, float, np.floating, np.integer)) and math.isfinite(x)

    if (limit is not None and (not finite(limit) or limit <= 0)) or not finite(noise) or noise <= 0:
        raise RangeError("Invalid conditional depth sensitivity options")
    return {"maxRelativeDepthStdDev": float(limit or 0.0), "pixelNoiseStdDevPx": float(noise)}


def _supported_scene_coordinate(value) -> bool:
    """`supportedSceneCoordinate`, observations.ts:26-31."""
    if not math.isfinite(value):
        return False
    magnitude = abs(value)
    return magnitude <= 1e9 and (value == 0 or magnitude >= 1e-30)


def observation_request(model: Mapping, max_error_px: float, options: Mapping,
                        last_only: bool) -> dict:
    """`observationRequest`, observations.ts:132-197: checks and host packing. Raises RangeError
    with the reference's messages before any device work."""
    depth = depth_sensitivity_controls(options)
    guarded = depth["maxRelativeDepthStdDev"] > 0
    validate_reconstruction(model)
    if (not finite_gpu_array(np.array([depth["maxRelativeDepthStdDev"],
                                       depth["pixelNoiseStdDevPx"]]))
            or not math.isfinite(max_error_px) or max_error_px < 1e-12 or max_error_px > 1e6):
        raise RangeError("Invalid GPU observation threshold")
    landmarks = model["landmarks"]
    lengths = [1 if last_only else len(p["observations"]) for p in landmarks]
    count = sum(lengths)
    original_used: dict = {}
    for p in landmarks:
        for o in p["observations"]:
            if o["viewId"] not in original_used:
                original_used[o["viewId"]] = len(original_used)
            # skipping a projection does not relax the original context's coordinate bounds
            if last_only and (not _supported_scene_coordinate(o["pixel"][0])
                              or not _supported_scene_coordinate(o["pixel"][1])):
                raise RangeError("Invalid GPU scene coordinate bounds")
    used = original_used
    if last_only:
        used = {}
        for p in landmarks:
            vid = p["observations"][-1]["viewId"]
            if vid not in used:
                used[vid] = len(used)
    if count >= 2**32 or len(landmarks) >= 2**32:
        raise RangeError("GPU observation evaluation exceeds budget")
    cameras = np.zeros(max(1, 18 * len(used)))
    points = np.zeros(max(1, 3 * len(landmarks)))
    pixels = np.zeros(max(1, 2 * count))
    indices = np.zeros(max(2, 2 * count), dtype=np.uint32)
    views = {v["id"]: v for v in model["views"]}
    for vid in original_used:
        v = views[vid]
        k = validate_gpu_camera(v["camera"])
        row = np.concatenate([np.asarray(v["pose"]["R"], dtype=np.float64).reshape(-1),
                              np.asarray(v["pose"]["t"], dtype=np.float64).reshape(-1),
                              [k["fx"], k["fy"], k["cx"], k["cy"], k["k1"], k["k2"]]])
        if not finite_gpu_array(row):
            raise RangeError("Invalid GPU scene coordinate bounds")
        index = used.get(vid)
        if index is not None:
            cameras[index * 18:(index + 1) * 18] = row
    at = 0
    for point_index, p in enumerate(landmarks):
        xyz = np.asarray(p["xyz"], dtype=np.float64).reshape(-1)
        if not finite_gpu_array(xyz):
            raise RangeError("Invalid GPU scene coordinate bounds")
        points[point_index * 3:point_index * 3 + 3] = xyz
        obs = p["observations"]
        for o in (obs[-1:] if last_only else obs):
            pixels[2 * at:2 * at + 2] = o["pixel"]
            indices[2 * at] = used[o["viewId"]]
            indices[2 * at + 1] = point_index
            at += 1
    if not finite_gpu_array(pixels):
        raise RangeError("Invalid GPU scene coordinate bounds")
    offsets = None
    if guarded:
        offsets = np.zeros(len(lengths) + 1, dtype=np.uint32)
        offsets[1:] = np.cumsum(lengths)
    return {"depth": depth, "guarded": guarded, "lengths": lengths, "count": count,
            "views": len(used), "points": len(landmarks), "cameras": cameras, "points_xyz": points,
            "pixels": pixels, "indices": indices, "offsets": offsets}


def unpack_observations(summary: np.ndarray, flags: np.ndarray, lengths, count: int,
                        guarded: bool) -> dict:
    """`unpackObservations`, observations.ts:285-313: the masks in landmark order and the summary."""
    masks, at = [], 0
    for n in lengths:
        masks.append(flags[at:at + n].astype(np.uint8))
        at += n
    result = {}
    if guarded:
        result["depthRejected"] = flags[count:count + len(lengths)].astype(np.uint8)
    inliers = float(summary[1])
    result.update({"masks": masks, "inlierCount": inliers,
                   "meanErrorPx": float(summary[0]) / inliers if inliers > 0 else math.inf})
    return result


def evaluate_request(device: HipDevice, request: dict, max_error_px: float) -> dict:
    """The device evaluation of a packed request, unpacked as the reference returns it."""
    flags, summary, elapsed = evaluate_arrays(device, request, max_error_px)
    result = unpack_observations(summary, flags, request["lengths"], request["count"],
                                 request["guarded"])
    result["execution"] = {"backend": "hip", "elapsedMs": elapsed}
    return result


def evaluate_arrays(device: HipDevice, request: dict, max_error_px: float):
    """`tcv_obs_evaluate` on a packed request: (flags, summary, milliseconds of the call)."""
    lib = _native.library()
    count, guarded = request["count"], request["guarded"]
    tracks = len(request["lengths"]) if guarded else 0
    control = _native.ObsControl(views=request["views"], points=request["points"],
                                 observations=count, tracks=tracks,
                                 max_error_px=float(max_error_px),
                                 max_relative_depth_std_dev=request["depth"]["maxRelativeDepthStdDev"],
                                 pixel_noise_std_dev_px=request["depth"]["pixelNoiseStdDevPx"])
    flags = np.zeros(max(1, count + tracks), dtype=np.uint32)
    summary = np.zeros(2)
    offsets = request["offsets"]

    def job():
        started = time.perf_counter()
        _native.check(lib.tcv_obs_evaluate(
            device.handle, ctypes.byref(control), request["cameras"].ctypes.data,
            request["points_xyz"].ctypes.data, request["pixels"].ctypes.data,
            request["indices"].ctypes.data, None if offsets is None else offsets.ctypes.data,
            flags.ctypes.data, summary.ctypes.data), "tcv_obs_evaluate")
        return (time.perf_counter() - started) * 1e3

    elapsed = device.enqueue(job)
    return flags, summary, elapsed


def evaluate_observations_gpu(device: HipDevice, model: Mapping, max_error_px: float = 3,
                              options: Mapping | None = None) -> dict:
    """`evaluateObservationsGpu` (observations.ts:34-41): every observation of the scene."""
    options = options or {}
    request = observation_request(model, max_error_px, options, False)
    return evaluate_request(device, request, max_error_px)


def evaluate_continued_observations_gpu(device: HipDevice, model: Mapping,
                                        max_error_px: float = 3,
                                        options: Mapping | None = None) -> dict:
    """`evaluateContinuedObservationsGpu` (observations.ts:49-57): only each landmark's last
    observation, unless the depth guard is on (it needs every observation)."""
    options = options or {}
    guarded = depth_sensitivity_controls(options)["maxRelativeDepthStdDev"] > 0
    request = observation_request(model, max_error_px, options, not guarded)
    return evaluate_request(device, request, max_error_px)
