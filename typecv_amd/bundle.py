"""`bundle_adjust_gpu` — the reference operator `bundleAdjustGpu` on the MI355X engine.

Host half of src/sfm/bundle/gpu/bundle.ts:803-839 (option checks and routing),
bundle-host.ts:31-332 (problem checks, packing, output unpacking), bundle-directions.ts:19-98
(bearing constraints) and src/geometry/validation.ts:6-46 (scene validation). The solve itself is
one call into the HIP engine through the C ABI (`tcv_ba_solve`); there is no CPU path.

Data contracts (src/contracts/scene.ts:22-40, camera.ts:2-20), as Python mappings:
  view      {"id": int, "camera": {"fx","fy","cx","cy"[,"k1","k2"]}, "pose": {"R": 9, "t": 3}}
            (q = R X + t, R row-major)
  landmark  {"trackId": int, "xyz": 3, "observations": [{"viewId", "featureId", "pixel": 2}]}
  input     {"views": [view...], "landmarks": [landmark...]}
Options use the reference names (`iterations`, `huberPx`, `anchorViewId`, `scaleViewId`,
`variableViewIds`, `refineSharedFocal`, `focalGroups`, `stallTolerance`, `directionalConstraints`,
`maxCgIterations`, `cgRelativeTolerance`, ...) plus `precision` ("fp64" | "mixed").

Errors mirror the reference's: `RangeError` (a ValueError) for rejected scenes and options, raised
before any device work, or after it for the device-detected input errors of `checkBundleOutput`.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
import threading
from typing import Any, Mapping

import numpy as np

from . import _native
from .packing import PackedScene, bundle_array_sizes, finite_gpu_array, pack_arrays

BUNDLE_DEFAULTS = {"iterations": 20, "huberPx": 2.0, "cgRelativeTolerance": 1e-6}
BUNDLE_SOLVER_STALL_TOLERANCE = 1e-9
BUNDLE_STATE = {
    "baseline": 2, "cost": 6, "initialCost": 8, "accepted": 10, "rejected": 11, "commit": 12,
    "inputError": 14, "cgSteps": 16, "failedSolves": 17, "initialRmsePx": 24, "finalRmsePx": 25,
    "size": 26, "focalScale": 26,
}
FIXED_OBSERVATION_WORDS = 42
JOINT_OBSERVATION_WORDS = 45
POINT_WORDS, CAMERA_WORDS = 13, 78
U32_MAX = 0xFFFFFFFF


class RangeError(ValueError):
    """The reference's RangeError."""


# ------------------------------------------------------------------ validation


def _finite(*xs) -> bool:
    return all(isinstance(x, (int, float, np.floating, np.integer)) and math.isfinite(x) for x in xs)


def _safe_int(x) -> bool:
    return isinstance(x, (int, np.integer)) and not isinstance(x, bool) and abs(int(x)) <= 2**53 - 1


def validate_camera(k: Mapping) -> None:
    """`validateCamera`, src/geometry/camera.ts:5-12."""
    if not (_finite(k["fx"], k["fy"], k["cx"], k["cy"], k.get("k1", 0) or 0, k.get("k2", 0) or 0)
            and k["fx"] > 0 and k["fy"] > 0):
        raise RangeError("Invalid camera: finite parameters and positive focal lengths required")


def validate_pose(p: Mapping) -> None:
    """`validatePose`, src/geometry/pose.ts:13-27."""
    R = np.asarray(p["R"], dtype=np.float64).reshape(-1)
    t = np.asarray(p["t"], dtype=np.float64).reshape(-1)
    if R.shape[0] != 9 or t.shape[0] != 3 or not np.all(np.isfinite(R)) or not np.all(np.isfinite(t)):
        raise RangeError("Invalid pose dimensions or values")
    Rm = R.reshape(3, 3)
    q = Rm @ Rm.T
    if np.any(np.abs(q - np.eye(3)) > 1e-6) or abs(np.linalg.det(Rm) - 1) > 1e-6:
        raise RangeError("Pose rotation must belong to SO(3)")


def validate_reconstruction(scene: Mapping) -> None:
    """`validateReconstruction`, src/geometry/validation.ts:6-46."""
    views = set()
    for v in scene["views"]:
        if not _safe_int(v["id"]) or v["id"] < 0 or v["id"] in views:
            raise RangeError("Invalid/duplicate registered view id")
        validate_camera(v["camera"])
        validate_pose(v["pose"])
        views.add(v["id"])
    points, owned = set(), set()
    for p in scene["landmarks"]:
        xyz = np.asarray(p["xyz"], dtype=np.float64).reshape(-1)
        if (not _safe_int(p["trackId"]) or p["trackId"] < 0 or p["trackId"] in points
                or xyz.shape[0] != 3 or not np.all(np.isfinite(xyz))):
            raise RangeError("Invalid/duplicate landmark")
        points.add(p["trackId"])
        if len(p["observations"]) < 2:
            raise RangeError("Landmarks require at least two observations")
        seen = set()
        for o in p["observations"]:
            key = (o["viewId"], o["featureId"])
            px = o["pixel"]
            if (o["viewId"] not in views or o["viewId"] in seen or not _safe_int(o["featureId"])
                    or o["featureId"] < 0 or len(px) != 2 or not _finite(px[0], px[1])
                    or key in owned):
                raise RangeError("Invalid or multiply owned landmark observation")
            seen.add(o["viewId"])
            owned.add(key)


def validate_gpu_camera(k: Mapping) -> dict:
    """`validateGpuCamera`, src/geometry/gpu/bounds.ts:5-17."""
    validate_camera(k)
    k1, k2 = k.get("k1", 0) or 0, k.get("k2", 0) or 0
    if (k["fx"] < 1e-6 or k["fy"] < 1e-6 or k["fx"] > 1e6 or k["fy"] > 1e6
            or max(abs(k["cx"]), abs(k["cy"])) > 1e9 or max(abs(k1), abs(k2)) > 1):
        raise RangeError("Camera exceeds the supported GPU geometry bounds")
    return {**k, "k1": k1, "k2": k2}


def append_bundle_directions(scene: Mapping, options: Mapping) -> dict:
    """`appendBundleDirections`, src/sfm/bundle/bundle-directions.ts:19-72."""
    constraints = options.get("directionalConstraints")
    if constraints is None:
        return scene
    if not isinstance(constraints, (list, tuple)):
        raise RangeError("directionalConstraints must be an array")
    ids = {p["trackId"] for p in scene["landmarks"]}
    views = {v["id"] for v in scene["views"]}
    features = {(o["viewId"], o["featureId"]) for p in scene["landmarks"] for o in p["observations"]}
    extra = []
    for p in constraints:
        d = p.get("direction")
        d = None if d is None else np.asarray(d, dtype=np.float64).reshape(-1)
        norm = float(np.hypot.reduce(d)) if d is not None and d.shape[0] == 3 else float("nan")
        if (not _safe_int(p.get("trackId")) or p["trackId"] < 0 or p["trackId"] in ids or d is None
                or d.shape[0] != 3 or not math.isfinite(norm) or norm <= 1e-12
                or not isinstance(p.get("observations"), (list, tuple))
                or len(p["observations"]) < 2):
            raise RangeError("Invalid or multiply owned directional constraint")
        ids.add(p["trackId"])
        seen = set()
        for o in p["observations"]:
            key = (o["viewId"], o["featureId"])
            px = o["pixel"]
            if (o["viewId"] not in views or o["viewId"] in seen or not _safe_int(o["featureId"])
                    or o["featureId"] < 0 or len(px) != 2 or not _finite(px[0], px[1])
                    or key in features):
                raise RangeError("Invalid or multiply owned directional observation")
            seen.add(o["viewId"])
            features.add(key)
        extra.append({"trackId": p["trackId"], "xyz": d / norm, "observations": p["observations"]})
    if not extra:
        return scene
    return {"views": scene["views"], "landmarks": list(scene["landmarks"]) + extra}


def bundle_stall_tolerance(options: Mapping) -> float:
    """`bundleStallTolerance`, src/sfm/bundle/bundle-shared.ts:8-14."""
    tol = options.get("stallTolerance", BUNDLE_SOLVER_STALL_TOLERANCE)
    if not _finite(tol) or tol < 0 or tol >= 1:
        raise RangeError("stallTolerance must be in [0, 1)")
    return float(tol)


def bundle_problem(scene: Mapping, options: Mapping, default_max_cg) -> dict:
    """`bundleProblem`, src/sfm/bundle/gpu/bundle-host.ts:31-113."""
    validate_reconstruction(scene)
    finite_points = len(scene["landmarks"])
    scene = append_bundle_directions(scene, options)
    C, P = len(scene["views"]), len(scene["landmarks"])
    M = sum(len(p["observations"]) for p in scene["landmarks"])
    if not all(0 <= n <= U32_MAX for n in (C, P, M)):
        raise RangeError("GPU BA scene dimensions exceed integer index bounds")
    iterations = options.get("iterations", BUNDLE_DEFAULTS["iterations"])
    huber = options.get("huberPx", BUNDLE_DEFAULTS["huberPx"])
    max_cg = options.get("maxCgIterations", None)
    if max_cg is None:
        max_cg = default_max_cg(C)
    tolerance = options.get("cgRelativeTolerance", BUNDLE_DEFAULTS["cgRelativeTolerance"])
    if (C < 2 or finite_points < 1 or not _safe_int(iterations) or iterations < 0
            or iterations > 100 or not _finite(huber) or huber < 1e-12 or huber > 1e6
            or not _safe_int(max_cg) or max_cg < 1 or max_cg > 256 or not _finite(tolerance)
            or tolerance < 1e-12 or tolerance > 1e-2):
        raise RangeError("Invalid GPU BA scene/options")
    by_id = {v["id"]: i for i, v in enumerate(scene["views"])}
    anchor_id = options.get("anchorViewId", scene["views"][0]["id"])
    scale_id = options.get("scaleViewId", scene["views"][1]["id"])
    if anchor_id not in by_id or scale_id not in by_id or anchor_id == scale_id:
        raise RangeError("Invalid GPU BA gauge cameras")
    variable_ids = options.get("variableViewIds")
    variable = None if variable_ids is None else set(variable_ids)
    if variable is not None and (len(variable) != len(variable_ids)
                                 or any(i == anchor_id or i not in by_id for i in variable)):
        raise RangeError("Invalid GPU BA variable camera window")
    active = [v["id"] != anchor_id and (variable is None or v["id"] in variable)
              for v in scene["views"]]
    fixed = sum(1 for a in active if not a)
    if fixed > 1 and variable is not None and scale_id in variable:
        raise RangeError("Local GPU BA scale camera must be fixed")
    connected = {anchor_id}
    grew = True
    finite = scene["landmarks"][:finite_points]
    while grew:
        grew = False
        for p in finite:
            if any(o["viewId"] in connected for o in p["observations"]):
                for o in p["observations"]:
                    if o["viewId"] not in connected:
                        connected.add(o["viewId"])
                        grew = True
    if any(v["id"] not in connected for v in scene["views"]):
        raise RangeError("GPU BA requires one connected observation graph")
    return {
        "input": scene, "finitePoints": finite_points, "C": C, "P": P, "M": M,
        "iterations": int(iterations), "huber": float(huber), "maxCg": int(max_cg),
        "tolerance": float(tolerance), "byId": by_id, "anchorId": anchor_id,
        "scaleId": scale_id, "active": active, "fixed": fixed,
    }


def pack_bundle_scene(model: Mapping, by_id: Mapping, active) -> PackedScene:
    """`packBundleScene`, src/sfm/bundle/gpu/bundle-host.ts:177-218 (each camera validated)."""
    views, landmarks = model["views"], model["landmarks"]
    C, P = len(views), len(landmarks)
    poses = np.empty((C, 12))
    models = np.empty((C, 6))
    for c, v in enumerate(views):
        k = validate_gpu_camera(v["camera"])
        poses[c, :9] = np.asarray(v["pose"]["R"], dtype=np.float64).reshape(9)
        poses[c, 9:] = np.asarray(v["pose"]["t"], dtype=np.float64).reshape(3)
        models[c] = [k["fx"], k["fy"], k["cx"], k["cy"], k["k1"], k["k2"]]
    points = np.array([np.asarray(p["xyz"], dtype=np.float64).reshape(3) for p in landmarks])
    lengths = np.array([len(p["observations"]) for p in landmarks], dtype=np.int64)
    obs_camera = np.array([by_id[o["viewId"]] for p in landmarks for o in p["observations"]],
                          dtype=np.uint32)
    pixels = np.array([[o["pixel"][0], o["pixel"][1]] for p in landmarks
                       for o in p["observations"]], dtype=np.float64)
    return pack_arrays(poses, models, points.reshape(P, 3), lengths, obs_camera,
                       pixels.reshape(-1, 2), np.asarray(active, dtype=np.uint32))


def check_bundle_output(values: np.ndarray, base: int) -> None:
    """`checkBundleOutput`, src/sfm/bundle/gpu/bundle-host.ts:278-283."""
    if values[base + BUNDLE_STATE["inputError"]] == 1:
        raise RangeError("GPU BA gauge requires a nonzero baseline")
    if values[base + BUNDLE_STATE["inputError"]] == 2:
        raise RangeError("GPU BA observations must initially project in front of their cameras")


# ------------------------------------------------------------------ device


class HipDevice:
    """The subset of the reference's `ShadeDevice` (src/device/gpu/device.ts:240-1002) that the BA
    operator uses: a FIFO of whole jobs, owned device memory, statistics and disposal."""

    backend = "hip"

    def __init__(self, device: int = 0, working_bytes: int = 0):
        lib = _native.library()
        handle = ctypes.c_void_p()
        _native.check(lib.tcv_ba_create(device, ctypes.byref(handle)), "tcv_ba_create")
        self._lib = lib
        self._handle = handle
        self._lock = threading.Lock()
        self.index = device
        self.upload_bytes = 0
        self.readback_bytes = 0
        self.jobs = 0
        self.working_bytes = 0
        if working_bytes:
            self.set_working_bytes(working_bytes)

    @classmethod
    def open(cls, device: int = 0, working_bytes: int = 0) -> "HipDevice":
        """`ShadeDevice.open({workingBytes})` (device.ts:357-360): `working_bytes` bounds the device
        memory of one bundle adjustment (0: no bound); a problem above it streams or is refused."""
        return cls(device, working_bytes)

    def set_working_bytes(self, working_bytes: int) -> None:
        if not _safe_int(working_bytes) or working_bytes < 0:
            raise RangeError("workingBytes must be a non-negative integer")
        _native.check(self._lib.tcv_ba_set_working_bytes(self.handle, int(working_bytes)),
                      "tcv_ba_set_working_bytes")
        self.working_bytes = int(working_bytes)

    @property
    def handle(self) -> ctypes.c_void_p:
        if self._handle is None:
            raise RuntimeError("The device is closed")
        return self._handle

    @property
    def statistics(self) -> dict:
        live = int(self._lib.tcv_ba_device_bytes(self._handle)) if self._handle else 0
        return {"liveBytes": live, "applicationUploadBytes": self.upload_bytes,
                "applicationReadbackBytes": self.readback_bytes, "jobs": self.jobs}

    def enqueue(self, job):
        """Run `job` after every earlier job of this device (device.ts:592-598)."""
        with self._lock:
            if self._handle is None:
                raise RuntimeError("The device is closed")
            self.jobs += 1
            return job()

    def drop(self) -> None:
        """Free every device array of the context (the reference's `drop`)."""
        if self._handle:
            _native.check(self._lib.tcv_ba_release(self._handle), "tcv_ba_release")

    def dispose(self) -> None:
        if self._handle is not None:
            self._lib.tcv_ba_destroy(self._handle)
            self._handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.dispose()


# ------------------------------------------------------------------ engine calls


def make_control(C: int, P: int, M: int, *, directional: int = 0, anchor: int = 0,
                 scale_view: int = 1, normalize_scale: bool = True, shared_focal: bool = False,
                 iterations: int = 20, max_cg: int = 24, precision: str = "fp64",
                 huber: float = 2.0, tolerance: float = 1e-6, stall_tolerance: float = 1e-9,
                 refinement_tolerance: float = 0.0, timing: bool = False,
                 implicit: bool = False, schur_cache: bool = False,
                 direct: bool = False) -> _native.Control:
    """The `tcv_ba_control` of a problem. `direct`: the direct route (dense reduced camera system,
    Cholesky, matrix-free certificate) for bundles of at most 2,048 cameras; `schur_cache`: the
    reference's `schurCache`, which takes the direct route only inside its guard (shared focal,
    at most 16 cameras, M (6C + 1) <= 262,144) and is ignored elsewhere."""
    flags = ((_native.FLAG_TIMING if timing else 0) | (_native.FLAG_IMPLICIT if implicit else 0)
             | (_native.FLAG_SCHUR_CACHE if schur_cache else 0)
             | (_native.FLAG_DIRECT if direct else 0))
    return _native.Control(
        cameras=C, points=P, directional_points=directional, observations=M, anchor=anchor,
        scale_view=scale_view, normalize_scale=int(normalize_scale),
        shared_focal=int(shared_focal), iterations=iterations, max_cg=max_cg,
        precision=_native.MIXED if precision == "mixed" else _native.FP64,
        flags=flags, huber=huber, relative_tolerance=tolerance,
        stall_tolerance=stall_tolerance, refinement_tolerance=refinement_tolerance)


def solve_packed(device: HipDevice, control: _native.Control, packed: PackedScene):
    """One engine solve of a packed problem from host arrays: (out array, report dict)."""
    lib = _native.library()
    scene = np.ascontiguousarray(packed.scene, dtype=np.float64)
    index = np.ascontiguousarray(packed.index, dtype=np.uint32)
    if scene.shape[0] != lib.tcv_ba_scene_size(ctypes.byref(control)):
        raise RangeError("scene array does not match the control sizes")
    if index.shape[0] != lib.tcv_ba_index_size(ctypes.byref(control)):
        raise RangeError("index array does not match the control sizes")
    out = np.empty(lib.tcv_ba_output_size(ctypes.byref(control)), dtype=np.float64)
    report = _native.Report()

    def job():
        status = lib.tcv_ba_solve(device.handle, ctypes.byref(control), scene.ctypes.data,
                                  index.ctypes.data, out.ctypes.data, ctypes.byref(report))
        _native.check(status, "tcv_ba_solve")
        device.upload_bytes += scene.nbytes + index.nbytes
        device.readback_bytes += out.nbytes
        return out

    device.enqueue(job)
    return out, report.as_dict()


def solve_batch(device: HipDevice, controls, packeds):
    """`tcv_ba_solve_batch`: many independent problems in one batched solve.
    Returns (list of out arrays, report dict, per-problem LM-loop time in ms)."""
    lib = _native.library()
    n = len(controls)
    if n != len(packeds) or n < 1:
        raise ValueError("one control per packed problem")
    arr = (_native.Control * n)(*controls)
    scenes = [np.ascontiguousarray(p.scene, dtype=np.float64) for p in packeds]
    indices = [np.ascontiguousarray(p.index, dtype=np.uint32) for p in packeds]
    outs = [np.empty(lib.tcv_ba_output_size(ctypes.byref(c)), dtype=np.float64) for c in controls]
    ptr = ctypes.c_void_p
    scene_ptrs = (ptr * n)(*[s.ctypes.data for s in scenes])
    index_ptrs = (ptr * n)(*[x.ctypes.data for x in indices])
    out_ptrs = (ptr * n)(*[o.ctypes.data for o in outs])
    finish = np.zeros(n)
    report = _native.Report()

    def job():
        _native.check(lib.tcv_ba_solve_batch(device.handle, n, arr, scene_ptrs, index_ptrs,
                                             out_ptrs, ctypes.byref(report), finish.ctypes.data),
                      "tcv_ba_solve_batch")

    device.enqueue(job)
    return outs, report.as_dict(), finish


def schur_probe(device: HipDevice, control: _native.Control, packed: PackedScene, lam: float,
                log_focal: float, p: np.ndarray, repeat: int = 1):
    """`tcv_ba_schur_probe`: (rhs, S p, M^-1 p, status, report)."""
    lib = _native.library()
    m = control.cameras * 6 + (1 if control.shared_focal else 0)
    p = np.ascontiguousarray(p, dtype=np.float64)
    if p.shape[0] != m:
        raise ValueError("vector length is not 6C (+1)")
    rhs, sp, z = np.empty(m), np.empty(m), np.empty(m)
    status = np.zeros(1, dtype=np.int32)
    report = _native.Report()
    scene = np.ascontiguousarray(packed.scene, dtype=np.float64)
    index = np.ascontiguousarray(packed.index, dtype=np.uint32)

    def job():
        _native.check(lib.tcv_ba_schur_probe(
            device.handle, ctypes.byref(control), scene.ctypes.data, index.ctypes.data,
            float(lam), float(log_focal), p.ctypes.data, rhs.ctypes.data, sp.ctypes.data,
            z.ctypes.data, status.ctypes.data, int(repeat), ctypes.byref(report)),
            "tcv_ba_schur_probe")

    device.enqueue(job)
    return rhs, sp, z, int(status[0]), report.as_dict()


# ------------------------------------------------------------------ derivatives


def at_solution(packed: PackedScene, out: np.ndarray) -> PackedScene:
    """The packed problem with its poses and points replaced by a solve's output coordinates."""
    scene = packed.scene.copy()
    n = packed.C * 12 + packed.P * 3
    scene[:n] = out[:n]
    return dataclasses.replace(packed, scene=scene)


def output_log_focal(control: _native.Control, out: np.ndarray) -> float:
    """The log focal scale word of a solve's output (0 on the fixed-focal route)."""
    if not control.shared_focal:
        return 0.0
    return float(out[control.cameras * 12 + control.points * 3 + 26])


def jacobian(device: HipDevice, control: _native.Control, packed: PackedScene,
             log_focal: float = 0.0) -> np.ndarray:
    """`tcv_ba_jacobian`: (M, 23) = r (2), d(u,v)/d(omega, v) (12), d(u,v)/dX (6),
    d(u,v)/d log f (2), w per observation."""
    lib = _native.library()
    scene = np.ascontiguousarray(packed.scene, dtype=np.float64)
    index = np.ascontiguousarray(packed.index, dtype=np.uint32)
    out = np.empty((packed.M, 23))
    report = _native.Report()

    def job():
        _native.check(lib.tcv_ba_jacobian(device.handle, ctypes.byref(control), scene.ctypes.data,
                                          index.ctypes.data, float(log_focal), out.ctypes.data,
                                          ctypes.byref(report)), "tcv_ba_jacobian")

    device.enqueue(job)
    return out


def polish(device: HipDevice, control: _native.Control, packed: PackedScene,
           log_focal: float = 0.0, iterations: int = 3, max_cg: int = 4000,
           tolerance: float = 1e-12, lm_iterations: int = 0):
    """`tcv_ba_polish`: up to `lm_iterations` damped steps, then up to `iterations` Newton steps,
    on the exact Hessian from the packed scene. Returns (out in the solve layout, report)."""
    lib = _native.library()
    scene = np.ascontiguousarray(packed.scene, dtype=np.float64)
    index = np.ascontiguousarray(packed.index, dtype=np.uint32)
    out = np.empty(lib.tcv_ba_output_size(ctypes.byref(control)))
    report = _native.Report()

    def job():
        _native.check(lib.tcv_ba_polish(device.handle, ctypes.byref(control), scene.ctypes.data,
                                        index.ctypes.data, float(log_focal), int(lm_iterations),
                                        int(iterations), int(max_cg), float(tolerance),
                                        out.ctypes.data, ctypes.byref(report)), "tcv_ba_polish")

    device.enqueue(job)
    return out, report.as_dict()


ADJOINT_MODES = {"exact": 0, "gaussNewton": 1}


def backward(device: HipDevice, control: _native.Control, packed: PackedScene, log_focal: float,
             cotangent: np.ndarray, max_cg: int = 4000, tolerance: float = 1e-10,
             mode: str = "exact"):
    """`tcv_ba_backward`: the pixel cotangent (M, 2) of a cotangent on the outputs
    (12C + 3P [+ 1]), the input-focal cotangent (C: d/d log of each camera's fx, fy), the adjoint
    lambda (6C + 3P [+ 1], chart) and the report. `mode`: "exact" (the exact Hessian; needs a
    stationary scene) or "gaussNewton" (the Gauss-Newton Hessian, defined at any scene)."""
    if mode not in ADJOINT_MODES:
        raise ValueError(f"adjoint mode must be one of {sorted(ADJOINT_MODES)}")
    lib = _native.library()
    n_out = control.cameras * 12 + control.points * 3 + (1 if control.shared_focal else 0)
    cot = np.ascontiguousarray(cotangent, dtype=np.float64)
    if cot.shape != (n_out,):
        raise ValueError(f"cotangent must have {n_out} entries")
    scene = np.ascontiguousarray(packed.scene, dtype=np.float64)
    index = np.ascontiguousarray(packed.index, dtype=np.uint32)
    pixels = np.empty((packed.M, 2))
    focal = np.empty(control.cameras)
    adjoint = np.empty(control.cameras * 6 + control.points * 3 + (1 if control.shared_focal else 0))
    report = _native.Report()

    def job():
        _native.check(lib.tcv_ba_backward(device.handle, ctypes.byref(control), scene.ctypes.data,
                                          index.ctypes.data, float(log_focal), cot.ctypes.data,
                                          int(max_cg), float(tolerance), ADJOINT_MODES[mode],
                                          pixels.ctypes.data, focal.ctypes.data,
                                          adjoint.ctypes.data, ctypes.byref(report)),
                      "tcv_ba_backward")

    device.enqueue(job)
    return pixels, focal, adjoint, report.as_dict()


# ------------------------------------------------------------------ the operator


def _check_option_enums(options: Mapping) -> None:
    """bundle.ts:808-831."""
    size = options.get("cameraProductWorkgroupSize")
    if size is not None and size not in (32, 64, 128, 256, "auto"):
        raise RangeError("cameraProductWorkgroupSize must be 32, 64, 128, 256 or auto")
    tile = options.get("cameraProductTileSize")
    if tile is not None and (not _safe_int(tile) or tile < 1 or tile > U32_MAX):
        raise RangeError("cameraProductTileSize must be a positive u32 integer")
    product = options.get("cameraProduct")
    if product is not None and product not in ("workgroups", "rows"):
        raise RangeError("cameraProduct must be 'workgroups' or 'rows'")
    cache = options.get("schurCache")
    if cache is not None and not isinstance(cache, bool):
        raise RangeError("schurCache must be a boolean")
    direct = options.get("directSolve")
    if direct is not None and not isinstance(direct, bool):
        raise RangeError("directSolve must be a boolean")
    reduction = options.get("schurReduction")
    if reduction is not None and reduction not in ("auto", "serial", "cooperative"):
        raise RangeError("Invalid Schur reduction mode")
    precision = options.get("precision", "fp64")
    if precision not in ("fp64", "mixed"):
        raise RangeError("precision must be 'fp64' or 'mixed'")


def bundle_adjust_gpu(device: HipDevice, scene: Mapping, options: Mapping | None = None) -> dict:
    """`bundleAdjustGpu(device, input, options)`: joint pose/point (/shared focal) LM bundle
    adjustment with matrix-free Schur PCG. Returns the reference's result record:
    {reconstruction, diagnostics, cgIterations, linearSolverFailures, execution
    [, directionalConstraints]}. The input is never mutated."""
    options = dict(options or {})
    _check_option_enums(options)
    shared = bool(options.get("refineSharedFocal"))
    if shared:
        groups = options.get("focalGroups")
        if groups and len({groups.get(v["id"]) for v in scene["views"]}) > 1:
            raise RangeError("Per-group focal refinement is not implemented on the GPU")
        default_max_cg = lambda C: min(256, max(32, C * 12 + 1))  # noqa: E731  joint-focal.ts:83
        words = JOINT_OBSERVATION_WORDS
    else:
        default_max_cg = lambda C: min(96, max(24, C * 12))  # noqa: E731  bundle.ts:94
        words = FIXED_OBSERVATION_WORDS
    stall = bundle_stall_tolerance(options)
    problem = bundle_problem(scene, options, default_max_cg)
    model = problem["input"]
    C, P, M = problem["C"], problem["P"], problem["M"]
    sizes = bundle_array_sizes(C, P, M)
    lengths = [sizes["sceneSize"], sizes["trialSize"], words * M + POINT_WORDS * P + CAMERA_WORDS * C,
               30 * C + 3 * P, sizes["indexSize"]]
    if shared:
        lengths.append(6 * C + 6 + 2 * M)
    if not all(0 <= n <= U32_MAX for n in lengths):
        raise RangeError("GPU BA array lengths exceed integer index bounds")
    packed = pack_bundle_scene(model, problem["byId"], problem["active"])
    if not finite_gpu_array(packed.scene):
        raise RangeError("GPU BA input exceeds supported coordinate bounds")
    control = make_control(
        C, P, M, directional=P - problem["finitePoints"], anchor=problem["byId"][problem["anchorId"]],
        scale_view=problem["byId"][problem["scaleId"]], normalize_scale=problem["fixed"] == 1,
        shared_focal=shared, iterations=problem["iterations"], max_cg=problem["maxCg"],
        precision=options.get("precision", "fp64"), huber=problem["huber"],
        tolerance=problem["tolerance"], stall_tolerance=stall,
        schur_cache=options.get("schurCache") is True,
        direct=options.get("directSolve") is True)
    before = device.statistics
    values, report = solve_packed(device, control, packed)
    after = device.statistics
    base = problem["C"] * 12 + problem["P"] * 3
    check_bundle_output(values, base)
    changed = values[base + BUNDLE_STATE["accepted"]] > 0
    focal = math.exp(values[base + BUNDLE_STATE["focalScale"]]) if shared and changed else 1.0
    views = []
    for c, v in enumerate(model["views"]):
        camera = dict(v["camera"])
        if shared and changed:
            camera["fx"] = v["camera"]["fx"] * focal
            camera["fy"] = v["camera"]["fy"] * focal
        if changed and problem["active"][c]:
            pose = {"R": values[c * 12: c * 12 + 9].copy(), "t": values[c * 12 + 9: c * 12 + 12].copy()}
        else:
            pose = {"R": np.array(v["pose"]["R"], dtype=np.float64),
                    "t": np.array(v["pose"]["t"], dtype=np.float64)}
        views.append({**v, "camera": camera, "pose": pose})
    landmarks = []
    for i, p in enumerate(model["landmarks"]):
        xyz = (values[C * 12 + i * 3: C * 12 + i * 3 + 3].copy() if changed
               else np.array(p["xyz"], dtype=np.float64))
        landmarks.append({**p, "xyz": xyz, "observations": [dict(o) for o in p["observations"]]})
    diagnostics = {
        "initialCost": float(values[base + BUNDLE_STATE["initialCost"]]),
        "finalCost": float(values[base + BUNDLE_STATE["cost"]]),
        "initialRmsePx": float(values[base + BUNDLE_STATE["initialRmsePx"]]),
        "finalRmsePx": float(values[base + BUNDLE_STATE["finalRmsePx"]]),
        "acceptedSteps": float(values[base + BUNDLE_STATE["accepted"]]),
        "rejectedSteps": float(values[base + BUNDLE_STATE["rejected"]]),
        "observations": M,
        "baseline": float(values[base + BUNDLE_STATE["baseline"]]),
    }
    result = {
        "reconstruction": {"views": views, "landmarks": landmarks},
        "diagnostics": diagnostics,
        "cgIterations": float(values[base + BUNDLE_STATE["cgSteps"]]),
        "linearSolverFailures": float(values[base + BUNDLE_STATE["failedSolves"]]),
        "execution": {
            "backend": "hip",
            "report": {"tier": "hip", **report},
            "applicationUploadBytes": after["applicationUploadBytes"] - before["applicationUploadBytes"],
            "applicationReadbackBytes": after["applicationReadbackBytes"] - before["applicationReadbackBytes"],
            "cpuStages": [],
        },
    }
    finite_points = problem["finitePoints"]
    if len(landmarks) != finite_points:  # separateBundleDirections, bundle-directions.ts:75-98
        directions = landmarks[finite_points:]
        result["reconstruction"]["landmarks"] = landmarks[:finite_points]
        result["directionalConstraints"] = [
            {"trackId": p["trackId"], "direction": p["xyz"], "observations": p["observations"]}
            for p in directions]
        result["diagnostics"]["directionalObservations"] = sum(len(p["observations"]) for p in directions)
        result["diagnostics"]["pointParameters"] = 3 * finite_points + 2 * len(directions)
    return result


# The reference's camelCase name.
bundleAdjustGpu = bundle_adjust_gpu


def bundle_adjust_gpu_backward(device: HipDevice, scene: Mapping, options: Mapping | None,
                               result: Mapping, cotangent: Mapping, *, polish_steps: int = 3,
                               tolerance: float = 1e-10, max_cg: int = 4000,
                               mode: str = "exact") -> dict:
    """The backward entry of `bundle_adjust_gpu`: for a cotangent on its outputs,
    {"views": [{"R": 3x3, "t": 3}] (result order), "landmarks": [{"xyz": 3}],
     "directionalConstraints": [{"direction": 3}] (when the result has them),
     "logFocal": d/d log f (shared-focal route)},
    returns the cotangents of the observation pixels and of each view's input focal by implicit
    differentiation at the solution:
    {"landmarks": [[2-vectors per observation]] (input landmark and observation order),
     "directionalConstraints": [...] likewise, "pixels": (M, 2) in packed order,
     "focal": [d/d log of view c's input fx and fy, scaled together] (result view order), "report"}.
    `polish_steps` Newton steps on the exact Hessian first make the returned solution stationary to
    rounding (0 differentiates at the result as given). `mode` "gaussNewton" solves with the
    Gauss-Newton Hessian instead of the exact one (defined at any scene; the usual linearised
    sensitivity, off the exact derivative by the residual-weighted curvature it drops). Entries the
    cotangent omits are zero."""
    options = dict(options or {})
    _check_option_enums(options)
    shared = bool(options.get("refineSharedFocal"))
    if shared:
        default_max_cg = lambda C: min(256, max(32, C * 12 + 1))  # noqa: E731
    else:
        default_max_cg = lambda C: min(96, max(24, C * 12))  # noqa: E731
    problem = bundle_problem(scene, options, default_max_cg)
    model = problem["input"]
    C, P, M = problem["C"], problem["P"], problem["M"]
    finite = problem["finitePoints"]
    packed = pack_bundle_scene(model, problem["byId"], problem["active"])
    control = make_control(
        C, P, M, directional=P - finite, anchor=problem["byId"][problem["anchorId"]],
        scale_view=problem["byId"][problem["scaleId"]], normalize_scale=problem["fixed"] == 1,
        shared_focal=shared, iterations=problem["iterations"], max_cg=problem["maxCg"],
        huber=problem["huber"], tolerance=problem["tolerance"])
    rec = result["reconstruction"]
    coords = packed.scene.copy()
    for c, v in enumerate(rec["views"]):
        coords[12 * c: 12 * c + 9] = np.asarray(v["pose"]["R"], dtype=np.float64).reshape(9)
        coords[12 * c + 9: 12 * c + 12] = np.asarray(v["pose"]["t"], dtype=np.float64).reshape(3)
    directions = result.get("directionalConstraints", [])
    for i, p in enumerate(list(rec["landmarks"]) + [{"xyz": d["direction"]} for d in directions]):
        coords[12 * C + 3 * i: 12 * C + 3 * i + 3] = np.asarray(p["xyz"], dtype=np.float64)
    at = dataclasses.replace(packed, scene=coords)
    log_focal = 0.0
    if shared:
        log_focal = math.log(rec["views"][0]["camera"]["fx"] / model["views"][0]["camera"]["fx"])
    report = {}
    if polish_steps:
        out, report = polish(device, control, at, log_focal, iterations=polish_steps,
                             max_cg=max_cg, tolerance=min(1e-12, tolerance))
        at = at_solution(at, out)
        log_focal = output_log_focal(control, out)
    cot = np.zeros(12 * C + 3 * P + (1 if shared else 0))
    for c, v in enumerate(cotangent.get("views", [])):
        if "R" in v:
            cot[12 * c: 12 * c + 9] = np.asarray(v["R"], dtype=np.float64).reshape(9)
        if "t" in v:
            cot[12 * c + 9: 12 * c + 12] = np.asarray(v["t"], dtype=np.float64).reshape(3)
    for i, p in enumerate(cotangent.get("landmarks", [])):
        cot[12 * C + 3 * i: 12 * C + 3 * i + 3] = np.asarray(p["xyz"], dtype=np.float64)
    for k, d in enumerate(cotangent.get("directionalConstraints", [])):
        i = finite + k
        cot[12 * C + 3 * i: 12 * C + 3 * i + 3] = np.asarray(d["direction"], dtype=np.float64)
    if shared:
        cot[-1] = float(cotangent.get("logFocal", 0.0))
    pixels, focal, adjoint, back = backward(device, control, at, log_focal, cot, max_cg=max_cg,
                                            tolerance=tolerance, mode=mode)
    lengths = [len(p["observations"]) for p in model["landmarks"]]
    starts = np.concatenate([[0], np.cumsum(lengths)])
    per_point = [[pixels[k].copy() for k in range(starts[i], starts[i + 1])] for i in range(P)]
    return {"landmarks": per_point[:finite], "directionalConstraints": per_point[finite:],
            "pixels": pixels, "focal": focal, "adjoint": adjoint, "logFocal": log_focal,
            "report": {"polish": report, "backward": back}}


bundleAdjustGpuBackward = bundle_adjust_gpu_backward
