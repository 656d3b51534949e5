"""TEST INFRASTRUCTURE: ctypes binding of the CPU oracle `oracle/libba_oracle.so`.

Only tests, the smoke check and the CPU-baseline leg of the benchmark import this module. It is
the checker of the HIP engine, never a path of the product.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

from typecv_amd import _native
from typecv_amd.packing import PackedScene

LIBRARY = Path(__file__).resolve().parent / "libba_oracle.so"
_lib = None


def library() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIBRARY.exists():
        from . import build
        build.build()
    lib = ctypes.CDLL(str(LIBRARY))
    P = ctypes.POINTER
    vp = ctypes.c_void_p
    lib.tcv_oracle_set_threads.argtypes = [ctypes.c_int]
    lib.tcv_oracle_set_threads.restype = ctypes.c_int
    lib.tcv_oracle_solve.argtypes = [P(_native.Control), vp, vp, vp, vp, ctypes.c_uint32]
    lib.tcv_oracle_solve.restype = ctypes.c_int
    lib.tcv_oracle_schur_probe.argtypes = [P(_native.Control), vp, vp, ctypes.c_double,
                                           ctypes.c_double, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                           vp, vp]
    lib.tcv_oracle_schur_probe.restype = ctypes.c_int
    lib.tcv_oracle_pcg.argtypes = [P(_native.Control), vp, vp, ctypes.c_double, ctypes.c_double,
                                   vp, vp]
    lib.tcv_oracle_pcg.restype = ctypes.c_int
    lib.tcv_oracle_project.argtypes = [vp, vp, vp]
    lib.tcv_oracle_project.restype = ctypes.c_int
    lib.tcv_oracle_radial_limit.argtypes = [vp]
    lib.tcv_oracle_radial_limit.restype = ctypes.c_double
    lib.tcv_oracle_pose_row.argtypes = [vp, vp, vp]
    lib.tcv_oracle_pose_row.restype = None
    lib.tcv_oracle_robust_cost.argtypes = [ctypes.c_double, ctypes.c_double]
    lib.tcv_oracle_robust_cost.restype = ctypes.c_double
    lib.tcv_oracle_solve3.argtypes = [vp, vp, vp]
    lib.tcv_oracle_solve3.restype = ctypes.c_int
    lib.tcv_oracle_solve6.argtypes = [vp, vp, vp]
    lib.tcv_oracle_solve6.restype = ctypes.c_int
    lib.tcv_oracle_move_pose.argtypes = [vp, vp, vp]
    lib.tcv_oracle_move_pose.restype = None
    lib.tcv_oracle_focal_scale.argtypes = [ctypes.c_double]
    lib.tcv_oracle_focal_scale.restype = ctypes.c_double
    lib.tcv_oracle_gradient.argtypes = [P(_native.Control), vp, vp, ctypes.c_double, vp, vp]
    lib.tcv_oracle_dense_solve.argtypes = [P(_native.Control), vp, vp, vp]
    lib.tcv_oracle_observations.argtypes = [vp, vp, vp, vp, ctypes.c_uint32, vp, ctypes.c_uint32,
                                            ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                            vp, vp, vp]
    lib.tcv_oracle_observations.restype = None
    _lib = lib
    return lib


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data


def set_threads(n: int) -> int:
    return library().tcv_oracle_set_threads(n)


def solve(control: _native.Control, packed: PackedScene, max_lm: int = 0):
    """The oracle's whole operator: (out array in packBundleResult layout, stats dict)."""
    lib = library()
    size = control.cameras * 12 + control.points * 3 + 26 + (1 if control.shared_focal else 0)
    out = np.empty(size)
    stats = np.zeros(4)
    scene = np.ascontiguousarray(packed.scene, dtype=np.float64)
    index = np.ascontiguousarray(packed.index, dtype=np.uint32)
    status = lib.tcv_oracle_solve(ctypes.byref(control), _ptr(scene), _ptr(index), _ptr(out),
                                  _ptr(stats), int(max_lm))
    if status != 0:
        raise RuntimeError(f"oracle solve failed ({status})")
    return out, {"lm_iterations": stats[0], "cg_iterations": stats[1], "wall_ms": stats[2],
                 "lm_ms": stats[3]}


def schur_probe(control: _native.Control, packed: PackedScene, lam: float, log_focal: float,
                p: np.ndarray, blocks: bool = False):
    """One assembled linear system: (rhs, S p, M^-1 p, status[, blocks dict])."""
    lib = library()
    C, P, M = control.cameras, control.points, control.observations
    shared = 1 if control.shared_focal else 0
    m, K = 6 * C + shared, 21 if shared else 18
    p = np.ascontiguousarray(p, dtype=np.float64)
    rhs, sp, z = np.empty(m), np.empty(m), np.empty(m)
    status = np.zeros(1, dtype=np.int32)
    out = None
    if blocks:
        out = {"J": np.empty((M, 24)), "cross": np.empty((M, K)), "points": np.empty((P, 13)),
               "cameras": np.empty((C, 36)), "inverse": np.empty((C, 36)),
               "focal": np.empty(6 * C + 6 + 2 * M)}
    scene = np.ascontiguousarray(packed.scene, dtype=np.float64)
    index = np.ascontiguousarray(packed.index, dtype=np.uint32)
    g = (lambda k: _ptr(out[k])) if blocks else (lambda k: None)
    r = lib.tcv_oracle_schur_probe(ctypes.byref(control), _ptr(scene), _ptr(index), float(lam),
                                   float(log_focal), _ptr(p), _ptr(rhs), _ptr(sp), _ptr(z),
                                   _ptr(status), g("J"), g("cross"), g("points"), g("cameras"),
                                   g("inverse"), g("focal"))
    if r != 0:
        raise RuntimeError(f"oracle probe failed ({r})")
    if blocks:
        return rhs, sp, z, int(status[0]), out
    return rhs, sp, z, int(status[0])


def pcg(control: _native.Control, packed: PackedScene, lam: float, log_focal: float = 0.0):
    """PCG solution x of one assembled system and {steps, converged, failed, ratio}."""
    m = control.cameras * 6 + (1 if control.shared_focal else 0)
    x, stats = np.empty(m), np.zeros(4)
    scene = np.ascontiguousarray(packed.scene, dtype=np.float64)
    index = np.ascontiguousarray(packed.index, dtype=np.uint32)
    r = library().tcv_oracle_pcg(ctypes.byref(control), _ptr(scene), _ptr(index), float(lam),
                                 float(log_focal), _ptr(x), _ptr(stats))
    if r != 0:
        raise RuntimeError(f"oracle pcg failed ({r})")
    return x, {"steps": stats[0], "converged": stats[1], "failed": stats[2], "ratio": stats[3]}


def project(model, q):
    """(ok, pixel[2], du[3], dv[3]) of numerics.shade.ts `project`."""
    lib = library()
    model = np.ascontiguousarray(model, dtype=np.float64)
    q = np.ascontiguousarray(q, dtype=np.float64)
    out = np.empty(8)
    ok = lib.tcv_oracle_project(_ptr(model), _ptr(q), _ptr(out))
    return bool(ok), out[:2].copy(), out[2:5].copy(), out[5:8].copy()


def radial_limit(model) -> float:
    model = np.ascontiguousarray(model, dtype=np.float64)
    return library().tcv_oracle_radial_limit(_ptr(model))


def pose_row(q, j):
    q = np.ascontiguousarray(q, dtype=np.float64)
    j = np.ascontiguousarray(j, dtype=np.float64)
    out = np.empty(6)
    library().tcv_oracle_pose_row(_ptr(q), _ptr(j), _ptr(out))
    return out


def robust_cost(error: float, huber: float) -> float:
    return library().tcv_oracle_robust_cost(float(error), float(huber))


def solve3(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(9)
    b = np.ascontiguousarray(b, dtype=np.float64).reshape(3)
    x = np.empty(3)
    ok = library().tcv_oracle_solve3(_ptr(a), _ptr(b), _ptr(x))
    return bool(ok), x


def solve6(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(36)
    b = np.ascontiguousarray(b, dtype=np.float64).reshape(6)
    x = np.empty(6)
    ok = library().tcv_oracle_solve6(_ptr(a), _ptr(b), _ptr(x))
    return bool(ok), x


def move_pose(p, d):
    p = np.ascontiguousarray(p, dtype=np.float64).reshape(12)
    d = np.ascontiguousarray(d, dtype=np.float64).reshape(6)
    out = np.empty(12)
    library().tcv_oracle_move_pose(_ptr(p), _ptr(d), _ptr(out))
    return out


def focal_scale(x: float) -> float:
    return library().tcv_oracle_focal_scale(float(x))


def gradient(control: _native.Control, scene: np.ndarray, index: np.ndarray,
             log_focal: float = 0.0):
    """(g, half cost, skipped): the chart gradient of 1/2 sum robustCost (6C + 3P + shared)."""
    n = control.cameras * 6 + control.points * 3 + (1 if control.shared_focal else 0)
    g = np.empty(n)
    cost = np.zeros(1)
    scene = np.ascontiguousarray(scene, dtype=np.float64)
    index = np.ascontiguousarray(index, dtype=np.uint32)
    r = library().tcv_oracle_gradient(ctypes.byref(control), _ptr(scene), _ptr(index),
                                      float(log_focal), _ptr(g), _ptr(cost))
    if r < 0:
        raise RuntimeError(f"oracle gradient failed ({r})")
    return g, float(cost[0]), bool(r)


def dense_solve(control: _native.Control, packed: PackedScene) -> np.ndarray:
    """The reference's dense CPU `bundleAdjust` (packBundleResult layout, see the C header)."""
    size = control.cameras * 12 + control.points * 3 + 26 + (1 if control.shared_focal else 0)
    out = np.empty(size)
    scene = np.ascontiguousarray(packed.scene, dtype=np.float64)
    index = np.ascontiguousarray(packed.index, dtype=np.uint32)
    r = library().tcv_oracle_dense_solve(ctypes.byref(control), _ptr(scene), _ptr(index), _ptr(out))
    if r != 0:
        raise RuntimeError(f"oracle dense solve failed ({r})")
    return out


def observations(cameras: np.ndarray, points: np.ndarray, pixels: np.ndarray,
                 indices: np.ndarray, offsets: np.ndarray | None, max_error: float,
                 max_relative_depth_std: float = 0.0, pixel_noise_std: float = 1.0):
    """The fixed-scene observation evaluation (`evaluateObservationsGpu`'s kernels, see the C
    source): (flags: count [+ tracks] u32, errors: count f64, summary: [error sum, inliers])."""
    count = len(indices) // 2
    tracks = 0 if offsets is None else len(offsets) - 1
    cameras = np.ascontiguousarray(cameras, dtype=np.float64)
    points = np.ascontiguousarray(points, dtype=np.float64)
    pixels = np.ascontiguousarray(pixels, dtype=np.float64)
    indices = np.ascontiguousarray(indices, dtype=np.uint32)
    off = None if offsets is None else np.ascontiguousarray(offsets, dtype=np.uint32)
    flags = np.zeros(max(1, count + tracks), dtype=np.uint32)
    errors = np.zeros(max(1, count))
    summary = np.zeros(2)
    library().tcv_oracle_observations(_ptr(cameras), _ptr(points), _ptr(pixels), _ptr(indices),
                                      count, _ptr(off), tracks, float(max_error),
                                      float(max_relative_depth_std), float(pixel_noise_std),
                                      _ptr(flags), _ptr(errors), _ptr(summary))
    return flags[:count + tracks], errors[:count], summary
