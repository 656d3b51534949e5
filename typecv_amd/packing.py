"""Packed scene and index arrays of the bundle adjustment (the engine's input contract).

Restates the layouts of the reference host half, src/sfm/bundle/gpu/bundle-host.ts:
`bundleIndexLayout` (:121-131), `bundleArraySizes` (:134-143) and `packBundleScene` (:177-218),
with strides POSE=12, MODEL=6, POINT=3, PIXEL=2 (src/sfm/scene/scene.ts:8-14).

Observations are numbered point-major in input order; each camera's list holds its observations in
ascending observation order. The arrays are built with numpy so that scenes of millions of
observations pack in well under a second.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

POSE_STRIDE = 12
MODEL_STRIDE = 6
POINT_STRIDE = 3
PIXEL_STRIDE = 2
STATE_WORDS = 26


def bundle_index_layout(C: int, P: int, M: int) -> dict:
    """Word offsets of the index sections (bundle-host.ts:121-131)."""
    return {
        "observationCamera": 0,
        "observationPoint": M,
        "pointStart": 2 * M,
        "cameraStart": 2 * M + P + 1,
        "cameraObservations": 2 * M + P + C + 2,
        "cameraActive": 3 * M + P + C + 2,
        "size": 3 * M + P + 2 * C + 2,
    }


def bundle_array_sizes(C: int, P: int, M: int) -> dict:
    """Element counts of the arrays both routes size alike (bundle-host.ts:134-143)."""
    return {
        "sceneSize": (POSE_STRIDE + MODEL_STRIDE) * C + POINT_STRIDE * P + PIXEL_STRIDE * M,
        "trialSize": POSE_STRIDE * C + POINT_STRIDE * P,
        "indexSize": bundle_index_layout(C, P, M)["size"],
        "cameraScratchSize": 36 * C,
        "stateSize": STATE_WORDS,
        "partialsSize": 3 * max(1, -(-M // 64)),
    }


@dataclass
class PackedScene:
    """A problem in the engine's array form: `scene` f64 and `index` u32 of packBundleScene."""

    C: int
    P: int
    M: int
    scene: np.ndarray
    index: np.ndarray

    def poses(self) -> np.ndarray:
        return self.scene[: self.C * 12].reshape(self.C, 12)

    def points(self) -> np.ndarray:
        return self.scene[self.C * 12: self.C * 12 + self.P * 3].reshape(self.P, 3)

    def models(self) -> np.ndarray:
        at = self.C * 12 + self.P * 3
        return self.scene[at: at + self.C * 6].reshape(self.C, 6)

    def pixels(self) -> np.ndarray:
        return self.scene[self.C * 18 + self.P * 3:].reshape(self.M, 2)

    def section(self, name: str) -> np.ndarray:
        lay = bundle_index_layout(self.C, self.P, self.M)
        lengths = {
            "observationCamera": self.M,
            "observationPoint": self.M,
            "pointStart": self.P + 1,
            "cameraStart": self.C + 1,
            "cameraObservations": self.M,
            "cameraActive": self.C,
        }
        return self.index[lay[name]: lay[name] + lengths[name]]


def pack_arrays(poses: np.ndarray, models: np.ndarray, points: np.ndarray,
                track_lengths: np.ndarray, obs_camera: np.ndarray, pixels: np.ndarray,
                active: np.ndarray) -> PackedScene:
    """Pack a scene given as arrays.

    poses (C,12) R row-major then t; models (C,6) fx fy cx cy k1 k2; points (P,3);
    track_lengths (P,) observations per point; obs_camera (M,) camera index of each observation in
    point-major input order; pixels (M,2); active (C,) variable-camera flags.
    """
    poses = np.ascontiguousarray(poses, dtype=np.float64).reshape(-1, 12)
    models = np.ascontiguousarray(models, dtype=np.float64).reshape(-1, 6)
    points = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    C, P = poses.shape[0], points.shape[0]
    track_lengths = np.asarray(track_lengths, dtype=np.int64)
    obs_camera = np.asarray(obs_camera, dtype=np.uint32)
    M = int(obs_camera.shape[0])
    if track_lengths.shape[0] != P or int(track_lengths.sum()) != M:
        raise ValueError("track lengths do not match the observations")
    scene = np.concatenate([
        poses.reshape(-1), points.reshape(-1), models.reshape(-1),
        np.ascontiguousarray(pixels, dtype=np.float64).reshape(-1)])
    lay = bundle_index_layout(C, P, M)
    index = np.zeros(lay["size"], dtype=np.uint32)
    index[:M] = obs_camera
    index[M:2 * M] = np.repeat(np.arange(P, dtype=np.uint32), track_lengths)
    point_start = np.zeros(P + 1, dtype=np.uint64)
    np.cumsum(track_lengths, out=point_start[1:])
    index[lay["pointStart"]: lay["pointStart"] + P + 1] = point_start
    counts = np.bincount(obs_camera, minlength=C).astype(np.uint64)
    camera_start = np.zeros(C + 1, dtype=np.uint64)
    np.cumsum(counts, out=camera_start[1:])
    index[lay["cameraStart"]: lay["cameraStart"] + C + 1] = camera_start
    # a stable sort keeps each camera's observations in ascending observation order
    index[lay["cameraObservations"]: lay["cameraObservations"] + M] = np.argsort(
        obs_camera, kind="stable").astype(np.uint32)
    index[lay["cameraActive"]: lay["cameraActive"] + C] = np.asarray(active, dtype=np.uint32)
    return PackedScene(C, P, M, scene, index)


def finite_gpu_array(array: np.ndarray) -> bool:
    """`finiteGpuArray` (src/geometry/gpu/bounds.ts:19-26): finite, |x| <= 1e9, 0 or |x| >= 1e-30."""
    a = np.abs(array)
    return bool(np.all(np.isfinite(array)) and np.all(a <= 1e9) and np.all((a == 0) | (a >= 1e-30)))
