"""The fixed-scene observation cases of the reference's indexed-observation tests, restated
(tests/typeshade/indexed-observations-cases.mjs:7-131): three to 65 views along x looking down
+z, landmarks on a line with pixels built from the exact Brown projection plus known offsets."""
from __future__ import annotations

import copy
import math


def observation_model(points: int = 43, views: int = 3, depth: float = 10, baseline: float = 0.35,
                      radial: bool = True) -> dict:
    """`observationModel` (indexed-observations-cases.mjs:7-47): pixel x carries the offset
    ((p + i) % 4) * 0.03 px, so every error is that offset."""
    camera = {"fx": 730.0, "fy": 710.0, "cx": 312.5, "cy": 247.25,
              "k1": 0.08 if radial else 0.0, "k2": -0.01 if radial else 0.0}
    registered = [{"id": 10 + 7 * i, "camera": dict(camera),
                   "pose": {"R": [1.0, 0, 0, 0, 1, 0, 0, 0, 1], "t": [-i * baseline, -0.0, 0.0]}}
                  for i in range(views)]
    landmarks = []
    for p in range(points):
        X = [-0.7 + 0.04 * p, 0.2 * math.sin(p), depth + 0.03 * p]
        obs = []
        for i, v in enumerate(registered):
            x = (X[0] - i * baseline) / X[2]
            y = X[1] / X[2]
            r2 = x * x + y * y
            d = 1 + camera["k1"] * r2 + camera["k2"] * r2 * r2
            obs.append({"viewId": v["id"], "featureId": p,
                        "pixel": [camera["fx"] * x * d + camera["cx"] + ((p + i) % 4) * 0.03,
                                  camera["fy"] * y * d + camera["cy"]]})
        landmarks.append({"trackId": 1000 + p, "xyz": X, "observations": obs})
    return {"views": registered, "landmarks": landmarks}


def observation_cases() -> list:
    """`observationCases` (indexed-observations-cases.mjs:49-131)."""
    mixed = observation_model(points=5)
    mixed["landmarks"][1]["observations"][2]["pixel"][0] += 40
    mixed["landmarks"][2]["xyz"][2] = -10
    unused = observation_model(points=2)
    extra = copy.deepcopy(unused["views"][0])
    extra["id"] = 9999
    extra["camera"]["fx"] = 1e8
    unused["views"].append(extra)
    shuffled = observation_model(points=5, views=4)
    shuffled["views"].reverse()
    shuffled["landmarks"].reverse()
    for p in shuffled["landmarks"]:
        p["observations"] = [p["observations"][i] for i in (2, 0, 3, 1)]
    calibrated = observation_model(points=5, views=4)
    for i, v in enumerate(calibrated["views"]):
        angle = i * 0.025
        c, s = math.cos(angle), math.sin(angle)
        v["pose"]["R"] = [c, 0, s, 0, 1, 0, -s, 0, c]
        v["camera"].update({"fx": 600 + 41 * i, "fy": 640 + 27 * i, "cx": 301 + 3 * i,
                            "k1": -0.03 + i * 0.025, "k2": 0.001 * i})
    for p in calibrated["landmarks"]:
        for i, o in enumerate(p["observations"]):
            v = calibrated["views"][i]
            k, X, R, t = v["camera"], p["xyz"], v["pose"]["R"], v["pose"]["t"]
            q = [R[3 * r] * X[0] + R[3 * r + 1] * X[1] + R[3 * r + 2] * X[2] + t[r]
                 for r in range(3)]
            x, y = q[0] / q[2], q[1] / q[2]
            r2 = x * x + y * y
            d = 1 + k["k1"] * r2 + k["k2"] * r2 * r2
            o["pixel"] = [k["fx"] * x * d + k["cx"] + 0.02, k["fy"] * y * d + k["cy"]]
    guard = {"maxRelativeDepthStdDev": 0.1}
    return [
        {"name": "empty", "model": {"views": [], "landmarks": []}},
        {"name": "empty-guarded", "model": {"views": [], "landmarks": []}, "options": guard},
        {"name": "unused-view-outside-gpu-bounds", "model": unused},
        {"name": "radial-tail-129", "model": observation_model()},
        {"name": "shuffled-view-point-observation-indices", "model": shuffled, "options": guard},
        {"name": "different-calibration-and-rotated-poses", "model": calibrated,
         "options": guard},
        {"name": "radial-tail-129-guarded", "model": observation_model(), "options": guard},
        {"name": "mixed-inliers-behind-camera", "model": mixed, "options": guard},
        {"name": "far-depth-rejected",
         "model": observation_model(points=3, baseline=0.0001, depth=2000),
         "options": {"maxRelativeDepthStdDev": 0.01}},
        {"name": "long-track-65", "model": observation_model(points=1, views=65, baseline=0.01),
         "options": guard},
    ]
