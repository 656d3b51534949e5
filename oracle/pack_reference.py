"""TEST INFRASTRUCTURE: a loop-for-loop restatement of the reference's `packBundleScene`
(src/sfm/bundle/gpu/bundle-host.ts:177-218) and `bundleIndexLayout` (:121-131), used to check the
vectorised packer of typecv_amd/packing.py bit for bit."""
from __future__ import annotations

import numpy as np


def pack_bundle_scene(model, by_id, active):
    views, landmarks = model["views"], model["landmarks"]
    C, P = len(views), len(landmarks)
    M = sum(len(p["observations"]) for p in landmarks)
    scene = np.zeros(18 * C + 3 * P + 2 * M)
    index = np.zeros(3 * M + P + 2 * C + 2, dtype=np.uint32)
    observation_point, point_start = M, 2 * M
    camera_start, camera_observations = 2 * M + P + 1, 2 * M + P + C + 2
    camera_active = 3 * M + P + C + 2
    lists = [[] for _ in range(C)]
    for c in range(C):
        v = views[c]
        scene[c * 12: c * 12 + 9] = v["pose"]["R"]
        scene[c * 12 + 9: c * 12 + 12] = v["pose"]["t"]
        k = v["camera"]
        at = 12 * C + 3 * P + c * 6
        scene[at: at + 6] = [k["fx"], k["fy"], k["cx"], k["cy"], k.get("k1", 0) or 0, k.get("k2", 0) or 0]
        index[camera_active + c] = 1 if active[c] else 0
    obs = 0
    for p in range(P):
        point = landmarks[p]
        scene[12 * C + 3 * p: 12 * C + 3 * p + 3] = point["xyz"]
        index[point_start + p] = obs
        for o in point["observations"]:
            c = by_id[o["viewId"]]
            index[obs] = c
            index[observation_point + obs] = p
            scene[18 * C + 3 * P + 2 * obs: 18 * C + 3 * P + 2 * obs + 2] = o["pixel"]
            lists[c].append(obs)
            obs += 1
    index[point_start + P] = obs
    cursor = 0
    for c in range(C):
        index[camera_start + c] = cursor
        index[camera_observations + cursor: camera_observations + cursor + len(lists[c])] = lists[c]
        cursor += len(lists[c])
    index[camera_start + C] = cursor
    return scene, index
