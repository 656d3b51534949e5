"""Seeded synthetic bundle-adjustment scenes.

Two families:

* The reference tests' closed-form generators, restated so that the same inputs can be run here:
  `pose_fixture` (tests/typeshade/geometry-cases.mjs:2-60), `bundle_fixture` (:166-219),
  `ring_bundle_fixture` (:252-314) and `joint_focal_fixture` (tests/typeshade/joint-focal-cases.mjs:2-87).
  They return the reference's object model (see typecv_amd/bundle.py).
* Benchmark-scale scenes generated directly in packed-array form (no per-observation Python
  objects): `ring_scene` — cameras on a radius-5 ring around a cube of points — and
  `trajectory_scene` — cameras along a smooth closed trajectory looking outward at points seen by
  5–20 nearby cameras (geometric visibility), with pixel noise, uniform outliers and a common focal
  error.
"""
from __future__ import annotations

import math

import numpy as np

from .packing import PackedScene, pack_arrays

# ------------------------------------------------------------------ reference test generators


def _xorshift(seed: int):
    """`seededRandom`, src/geometry/numerics.ts:133-141 (a zero seed takes 0x9e3779b9)."""
    state = [(seed & 0xFFFFFFFF) or 0x9E3779B9]

    def next_value() -> float:
        s = state[0]
        s ^= (s << 13) & 0xFFFFFFFF
        s ^= s >> 17
        s ^= (s << 5) & 0xFFFFFFFF
        state[0] = s
        return s / 2**32

    return next_value


def pose_fixture(count=97, noise=0.3, outliers=0.25, planar=False, scale=1, radial=True):
    """`poseFixture`, tests/typeshade/geometry-cases.mjs:2-60 (world points, pixels, camera)."""
    nxt = _xorshift(731)
    a, b = 0.11, -0.03
    ca, sa, cb, sb = math.cos(a), math.sin(a), math.cos(b), math.sin(b)
    R = [ca, sa * sb, sa * cb, 0, cb, -sb, -sa, ca * sb, ca * cb]
    t = [0.3 * scale, -0.1 * scale, 0.2 * scale]
    camera = {"fx": 900, "fy": 920, "cx": 480, "cy": 270,
              "k1": -0.05 if radial else 0, "k2": 0.005 if radial else 0}
    world = np.zeros(count * 3)
    pixels = np.zeros(count * 2)
    for i in range(count):
        x0 = (nxt() - 0.5) * 4 * scale
        x1 = (nxt() - 0.5) * 3 * scale
        x2 = (6 if planar else 4 + nxt() * 5) * scale
        X = [x0, x1, x2]
        q = [t[j] + R[j * 3] * X[0] + R[j * 3 + 1] * X[1] + R[j * 3 + 2] * X[2] for j in range(3)]
        x, y = q[0] / q[2], q[1] / q[2]
        r2 = x * x + y * y
        d = 1 + camera["k1"] * r2 + camera["k2"] * r2 * r2
        world[3 * i: 3 * i + 3] = X
        u = camera["fx"] * x * d + camera["cx"] + (nxt() - 0.5) * noise * 2
        v = camera["fy"] * y * d + camera["cy"] + (nxt() - 0.5) * noise * 2
        pixels[2 * i: 2 * i + 2] = [u, v]
        if i < math.floor(count * outliers):
            pixels[2 * i: 2 * i + 2] = [50 + i * 17, 100 + i * 7]
    return {"world": world, "pixels": pixels, "camera": camera, "truth": {"R": R, "t": t}}


def _pose(R, t):
    return {"R": np.array(R, dtype=np.float64), "t": np.array(t, dtype=np.float64)}


def update_pose(pose, delta) -> dict:
    """`updatePose`, src/geometry/pose.ts:53-57: R' = exp(w) R, t' = exp(w) t + v."""
    dR = _rotation_exp(delta[:3]).reshape(3, 3)
    R = np.asarray(pose["R"], dtype=np.float64).reshape(3, 3)
    t = np.asarray(pose["t"], dtype=np.float64)
    return _pose((dR @ R).reshape(9), dR @ t + np.asarray(delta[3:6], dtype=np.float64))


def _rotation_exp(w) -> np.ndarray:
    """`rotationExp`, src/geometry/pose.ts:40-51: the left SO(3) exponential, row-major."""
    x, y, z = (float(v) for v in w)
    theta = math.sqrt(x * x + y * y + z * z)
    a = 1 - theta * theta / 6 if theta < 1e-7 else math.sin(theta) / theta
    b = 0.5 - theta * theta / 24 if theta < 1e-7 else (1 - math.cos(theta)) / (theta * theta)
    W = np.array([[0, -z, y], [z, 0, -x], [-y, x, 0]])
    return (np.eye(3) + a * W + b * (W @ W)).reshape(9)


def reference_scene(count=90, views=6, noise=0.0, seed=731, camera=None):
    """`scene`, tests/scene.ts:10-61: `count` points in a box in front of `views` cameras stepping
    along x (ids 10, 20, ...), every point observed by every view with uniform pixel noise of
    half-width `noise`. Returns (reconstruction, rng) with the generator left where the
    reference's is."""
    camera = dict(camera or {"fx": 850, "fy": 870, "cx": 480, "cy": 270})
    camera.setdefault("k1", 0.0)
    camera.setdefault("k2", 0.0)
    rng = _xorshift(seed)
    points = [np.array([(rng() - 0.5) * 4, (rng() - 0.5) * 3, 4 + rng() * 5]) for _ in range(count)]
    registered = []
    for i in range(views):
        R = _rotation_exp([0.008 * i, -0.018 * i, 0.003 * i])
        C = np.array([i * 0.32, math.sin(i) * 0.04, 0.02 * i])
        t = -(R.reshape(3, 3) @ C)
        registered.append({"id": i * 10 + 10, "camera": dict(camera), "pose": _pose(R, t)})
    keypoints = []
    for v in registered:
        kp = []
        for X in points:
            u, w = _project_pinhole_radial(camera, v["pose"]["R"], v["pose"]["t"], X)
            kp.append((u + (rng() - 0.5) * noise * 2, w + (rng() - 0.5) * noise * 2))
        keypoints.append(kp)
    landmarks = [{"trackId": i, "xyz": X.copy(),
                  "observations": [{"viewId": v["id"], "featureId": i,
                                    "pixel": [keypoints[j][i][0], keypoints[j][i][1]]}
                                   for j, v in enumerate(registered)]}
                 for i, X in enumerate(points)]
    return {"views": registered, "landmarks": landmarks}, rng


def _project_pinhole_radial(camera, R, t, X):
    q = [t[j] + R[j * 3] * X[0] + R[j * 3 + 1] * X[1] + R[j * 3 + 2] * X[2] for j in range(3)]
    x, y = q[0] / q[2], q[1] / q[2]
    r2 = x * x + y * y
    d = 1 + camera["k1"] * r2 + camera["k2"] * r2 * r2
    return camera["fx"] * x * d + camera["cx"], camera["fy"] * y * d + camera["cy"]


def bundle_fixture(count=40, views=4, outlier=False, shift=(0.0, 0.0, 0.0)):
    """`bundleFixture`, tests/typeshade/geometry-cases.mjs:166-219: `views` cameras along x with
    perturbed poses (camera 0 exact), `count` points with sinusoidal pixel noise."""
    data = pose_fixture(count=count, noise=0, outliers=0)
    model = {"views": [], "landmarks": []}
    truth = []
    for c in range(views):
        a = 0.02 * c
        ca, sa = math.cos(a), math.sin(a)
        R = [ca, 0, sa, 0, 1, 0, -sa, 0, ca]
        t = [-0.32 * c, 0.02 * math.sin(c), 0.015 * c]
        truth.append((R, t))
        perturbed = a + (0.004 * math.sin(c) if c else 0)
        cp, sp = math.cos(perturbed), math.sin(perturbed)
        Rp = [cp, 0, sp, 0, 1, 0, -sp, 0, cp]
        tp = [t[j] + (0.015 * math.sin(c + j) if c else 0) - Rp[j * 3] * shift[0]
              - Rp[j * 3 + 1] * shift[1] - Rp[j * 3 + 2] * shift[2] for j in range(3)]
        model["views"].append({"id": 10 + 10 * c, "camera": dict(data["camera"]), "pose": _pose(Rp, tp)})
    for p in range(count):
        X = list(data["world"][p * 3: p * 3 + 3])
        observations = []
        for c, (R, t) in enumerate(truth):
            u, v = _project_pinhole_radial(data["camera"], R, t, X)
            pixel = [u + 0.1 * math.sin(p * 3 + c), v + 0.1 * math.cos(p + c * 3)]
            if outlier and p == 0 and c == 2:
                pixel = [pixel[0] + 40, pixel[1] - 20]
            observations.append({"viewId": model["views"][c]["id"], "featureId": p, "pixel": pixel})
        model["landmarks"].append({
            "trackId": p,
            "xyz": np.array([X[j] + shift[j] + 0.035 * math.sin(p + j) for j in range(3)]),
            "observations": observations,
        })
    return model


def ring_bundle_fixture(views=320, points=4800, track=5, global_points=0):
    """`ringBundleFixture`, tests/typeshade/geometry-cases.mjs:252-314: a ring of cameras around a
    cloud of points, each normally seen by `track` consecutive cameras."""
    nxt = _xorshift(977)
    camera = {"fx": 900, "fy": 920, "cx": 480, "cy": 270, "k1": -0.05, "k2": 0.005}

    def rotation(a):
        return [math.cos(a), 0, math.sin(a), 0, 1, 0, -math.sin(a), 0, math.cos(a)]

    truth = []
    model = {"views": [], "landmarks": []}
    for c in range(views):
        a = 2 * math.pi * c / views
        R, t = rotation(a), [0, 0, 10]
        truth.append((R, t))
        Rp = rotation(a + (0.002 * math.sin(c) if c else 0))
        tp = [t[j] + (0.01 * math.sin(c + j) if c else 0) for j in range(3)]
        model["views"].append({"id": 1000 + c, "camera": dict(camera), "pose": _pose(Rp, tp)})
    features = [0] * views
    for p in range(points):
        X = [(nxt() - 0.5) * 3, (nxt() - 0.5) * 3, (nxt() - 0.5) * 3]
        observations = []
        for k in range(views if p < global_points else track):
            c = (p + k) % views
            R, t = truth[c]
            u, v = _project_pinhole_radial(camera, R, t, X)
            observations.append({
                "viewId": model["views"][c]["id"], "featureId": features[c],
                "pixel": [u + 0.1 * math.sin(p * 3 + c), v + 0.1 * math.cos(p + c * 3)],
            })
            features[c] += 1
        model["landmarks"].append({
            "trackId": p, "xyz": np.array([X[j] + 0.02 * math.sin(p + j) for j in range(3)]),
            "observations": observations,
        })
    return model


def joint_focal_fixture(views=6, points=60, focal_scale=0.75, noise=0.0, radial=True):
    """`jointFocalFixture`, tests/typeshade/joint-focal-cases.mjs:2-87: returns
    {input, truth, expectedScale}; the input focal lengths are scaled by `focal_scale`."""

    def multiply(a, b):
        return [sum(a[(i // 3) * 3 + k] * b[k * 3 + i % 3] for k in range(3)) for i in range(9)]

    def rotation(x, y, z):
        cx, sx, cy, sy, cz, sz = math.cos(x), math.sin(x), math.cos(y), math.sin(y), math.cos(z), math.sin(z)
        return multiply([cz, -sz, 0, sz, cz, 0, 0, 0, 1],
                        multiply([cy, 0, sy, 0, 1, 0, -sy, 0, cy], [1, 0, 0, 0, cx, -sx, 0, sx, cx]))

    camera = {"fx": 800, "fy": 840, "cx": 511.5, "cy": 383.5,
              "k1": -0.04 if radial else 0, "k2": 0.004 if radial else 0}
    truth = {"views": [], "landmarks": []}
    inp = {"views": [], "landmarks": []}
    for c in range(views):
        angles = [0.03 * math.sin(c), 0.075 * c, 0.035 * math.sin(c * 0.7)]
        R = rotation(*angles)
        t = [-0.4 * c, 0.12 * math.sin(c * 1.3), 0.08 * (math.cos(c) - 1)]
        truth["views"].append({"id": 10 + 10 * c, "camera": dict(camera), "pose": _pose(R, t)})
        inp["views"].append({
            "id": 10 + 10 * c,
            "camera": {**camera, "fx": camera["fx"] * focal_scale, "fy": camera["fy"] * focal_scale},
            "pose": _pose(rotation(*[v + (0.003 * math.sin(c + j) if c else 0) for j, v in enumerate(angles)]),
                          [v + (0.012 * math.sin(c + j) if c else 0) for j, v in enumerate(t)]),
        })
    seed = [117]

    def nxt():
        seed[0] = (seed[0] * 1664525 + 1013904223) & 0xFFFFFFFF
        return seed[0] / 2**32

    for p in range(points):
        X = [(nxt() - 0.5) * 3, (nxt() - 0.5) * 2, 3 + nxt() * 6]
        observations = []
        for c, v in enumerate(truth["views"]):
            R, t = v["pose"]["R"], v["pose"]["t"]
            q = [t[j] + sum(R[j * 3 + k] * X[k] for k in range(3)) for j in range(3)]
            x, y = q[0] / q[2], q[1] / q[2]
            r = x * x + y * y
            d = 1 + camera["k1"] * r + camera["k2"] * r * r
            observations.append({
                "viewId": v["id"], "featureId": p,
                "pixel": [camera["fx"] * x * d + camera["cx"] + noise * math.sin(p + c),
                          camera["fy"] * y * d + camera["cy"] + noise * math.cos(p * 2 + c)],
            })
        truth["landmarks"].append({"trackId": p, "xyz": np.array(X), "observations": observations})
        inp["landmarks"].append({"trackId": p,
                                 "xyz": np.array([v + 0.025 * math.sin(p + j) for j, v in enumerate(X)]),
                                 "observations": observations})
    return {"input": inp, "truth": truth, "expectedScale": 1 / focal_scale}


# ------------------------------------------------------------------ conversions


def packed_to_reconstruction(packed: PackedScene, view_ids=None) -> dict:
    """The object model of a packed scene (small scenes; every observation becomes a dict)."""
    C, P = packed.C, packed.P
    ids = list(range(C)) if view_ids is None else list(view_ids)
    poses, models, points, pixels = packed.poses(), packed.models(), packed.points(), packed.pixels()
    obs_camera = packed.section("observationCamera")
    starts = packed.section("pointStart")
    views = [{"id": ids[c],
              "camera": {"fx": models[c, 0], "fy": models[c, 1], "cx": models[c, 2],
                         "cy": models[c, 3], "k1": models[c, 4], "k2": models[c, 5]},
              "pose": {"R": poses[c, :9].copy(), "t": poses[c, 9:].copy()}} for c in range(C)]
    feature = [0] * C
    landmarks = []
    for p in range(P):
        obs = []
        for i in range(int(starts[p]), int(starts[p + 1])):
            c = int(obs_camera[i])
            obs.append({"viewId": ids[c], "featureId": feature[c],
                        "pixel": [float(pixels[i, 0]), float(pixels[i, 1])]})
            feature[c] += 1
        landmarks.append({"trackId": p, "xyz": points[p].copy(), "observations": obs})
    return {"views": views, "landmarks": landmarks}


# ------------------------------------------------------------------ benchmark scenes


def _look_rotation(forward: np.ndarray, up: np.ndarray) -> np.ndarray:
    """Rows: camera x (right), y (down), z (forward) in world coordinates, for each camera."""
    z = forward / np.linalg.norm(forward, axis=-1, keepdims=True)
    y = -(up - np.sum(up * z, axis=-1, keepdims=True) * z)
    y /= np.linalg.norm(y, axis=-1, keepdims=True)
    x = np.cross(y, z)
    return np.stack([x, y, z], axis=-2)


def _axis_angle(axis: np.ndarray, angle: np.ndarray) -> np.ndarray:
    """Rotation matrices of unit axes (n,3) and angles (n,) by Rodrigues' formula."""
    axis = axis / np.linalg.norm(axis, axis=-1, keepdims=True)
    K = np.zeros(axis.shape[:-1] + (3, 3))
    K[..., 0, 1], K[..., 0, 2] = -axis[..., 2], axis[..., 1]
    K[..., 1, 0], K[..., 1, 2] = axis[..., 2], -axis[..., 0]
    K[..., 2, 0], K[..., 2, 1] = -axis[..., 1], axis[..., 0]
    s, c = np.sin(angle)[..., None, None], np.cos(angle)[..., None, None]
    return np.eye(3) + s * K + (1 - c) * (K @ K)


def _perturb_poses(rng, R, centers, rotation_deg, translation, fixed=(0,)):
    C = R.shape[0]
    axes = rng.normal(size=(C, 3))
    dR = _axis_angle(axes, np.full(C, math.radians(rotation_deg)))
    dirs = rng.normal(size=(C, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    Rp = dR @ R
    t_true = -np.einsum("cij,cj->ci", R, centers)
    tp = t_true + translation * dirs
    for c in fixed:
        Rp[c], tp[c] = R[c], t_true[c]
    return Rp, tp, t_true


def _project(R, t, X, f, cx, cy):
    q = np.einsum("...ij,...j->...i", R, X) + t
    with np.errstate(divide="ignore", invalid="ignore"):
        u = f * q[..., 0] / q[..., 2] + cx
        v = f * q[..., 1] / q[..., 2] + cy
    return u, v, q[..., 2]


def ring_scene(seed: int = 1, cameras: int = 10, points: int = 2000, radius: float = 5.0,
               focal: float = 800.0, width: int = 640, height: int = 480, noise_px: float = 0.5,
               rotation_deg: float = 2.0, translation: float = 0.05, point_sigma: float = 0.03):
    """`cameras` pinhole cameras evenly on a horizontal ring of `radius` looking at the origin,
    `points` points uniform in [-1, 1]^3, an observation wherever a point projects in frame, pixel
    noise N(0, noise_px^2), poses of cameras 1.. perturbed by `rotation_deg` about a random axis
    and `translation` in a random direction, points by N(0, point_sigma^2). Camera 0 is the anchor,
    camera 1 the scale view. Returns (PackedScene, info)."""
    rng = np.random.default_rng(seed)
    angles = 2 * np.pi * np.arange(cameras) / cameras
    centers = np.stack([radius * np.cos(angles), np.zeros(cameras), radius * np.sin(angles)], axis=1)
    R = _look_rotation(-centers, np.array([0.0, 1.0, 0.0]))
    X = rng.uniform(-1, 1, size=(points, 3))
    t_true = -np.einsum("cij,cj->ci", R, centers)
    u, v, z = _project(R[None], t_true[None], X[:, None, :], focal, width / 2, height / 2)
    visible = (z > 1e-6) & (u >= 0) & (u < width) & (v >= 0) & (v < height)
    keep = visible.sum(axis=1) >= 2
    X, u, v, visible = X[keep], u[keep], v[keep], visible[keep]
    lengths = visible.sum(axis=1)
    pid, cam = np.nonzero(visible)  # point-major, cameras ascending within a point
    pixels = np.stack([u[pid, cam], v[pid, cam]], axis=1) + rng.normal(0, noise_px, size=(pid.size, 2))
    Rp, tp, _ = _perturb_poses(rng, R, centers, rotation_deg, translation)
    Xp = X + rng.normal(0, point_sigma, size=X.shape)
    poses = np.concatenate([Rp.reshape(cameras, 9), tp], axis=1)
    models = np.tile([focal, focal, width / 2, height / 2, 0.0, 0.0], (cameras, 1))
    active = np.ones(cameras, dtype=np.uint32)
    active[0] = 0
    packed = pack_arrays(poses, models, Xp, lengths, cam.astype(np.uint32), pixels, active)
    truth = {"R": R, "t": t_true, "X": X}
    return packed, {"truth": truth, "anchor": 0, "scale_view": 1}


def trajectory_scene(cameras: int = 500, points: int = 200_000, seed: int = 2,
                     spacing: float = 0.25, focal: float = 800.0, width: int = 640,
                     height: int = 480, depth=(4.0, 12.0), track=(5, 20), window: int = 24,
                     noise_px: float = 0.7, outliers: float = 0.02, focal_error: float = 0.05,
                     rotation_deg: float = 0.5, translation: float = 0.05,
                     point_sigma: float = 0.03, chunk: int = 131072):
    """Cameras along a seeded smooth closed trajectory (a loop of circumference cameras·spacing
    with low-order radius, height and yaw undulations) looking outward; each point is placed in the
    frustum of a home camera at a depth in `depth` and observed by the 5–20 nearest cameras of a
    ±`window` neighbourhood in which it projects in frame (geometric visibility; points sorted by
    home camera). Pixels get N(0, noise_px^2) noise, a fraction `outliers` is replaced by uniform
    in-frame pixels, every fx, fy is scaled by one common factor U(1-focal_error, 1+focal_error),
    poses of cameras 1.. are perturbed by `rotation_deg` / `translation`, points by
    N(0, point_sigma^2). Returns (PackedScene, info)."""
    rng = np.random.default_rng(seed)
    C = cameras
    # a loop of radius >= 20 with cameras `spacing` apart: closed when they go round, else an arc
    radius = max(C * spacing / (2 * np.pi), 20.0)
    closed = C * spacing >= 2 * np.pi * radius
    theta = np.arange(C) * (spacing / radius)
    k = np.arange(1, 4)
    a, phase = rng.uniform(-0.04, 0.04, 3), rng.uniform(0, 2 * np.pi, 3)
    b, phase2 = rng.uniform(-0.3, 0.3, 3), rng.uniform(0, 2 * np.pi, 3)
    yaw_a, phase3 = rng.uniform(-0.15, 0.15, 3), rng.uniform(0, 2 * np.pi, 3)
    r = radius * (1 + (a[None] * np.sin(k[None] * theta[:, None] + phase[None])).sum(1))
    h = (b[None] * np.sin(k[None] * theta[:, None] + phase2[None])).sum(1)
    yaw = (yaw_a[None] * np.sin(k[None] * theta[:, None] + phase3[None])).sum(1)
    centers = np.stack([r * np.cos(theta), h, r * np.sin(theta)], axis=1)
    heading = theta + yaw
    forward = np.stack([np.cos(heading), np.zeros(C), np.sin(heading)], axis=1)
    R = _look_rotation(forward, np.array([0.0, 1.0, 0.0]))
    t_true = -np.einsum("cij,cj->ci", R, centers)
    window = min(window, (C - 1) // 2)
    cx, cy = width / 2, height / 2
    # candidate offsets ordered by distance from the home camera: 0, -1, +1, -2, +2, ...
    offsets = np.array([0] + [s * d for d in range(1, window + 1) for s in (-1, 1)])
    home = np.sort(rng.integers(0, C, size=points))
    target = np.minimum(track[0] + np.floor((track[1] - track[0] + 1) * rng.uniform(size=points) ** 2),
                        track[1]).astype(np.int64)
    X = np.empty((points, 3))
    lengths = np.empty(points, dtype=np.int64)
    obs_pid, obs_cam, obs_u, obs_v = [], [], [], []
    todo = np.arange(points)
    for attempt in range(64):
        if not todo.size:
            break
        retry = []
        for s in range(0, todo.size, chunk):
            ids = todo[s: s + chunk]
            n = ids.size
            hc = home[ids]
            d = rng.uniform(depth[0], depth[1], n)
            uu, vv = rng.uniform(0, width, n), rng.uniform(0, height, n)
            ray = np.stack([(uu - cx) / focal, (vv - cy) / focal, np.ones(n)], axis=1) * d[:, None]
            Xw = centers[hc] + np.einsum("nji,nj->ni", R[hc], ray)
            raw = hc[:, None] + offsets[None, :]
            inside = (raw >= 0) & (raw < C)
            cand = raw % C
            u, v, z = _project(R[cand], t_true[cand], Xw[:, None, :], focal, cx, cy)
            vis = (z > 0.5) & (u >= 0) & (u < width) & (v >= 0) & (v < height)
            if not closed:
                vis &= inside
            rank = np.cumsum(vis, axis=1)
            take = vis & (rank <= target[ids][:, None])
            got = take.sum(axis=1)
            ok = got >= track[0]
            retry.append(ids[~ok])
            X[ids[ok]] = Xw[ok]
            lengths[ids[ok]] = got[ok]
            rows, cols = np.nonzero(take & ok[:, None])
            obs_pid.append(ids[rows])
            obs_cam.append(cand[rows, cols])
            obs_u.append(u[rows, cols])
            obs_v.append(v[rows, cols])
        todo = np.concatenate(retry)
    if todo.size:
        raise ValueError(f"{todo.size} points found fewer than {track[0]} observing cameras")
    pid = np.concatenate(obs_pid)
    cam = np.concatenate(obs_cam)
    order = np.lexsort((cam, pid))  # point-major, cameras ascending within a point
    obs_camera = cam[order].astype(np.uint32)
    pixels = np.stack([np.concatenate(obs_u)[order], np.concatenate(obs_v)[order]], axis=1)
    M = obs_camera.size
    pixels += rng.normal(0, noise_px, size=pixels.shape)
    bad = rng.uniform(size=M) < outliers
    pixels[bad] = np.stack([rng.uniform(0, width, bad.sum()), rng.uniform(0, height, bad.sum())], 1)
    factor = rng.uniform(1 - focal_error, 1 + focal_error)
    Rp, tp, _ = _perturb_poses(rng, R, centers, rotation_deg, translation)
    Xp = X + rng.normal(0, point_sigma, size=X.shape)
    poses = np.concatenate([Rp.reshape(C, 9), tp], axis=1)
    models = np.tile([focal * factor, focal * factor, cx, cy, 0.0, 0.0], (C, 1))
    active = np.ones(C, dtype=np.uint32)
    active[0] = 0
    packed = pack_arrays(poses, models, Xp, lengths, obs_camera, pixels, active)
    return packed, {"truth": {"R": R, "t": t_true, "X": X, "focal": focal}, "focal_factor": factor,
                    "outliers": int(bad.sum()), "anchor": 0, "scale_view": 1}


def local_bundles(count: int = 256, seed: int = 4, cameras=(5, 40), points=(500, 20_000),
                  track=(3, 12), **scene_options):
    """`count` independent local bundles of uneven size: cameras ~ U{cameras}, points ~ U{points}
    per bundle, each point seen by 3–12 cameras of its bundle (at most the bundle's cameras), the
    trajectory generator's noise, outliers and focal error otherwise. Returns a list of
    (PackedScene, info)."""
    rng = np.random.default_rng(seed)
    out = []
    for b in range(count):
        C = int(rng.integers(cameras[0], cameras[1] + 1))
        P = int(rng.integers(points[0], points[1] + 1))
        hi = min(track[1], C)
        lo = min(track[0], hi)
        out.append(trajectory_scene(cameras=C, points=P, seed=int(rng.integers(1 << 30)),
                                    track=(lo, hi), **scene_options))
    return out
