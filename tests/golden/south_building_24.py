"""Builds `tests/golden/south_building_24.npz`: the real-data bundle-adjustment problem of the
reference's budget measurements (24 views of COLMAP's south-building sample with their SIFT
matches, 3,000 triangulated landmarks, 12,706 observations; `docs/experiments/gpu-budget/
south-building-24-ba-{fixed,shared}.json` record its initial and final RMSE).

This script restates, in Python, the scene construction of the reference's
`scripts/measure-ba-budget.mjs:53-148` — orthonormalised reference poses, union-find tracks over the
pairwise matches, `triangulateTrack` (`src/geometry/multiview.ts:194-300`, maxErrorPx 2, minAngleDeg
1) with its helpers (`triangulateLinear` :107-132, `nullspace` numerics.ts:2-69, `refineTrackPoint`
:137-172, `angleAt` :173-192, `unprojectPixel` camera.ts:64-105, `cameraProjection` camera.ts:28-62,
`observationJacobian` pose.ts:61-88, `solveSystem` numerics.ts:71-100), the 3,000 landmarks with
the most observations, and the deterministic perturbation of poses and points. It reads the
reference's data files (`tests/fixtures/lab-images/south-building-24/sample.json`,
`colmap-sift.bin`) and runs only where they are; the problem it writes is the fixture the tests
load.

    python tests/golden/south_building_24.py [reference root]
"""
from __future__ import annotations

import json
import math
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
OUT = HERE / "south_building_24.npz"
MAX_POINTS = 3000
MAX_ERROR_PX = 2.0
MIN_ANGLE_DEG = 1.0
MAX_PAIRS = 32  # TRIANGULATION_DEFAULTS.maxPairHypotheses, operator-defaults.ts:23-27


def unit(v):
    n = math.hypot(*v)
    return [x / n for x in v]


def orthonormal(R):
    """measure-ba-budget.mjs:47-52: Gram-Schmidt of the first two rows, third by cross product."""
    x = unit(R[0:3])
    b = R[3:6]
    d = x[0] * b[0] + x[1] * b[1] + x[2] * b[2]
    y = unit([b[i] - d * x[i] for i in range(3)])
    z = [x[1] * y[2] - x[2] * y[1], x[2] * y[0] - x[0] * y[2], x[0] * y[1] - x[1] * y[0]]
    return x + y + z


def radial_limit_squared(k1, k2):
    if k2 == 0:
        return -1 / (3 * k1) if k1 < 0 else math.inf
    A, Bq = 5 * k2, 3 * k1
    D = Bq * Bq - 4 * A
    if D < 0:
        return math.inf
    q = -0.5 * (Bq + (1 if Bq >= 0 else -1) * math.sqrt(D))
    roots = [r for r in ((q / A) if A else math.inf, (1 / q) if q else math.inf)
             if r > 0 and math.isfinite(r)]
    return min(roots) if roots else math.inf


def camera_projection(k, q):
    """Pixel and 2x3 derivative with respect to camera-frame XYZ, or None."""
    X, Y, Z = q
    if not (Z > 1e-12) or not all(math.isfinite(v) for v in q):
        return None
    x, y = X / Z, Y / Z
    r2 = x * x + y * y
    k1, k2 = k["k1"], k["k2"]
    if r2 >= radial_limit_squared(k1, k2):
        return None
    d = 1 + k1 * r2 + k2 * r2 * r2
    g = 2 * (k1 + 2 * k2 * r2)
    xx, xy, yy = d + g * x * x, g * x * y, d + g * y * y
    if not (d > 0 and 1 + 3 * k1 * r2 + 5 * k2 * r2 * r2 > 0):
        return None
    pixel = (k["fx"] * x * d + k["cx"], k["fy"] * y * d + k["cy"])
    J = (k["fx"] * xx / Z, k["fx"] * xy / Z, -k["fx"] * (xx * x + xy * y) / Z,
         k["fy"] * xy / Z, k["fy"] * yy / Z, -k["fy"] * (xy * x + yy * y) / Z)
    if not all(math.isfinite(v) for v in pixel + J):
        return None
    return pixel, J


def transform(pose, X):
    R, t = pose["R"], pose["t"]
    return [R[i * 3] * X[0] + R[i * 3 + 1] * X[1] + R[i * 3 + 2] * X[2] + t[i] for i in range(3)]


def center(pose):
    R, t = pose["R"], pose["t"]
    return [-R[i] * t[0] - R[i + 3] * t[1] - R[i + 6] * t[2] for i in range(3)]


def project(o, X):
    p = camera_projection(o["camera"], transform(o["pose"], X))
    return p[0] if p else None


def point_jacobian(o, X):
    q = transform(o["pose"], X)
    p = camera_projection(o["camera"], q)
    if p is None:
        return None
    pixel, J = p
    R = o["pose"]["R"]
    point = [sum(J[i * 3 + k] * R[k * 3 + j] for k in range(3)) for i in range(2) for j in range(3)]
    return pixel, point


def unproject(k, pixel):
    xd = (pixel[0] - k["cx"]) / k["fx"]
    yd = (pixel[1] - k["cy"]) / k["fy"]
    rd = math.hypot(xd, yd)
    if not math.isfinite(rd):
        return None
    if rd == 0:
        return (0.0, 0.0)
    k1, k2 = k["k1"], k["k2"]
    limit2 = radial_limit_squared(k1, k2)
    if math.isfinite(limit2) and rd >= math.sqrt(limit2) * (1 + k1 * limit2 + k2 * limit2 * limit2):
        return None
    r = rd
    for _ in range(50):
        r2 = r * r
        d = 1 + k1 * r2 + k2 * r2 * r2
        slope = 1 + 3 * k1 * r2 + 5 * k2 * r2 * r2
        if not (slope > 1e-10 and d > 0):
            return None
        residual = r * d - rd
        if abs(residual) <= 1e-12 * max(1, rd):
            return (xd * r / rd, yd * r / rd)
        step, accepted = residual / slope, False
        for _ in range(25):
            nxt = r - step
            n2 = nxt * nxt
            if (nxt >= 0 and n2 < limit2 and 1 + 3 * k1 * n2 + 5 * k2 * n2 * n2 > 0
                    and abs(nxt * (1 + k1 * n2 + k2 * n2 * n2) - rd) < abs(residual)):
                r, accepted = nxt, True
                break
            step *= 0.5
        if not accepted:
            return None
    return None


def nullspace(rows, columns=4):
    """One-sided Jacobi SVD of the scaled design matrix: the right singular vector of the smallest
    singular value and the singular values in ascending order."""
    scale = max(abs(x) for row in rows for x in row)
    if not scale > 0:
        raise ValueError("zero design matrix")
    a = np.array(rows, dtype=np.float64) / scale
    v = np.eye(columns)
    for _ in range(160):
        changed = False
        for p in range(columns):
            for q in range(p + 1, columns):
                ap, aq = a[:, p], a[:, q]
                aa, bb, ab = float(ap @ ap), float(aq @ aq), float(ap @ aq)
                if abs(ab) <= 1e-13 * math.sqrt(aa * bb) or aa < 1e-30 or bb < 1e-30:
                    continue
                tau = (bb - aa) / (2 * ab)
                t = (1 if tau >= 0 else -1) / (abs(tau) + math.hypot(1, tau))
                c = 1 / math.hypot(1, t)
                s = c * t
                x, y = a[:, p].copy(), a[:, q].copy()
                a[:, p], a[:, q] = c * x - s * y, s * x + c * y
                x, y = v[:, p].copy(), v[:, q].copy()
                v[:, p], v[:, q] = c * x - s * y, s * x + c * y
                changed = True
        if not changed:
            break
    else:
        raise ValueError("Jacobi SVD did not converge")
    values = np.sqrt((a * a).sum(axis=0))
    order = sorted(range(columns), key=lambda i: values[i])
    return v[:, order[0]].copy(), values[order] * scale


def triangulate_linear(obs):
    if len(obs) < 2:
        return None
    rows = []
    for o in obs:
        xy = unproject(o["camera"], o["pixel"])
        if xy is None:
            return None
        R, t = o["pose"]["R"], o["pose"]["t"]
        for k in range(2):
            row = [xy[k] * R[6 + j] - R[3 * k + j] for j in range(3)]
            row.append(xy[k] * t[2] - t[k])
            norm = math.hypot(*row)
            if not norm > 0:
                return None
            rows.append([x / norm for x in row])
    try:
        vec, s = nullspace(rows, 4)
    except ValueError:
        return None
    if s[1] < 1e-10 * s[3] or abs(vec[3]) < 1e-12:
        return None
    X = [vec[i] / vec[3] for i in range(3)]
    return X if all(math.isfinite(x) for x in X) else None


def error(o, X):
    p = project(o, X)
    return math.hypot(p[0] - o["pixel"][0], p[1] - o["pixel"][1]) if p else math.inf


def solve_system(A, b):
    n = len(b)
    a = [[A[i * n + j] for j in range(n)] + [b[i]] for i in range(n)]
    scale = max(abs(a[i][j]) for i in range(n) for j in range(n))
    if not scale > 0:
        return None
    for j in range(n):
        p = j
        for i in range(j + 1, n):
            if abs(a[i][j]) > abs(a[p][j]):
                p = i
        if abs(a[p][j]) < 1e-14 * scale:
            return None
        a[p], a[j] = a[j], a[p]
        for i in range(j + 1, n):
            f = a[i][j] / a[j][j]
            for k in range(j, n + 1):
                a[i][k] -= f * a[j][k]
    out = [0.0] * n
    for i in range(n - 1, -1, -1):
        s = a[i][n]
        for j in range(i + 1, n):
            s -= a[i][j] * out[j]
        out[i] = s / a[i][i]
    return out if all(math.isfinite(x) for x in out) else None


def refine_point(X, obs, iterations=10):
    X = list(X)
    lam = 1e-3

    def cost(P):
        return sum(error(o, P) ** 2 for o in obs)

    value = cost(X)
    if not math.isfinite(value):
        return X
    for _ in range(iterations):
        A, b = [0.0] * 9, [0.0] * 3
        for o in obs:
            j = point_jacobian(o, X)
            if j is None:
                return X
            pixel, point = j
            for r in range(2):
                e = pixel[r] - o["pixel"][r]
                for k in range(3):
                    b[k] -= point[r * 3 + k] * e
                    for l in range(3):
                        A[k * 3 + l] += point[r * 3 + k] * point[r * 3 + l]
        for k in range(3):
            A[k * 3 + k] += lam * max(A[k * 3 + k], 1e-10)
        step = solve_system(A, b)
        if step is None:
            break
        cand = [X[k] + step[k] for k in range(3)]
        nxt = cost(cand)
        if nxt < value:
            change = value - nxt
            X, value = cand, nxt
            lam = max(1e-10, lam / 3)
            if change < 1e-12 * max(1, value):
                break
        else:
            lam *= 10
    return X


def angle_at(X, a, b):
    ca, cb = center(a), center(b)
    u = [X[i] - ca[i] for i in range(3)]
    v = [X[i] - cb[i] for i in range(3)]
    cr = [u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]]
    return math.atan2(math.hypot(*cr), sum(u[i] * v[i] for i in range(3))) * 180 / math.pi


def triangulate_track(obs):
    """triangulateTrack with maxErrorPx 2, minAngleDeg 1, no depth-sensitivity bound: (xyz, inlier
    mask) or None."""
    if len(obs) < 2:
        return None
    pairs = []
    for i in range(len(obs)):
        for j in range(i + 1, len(obs)):
            a, b = center(obs[i]["pose"]), center(obs[j]["pose"])
            pairs.append((i, j, math.hypot(*[a[k] - b[k] for k in range(3)])))
    pairs.sort(key=lambda p: (-p[2], p[0], p[1]))
    candidates = [triangulate_linear(obs)] + [triangulate_linear([obs[i], obs[j]])
                                              for i, j, _ in pairs[:MAX_PAIRS]]
    best, best_mask, best_count, best_cost = None, None, 0, math.inf
    for X in candidates:
        if X is None:
            continue
        mask, count, cost = [0] * len(obs), 0, 0.0
        for i, o in enumerate(obs):
            e = error(o, X)
            if e <= MAX_ERROR_PX:
                mask[i] = 1
                count += 1
                cost += e * e
        if count > best_count or (count == best_count and cost < best_cost):
            best, best_count, best_cost, best_mask = X, count, cost, mask
    if best is None or best_count < 2:
        return None
    selected = [o for i, o in enumerate(obs) if best_mask[i]]
    linear = triangulate_linear(selected)

    def cost_of(P):
        return sum(error(o, P) ** 2 for o in selected)

    if linear is not None and cost_of(linear) < cost_of(best):
        best = linear
    refined = refine_point(best, selected)
    if cost_of(refined) < cost_of(best):
        best = refined
    inliers = [1 if error(o, best) <= MAX_ERROR_PX else 0 for o in obs]
    accepted = [o for i, o in enumerate(obs) if inliers[i]]
    if len(accepted) < 2:
        return None
    max_angle = 0.0
    for i in range(len(accepted)):
        for j in range(i + 1, len(accepted)):
            max_angle = max(max_angle, angle_at(best, accepted[i]["pose"], accepted[j]["pose"]))
    if not math.isfinite(max_angle) or max_angle < MIN_ANGLE_DEG:
        return None
    return best, inliers


def wave(i, k):
    return math.sin(12.9898 * (i + 1) + 78.233 * (k + 1))


def build(root: Path) -> dict:
    directory = root / "tests" / "fixtures" / "lab-images" / "south-building-24"
    sample = json.loads((directory / "sample.json").read_text())
    raw = (directory / "colmap-sift.bin").read_bytes()
    feats = sample["features"]
    keypoints = np.frombuffer(raw, dtype="<f4", count=feats["keypointBytes"] // 4)
    matches = np.frombuffer(raw, dtype="<u2", offset=feats["keypointBytes"])
    views = [{"id": f["id"], "camera": {k: float(f["camera"].get(k, 0.0))
                                        for k in ("fx", "fy", "cx", "cy", "k1", "k2")},
              "pose": {"R": orthonormal(f["reference"]["R"]), "t": list(f["reference"]["t"])}}
             for f in sample["frames"]]
    blocks = {b["id"]: b for b in feats["keypoints"]}
    base, owner = {}, []
    for f in sample["frames"]:
        base[f["id"]] = len(owner)
        owner.extend([f["id"]] * blocks[f["id"]]["count"])
    parent = list(range(len(owner)))

    def find(i):
        while parent[i] != i:
            parent[i] = parent[parent[i]]
            i = parent[i]
        return i

    for pair in feats["pairs"]:
        for i in range(pair["count"]):
            a = find(base[pair["first"]] + int(matches[pair["offset"] + 2 * i]))
            b = find(base[pair["second"]] + int(matches[pair["offset"] + 2 * i + 1]))
            if a != b:
                parent[max(a, b)] = min(a, b)
    tracks = {}
    for i in range(len(owner)):
        tracks.setdefault(find(i), []).append(i)
    by_id = {v["id"]: v for v in views}
    landmarks = []
    for members in tracks.values():
        if len(members) < 2:
            continue
        owners = [owner[i] for i in members]
        if len(set(owners)) != len(owners):
            continue
        obs = []
        for j, i in enumerate(members):
            view = by_id[owners[j]]
            local = blocks[view["id"]]["offset"] + 2 * (i - base[view["id"]])
            obs.append({"viewId": view["id"], "featureId": i, "camera": view["camera"],
                        "pose": view["pose"],
                        "pixel": (float(keypoints[local]), float(keypoints[local + 1]))})
        track = triangulate_track(obs)
        if track is None:
            continue
        xyz, inliers = track
        kept = [o for j, o in enumerate(obs) if inliers[j]]
        if len(kept) >= 2:
            landmarks.append({"xyz": xyz, "observations": kept})
    landmarks.sort(key=lambda l: -len(l["observations"]))  # stable, as Array.prototype.sort
    landmarks = landmarks[:MAX_POINTS]
    R = np.array([v["pose"]["R"] for v in views])
    t = np.array([[x + (0.004 * wave(i, j) if i else 0.0) for j, x in enumerate(v["pose"]["t"])]
                  for i, v in enumerate(views)])
    xyz = np.array([[x + 0.004 * wave(i + 1000, j) for j, x in enumerate(l["xyz"])]
                    for i, l in enumerate(landmarks)])
    starts = np.cumsum([0] + [len(l["observations"]) for l in landmarks]).astype(np.int64)
    return {
        "view_id": np.array([v["id"] for v in views], dtype=np.int64),
        "camera": np.array([[v["camera"][k] for k in ("fx", "fy", "cx", "cy", "k1", "k2")]
                            for v in views]),
        "R": R, "t": t, "xyz": xyz, "obs_start": starts,
        "obs_view": np.array([o["viewId"] for l in landmarks for o in l["observations"]],
                             dtype=np.int64),
        "obs_feature": np.array([o["featureId"] for l in landmarks for o in l["observations"]],
                                dtype=np.int64),
        "obs_pixel": np.array([o["pixel"] for l in landmarks for o in l["observations"]],
                              dtype=np.float32),
    }


def main(argv):
    root = Path(argv[1]) if len(argv) > 1 else Path("/root/reference")
    arrays = build(root)
    np.savez_compressed(OUT, **arrays)
    print(f"{OUT.name}: {len(arrays['view_id'])} views, {len(arrays['xyz'])} points, "
          f"{len(arrays['obs_view'])} observations")


if __name__ == "__main__":
    main(sys.argv)
