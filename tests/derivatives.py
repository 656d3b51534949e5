"""Host-side derivative checks at full size, built on the engine's per-observation linearisation
(`tcv_ba_jacobian`: r, d(u,v)/d(omega, v), d(u,v)/dX, d(u,v)/d log f, w), which
`jacobian_vs_differences` checks against central differences of the oracle's projection.

* `gn_product`: the Gauss-Newton Hessian sum_i J_i^T D_i J_i applied to a chart vector, with
  D_i = d(w r)/dr (the Huber curvature: w (I - h h^T) outside the radius, h = r / |r|);
* `pixel_cotangent`, `focal_cotangent`: D_i J_i lambda and the per-camera input-focal sums
  -sum_i (D_i J_i lambda) . Jf_i [- (J_i lambda) . (w_i r_i)], from the adjoint lambda;
* `point_cotangent`: a cotangent on the finite points only, orthogonal to the scale gauge (the
  similarity about the anchor centre), so the gauge map's adjoint leaves it unchanged and the
  adjoint system's right-hand side is the cotangent itself.
"""
from __future__ import annotations

import numpy as np

from oracle import oracle as O


def parts(J: np.ndarray):
    """(r, Jc, Jp, Jf, w, ok) of a (M, 23) linearisation."""
    r = J[:, 0:2]
    Jc = J[:, 2:14].reshape(-1, 2, 6)
    Jp = J[:, 14:20].reshape(-1, 2, 3)
    Jf = J[:, 20:22]
    w = J[:, 22]
    return r, Jc, Jp, Jf, w, w >= 0


def robust(r: np.ndarray, w: np.ndarray, huber: float, t: np.ndarray) -> np.ndarray:
    """D t per observation, D = d(w r)/dr."""
    e = np.hypot(r[:, 0], r[:, 1])
    out = t.copy()
    far = e > huber
    h = r[far] / e[far, None]
    d = (h * t[far]).sum(1)
    out[far] = w[far, None] * (t[far] - h * d[:, None])
    return out


def apply_j(control, index, J, lam) -> np.ndarray:
    """J_i lambda per observation (M, 2)."""
    C, P, M = control.cameras, control.points, control.observations
    _, Jc, Jp, Jf, _, ok = parts(J)
    cams = np.asarray(index[:M], dtype=np.int64)
    pts = np.asarray(index[M: 2 * M], dtype=np.int64)
    lc = lam[: 6 * C].reshape(C, 6)[cams]
    lp = lam[6 * C: 6 * C + 3 * P].reshape(P, 3)[pts]
    t = np.einsum("mij,mj->mi", Jc, lc) + np.einsum("mij,mj->mi", Jp, lp)
    if control.shared_focal:
        t += Jf * lam[-1]
    t[~ok] = 0
    return t


def gn_product(control, index, J, lam) -> np.ndarray:
    """sum_i J_i^T D_i J_i lambda in the chart layout (6C + 3P [+ 1])."""
    C, P, M = control.cameras, control.points, control.observations
    r, Jc, Jp, Jf, w, ok = parts(J)
    t = robust(r, w, control.huber, apply_j(control, index, J, lam))
    t[~ok] = 0
    cams = np.asarray(index[:M], dtype=np.int64)
    pts = np.asarray(index[M: 2 * M], dtype=np.int64)
    out = np.zeros_like(lam)
    cam = np.zeros((C, 6))
    np.add.at(cam, cams, np.einsum("mij,mi->mj", Jc, t))
    pt = np.zeros((P, 3))
    np.add.at(pt, pts, np.einsum("mij,mi->mj", Jp, t))
    out[: 6 * C] = cam.reshape(-1)
    out[6 * C: 6 * C + 3 * P] = pt.reshape(-1)
    if control.shared_focal:
        out[-1] = float((Jf * t).sum())
    return out


def pixel_cotangent(control, index, J, lam) -> np.ndarray:
    r, _, _, _, w, ok = parts(J)
    t = robust(r, w, control.huber, apply_j(control, index, J, lam))
    t[~ok] = 0
    return t


def focal_cotangent(control, index, J, lam, exact: bool) -> np.ndarray:
    C, M = control.cameras, control.observations
    r, _, _, Jf, w, ok = parts(J)
    j = apply_j(control, index, J, lam)
    t = robust(r, w, control.huber, j)
    share = -(t * Jf).sum(1)
    if exact:
        share -= (j * (w[:, None] * r)).sum(1)
    share[~ok] = 0
    out = np.zeros(C)
    np.add.at(out, np.asarray(index[:M], dtype=np.int64), share)
    return out


def camera_center(scene: np.ndarray, c: int) -> np.ndarray:
    R = scene[12 * c: 12 * c + 9].reshape(3, 3)
    return -R.T @ scene[12 * c + 9: 12 * c + 12]


def point_cotangent(control, scene: np.ndarray, rng) -> tuple[np.ndarray, np.ndarray]:
    """(output cotangent 12C + 3P [+ 1], chart cotangent 6C + 3P [+ 1]): N(0, 1) on the finite
    points with its component along X_p - o removed (o the anchor's centre; with the scale gauge
    off nothing is removed), zero elsewhere."""
    C, P = control.cameras, control.points
    sh = 1 if control.shared_focal else 0
    finite = P - control.directional_points
    X = scene[12 * C: 12 * C + 3 * finite].reshape(finite, 3)
    v = rng.normal(size=(finite, 3))
    if control.normalize_scale:
        n = X - camera_center(scene, control.anchor)
        v -= n * ((v * n).sum() / (n * n).sum())
    out = np.zeros(12 * C + 3 * P + sh)
    out[12 * C: 12 * C + 3 * finite] = v.reshape(-1)
    chart = np.zeros(6 * C + 3 * P + sh)
    chart[6 * C: 6 * C + 3 * finite] = v.reshape(-1)
    return out, chart


def jacobian_vs_differences(control, at, index, J, log_focal: float, sample, h: float = 1e-6):
    """The largest |J - central differences| / max(1, |J|) over the sampled observations' 2 x 10
    blocks (pose chart, point, log focal), the differences taken of the oracle's projection with
    the chart's movePose, step h (the reference's geometry tests' step). On the fixed route the
    focal column is d(u,v)/d log f of the camera's own focal (log_focal 0)."""
    C, P, M = control.cameras, control.points, control.observations
    cams = np.asarray(index[:M], dtype=np.int64)
    pts = np.asarray(index[M: 2 * M], dtype=np.int64)
    models = at.scene[12 * C + 3 * P: 18 * C + 3 * P].reshape(C, 6)
    active = np.asarray(index[3 * M + P + C + 2: 3 * M + P + 2 * C + 2]).astype(bool)

    def pixel(pose, X, model, lfs):
        s = O.focal_scale(lfs)
        k = model.copy()
        k[0] *= s
        k[1] *= s
        R = pose[:9].reshape(3, 3)
        ok, px, _, _ = O.project(k, R @ X + pose[9:])
        return np.asarray(px)

    worst = 0.0
    for i in sample:
        c, p = cams[i], pts[i]
        pose = at.scene[12 * c: 12 * c + 12]
        X = at.scene[12 * C + 3 * p: 12 * C + 3 * p + 3]
        cols = []
        for k in range(6):
            if not active[c]:
                cols.append(np.zeros(2))
                continue
            d = np.zeros(6)
            d[k] = h
            cols.append((pixel(O.move_pose(pose, d), X, models[c], log_focal)
                         - pixel(O.move_pose(pose, -d), X, models[c], log_focal)) / (2 * h))
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            cols.append((pixel(pose, X + e, models[c], log_focal)
                         - pixel(pose, X - e, models[c], log_focal)) / (2 * h))
        cols.append((pixel(pose, X, models[c], log_focal + h)
                     - pixel(pose, X, models[c], log_focal - h)) / (2 * h))
        fd = np.array(cols).T
        an = np.concatenate([J[i, 2:14].reshape(2, 6), J[i, 14:20].reshape(2, 3),
                             J[i, 20:22].reshape(2, 1)], axis=1)
        worst = max(worst, float(np.max(np.abs(fd - an)) / max(1.0, np.max(np.abs(an)))))
    return worst
