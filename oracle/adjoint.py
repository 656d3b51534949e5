"""TEST INFRASTRUCTURE: dense references for the derivative entries, on tiny problems.

Only tests import this module. Everything here is built on the oracle's chart gradient
(`tcv_oracle_gradient`, the forward's linearizeBundle Jacobians and Huber weights) and its movePose:
  * `move` applies a chart step (movePose per variable camera, X + dX, log f + df);
  * `fd_hessian` is the central-difference Jacobian of that gradient, column by column: the exact
    Hessian of the half robust cost, independent of the engine's analytic second derivatives;
  * `newton` runs dense Newton steps on it (a stationary point to rounding, for small problems
    with no gauge freedom);
  * `dense_vjp` is implicit differentiation with the dense Hessian: lambda = H^-1 v on the
    variable rows, then D_i J_i lambda per observation, for a cotangent given in the chart;
  * `dense_focal_vjp` is its input-focal part, -lambda . dg/d(eps_c) for eps_c the log of camera
    c's input fx and fy, with dg/d(eps_c) by central differences of the same gradient (independent
    of the engine's closed form in D_i J_i lambda, Jf_i and w_i r_i).
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import oracle as O


def sizes(control):
    C, P = control.cameras, control.points
    sh = 1 if control.shared_focal else 0
    return C, P, sh, 6 * C + 3 * P + sh


def variable_rows(control, index) -> np.ndarray:
    """Chart rows that move: the variable cameras' 6, every point's 3, the focal row."""
    C, P, sh, n = sizes(control)
    M = control.observations
    active = np.asarray(index[3 * M + P + C + 2: 3 * M + P + 2 * C + 2]) != 0
    keep = np.ones(n, dtype=bool)
    for c in range(C):
        if not active[c]:
            keep[6 * c: 6 * c + 6] = False
    return np.nonzero(keep)[0]


def move(control, scene: np.ndarray, log_focal: float, delta: np.ndarray):
    """The scene and log focal scale after the chart step `delta` (6C + 3P + shared)."""
    C, P, sh, n = sizes(control)
    out = scene.copy()
    for c in range(C):
        d = delta[6 * c: 6 * c + 6]
        if np.any(d != 0):
            out[12 * c: 12 * c + 12] = O.move_pose(scene[12 * c: 12 * c + 12], d)
    out[12 * C: 12 * C + 3 * P] += delta[6 * C: 6 * C + 3 * P]
    return out, log_focal + (delta[-1] if sh else 0.0)


def fd_hessian(control, scene, index, log_focal: float = 0.0, h: float = 1e-6) -> np.ndarray:
    """The dense Hessian of the half robust cost by central differences of the chart gradient,
    over the variable rows (`variable_rows`), symmetrised."""
    C, P, sh, n = sizes(control)
    rows = variable_rows(control, index)
    H = np.empty((len(rows), len(rows)))
    for j, k in enumerate(rows):
        e = np.zeros(n)
        e[k] = h
        sp, fp = move(control, scene, log_focal, e)
        sm, fm = move(control, scene, log_focal, -e)
        gp, _, _ = O.gradient(control, sp, index, fp)
        gm, _, _ = O.gradient(control, sm, index, fm)
        H[:, j] = ((gp - gm) / (2 * h))[rows]
    return 0.5 * (H + H.T)


def newton(control, scene, index, log_focal: float = 0.0, steps: int = 4):
    """Dense Newton steps on the FD Hessian from (scene, log focal); for a problem without gauge
    freedom (two or more fixed cameras, no directional points)."""
    C, P, sh, n = sizes(control)
    rows = variable_rows(control, index)
    for _ in range(steps):
        g, _, _ = O.gradient(control, scene, index, log_focal)
        H = fd_hessian(control, scene, index, log_focal)
        d = np.zeros(n)
        d[rows] = -np.linalg.solve(H, g[rows])
        scene, log_focal = move(control, scene, log_focal, d)
    return scene, log_focal


def dense_vjp(control, packed, scene, log_focal: float, chart_cotangent: np.ndarray,
              huber: float):
    """Pixel cotangent (2M) of implicit differentiation with the dense FD Hessian, for a cotangent
    given in the chart (6C + 3P + shared), and the adjoint lambda (chart layout)."""
    C, P, sh, n = sizes(control)
    M = control.observations
    index = packed.index
    rows = variable_rows(control, index)
    H = fd_hessian(control, scene, index, log_focal)
    lam = np.zeros(n)
    lam[rows] = np.linalg.solve(H, chart_cotangent[rows])
    probe = dataclasses.replace(packed, scene=scene)
    _, _, _, _, blocks = O.schur_probe(control, probe, 0.0, log_focal,
                                       np.zeros(6 * C + sh), blocks=True)
    J, focal = blocks["J"], blocks["focal"]
    cams = np.asarray(index[:M], dtype=np.int64)
    pts = np.asarray(index[M: 2 * M], dtype=np.int64)
    out = np.zeros((M, 2))
    for i in range(M):
        b = J[i]
        if b[23] != 0:
            continue
        lc = lam[6 * cams[i]: 6 * cams[i] + 6]
        lp = lam[6 * C + 3 * pts[i]: 6 * C + 3 * pts[i] + 3]
        t = np.array([b[0:6] @ lc + b[12:15] @ lp, b[6:12] @ lc + b[15:18] @ lp])
        if sh:
            t += focal[6 * C + 6 + 2 * i: 6 * C + 8 + 2 * i] * lam[-1]
        r = b[18:20]
        e = np.hypot(r[0], r[1])
        if e > huber:
            rh = r / e
            t = b[20] * (t - rh * (rh @ t))
        out[i] = t
    return out.reshape(-1), lam


def dense_focal_vjp(control, scene, index, log_focal: float, lam: np.ndarray,
                    h: float = 1e-6) -> np.ndarray:
    """Input-focal cotangent (C) of implicit differentiation for the adjoint `lam` (chart layout,
    lam = H^-1 v with H the central-difference Jacobian of the chart gradient g, as `dense_vjp`):
    d<v, theta*>/d eps_c = -lam . dg/d eps_c, eps_c scaling camera c's input fx and fy by e^eps."""
    C, P, sh, n = sizes(control)
    rows = variable_rows(control, index)
    base = 12 * C + 3 * P
    out = np.zeros(C)
    for c in range(C):
        sp, sm = scene.copy(), scene.copy()
        sp[base + 6 * c: base + 6 * c + 2] *= np.exp(h)
        sm[base + 6 * c: base + 6 * c + 2] *= np.exp(-h)
        gp, _, _ = O.gradient(control, sp, index, log_focal)
        gm, _, _ = O.gradient(control, sm, index, log_focal)
        out[c] = -float(lam[rows] @ ((gp - gm) / (2 * h))[rows])
    return out
