// ba_math.h — scalar device math of the bundle adjustment: projection with Brown radial
// distortion and its derivative, robust cost, the equilibrated 3x3 / 6x6 Cholesky solves, the
// left SO(3) pose update and the shared focal exponential. Each function restates the reference
// helper it names (src/device/gpu/numerics.shade.ts; joint-focal.shade.ts:110-120) so that the
// device computes the same quantities as the reference's kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tcvba {

struct Model {
  double fx, fy, cx, cy, k1, k2;
};

// radialLimit, numerics.shade.ts:79-99: the squared radius where the radial map stops being
// injective (the first zero of d(r·D(r))/dr), or 1e30.
__host__ __device__ __forceinline__ double radial_limit(const Model& k) {
  if (k.k2 == 0) {
    if (k.k1 < 0) return -1 / (3 * k.k1);
    return 1e30;
  }
  const double a = 5 * k.k2, b = 3 * k.k1;
  const double d = b * b - 4 * a;
  if (d < 0) return 1e30;
  const double sign = b < 0 ? -1.0 : 1.0;
  const double q = -0.5 * (b + sign * sqrt(d));
  double limit = 1e30;
  const double first = q / a;
  if (first > 0) limit = fmin(limit, first);
  if (q != 0) {
    const double second = 1 / q;
    if (second > 0) limit = fmin(limit, second);
  }
  return limit;
}

// project, numerics.shade.ts:141-173: pixel and the rows d(pixel)/d(q) of camera-frame point q.
__host__ __device__ __forceinline__ bool project(const Model& k, const double q[3], double& px, double& py,
                                        double du[3], double dv[3]) {
  if (!(q[2] > 1e-12)) return false;
  const double x = q[0] / q[2], y = q[1] / q[2], r2 = x * x + y * y;
  if (!(r2 < radial_limit(k) && r2 < 1e8)) return false;
  const double d = 1 + k.k1 * r2 + k.k2 * r2 * r2, g = 2 * (k.k1 + 2 * k.k2 * r2);
  if (!(d > 0 && 1 + 3 * k.k1 * r2 + 5 * k.k2 * r2 * r2 > 0)) return false;
  const double xx = d + g * x * x, xy = g * x * y, yy = d + g * y * y;
  px = k.fx * x * d + k.cx;
  py = k.fy * y * d + k.cy;
  du[0] = (k.fx * xx) / q[2];
  du[1] = (k.fx * xy) / q[2];
  du[2] = (-k.fx * (xx * x + xy * y)) / q[2];
  dv[0] = (k.fy * xy) / q[2];
  dv[1] = (k.fy * yy) / q[2];
  dv[2] = (-k.fy * (xy * x + yy * y)) / q[2];
  const double mp = fmax(fabs(px), fabs(py));
  const double mu = fmax(fmax(fabs(du[0]), fabs(du[1])), fabs(du[2]));
  const double mv = fmax(fmax(fabs(dv[0]), fabs(dv[1])), fabs(dv[2]));
  return mp < 1e12 && mu < 1e12 && mv < 1e12;
}

// The residual-weighted second derivative of the projection: G = a0 d2u/dq2 + a1 d2v/dq2 (3x3,
// row-major, symmetric) at a camera-frame point q that `project` accepted. With x = q0/q2,
// y = q1/q2, r2 = x^2 + y^2, D = 1 + k1 r2 + k2 r2^2, D' = k1 + 2 k2 r2, D'' = 2 k2, the image-plane
// map m = (x D, y D) has
//   d2 m_x/dx2 = 6x D' + 4x^3 D'',  d2 m_x/dxdy = 2y D' + 4x^2 y D'',  d2 m_x/dy2 = 2x D' + 4x y^2 D''
// (m_y by exchanging x and y), and the chain through (x, y)(q) adds the curvature of x = q0/q2 and
// y = q1/q2: d2x/dq0dq2 = -1/q2^2, d2x/dq2^2 = 2x/q2^2 (likewise y).
__host__ __device__ __forceinline__ void projection_curvature(const Model& k, const double q[3], double a0,
                                                     double a1, double G[9]) {
  const double s = 1 / q[2], x = q[0] * s, y = q[1] * s, r2 = x * x + y * y;
  const double D = 1 + k.k1 * r2 + k.k2 * r2 * r2, D1 = k.k1 + 2 * k.k2 * r2, D2 = 2 * k.k2;
  const double a = a0 * k.fx, c = a1 * k.fy;
  // first derivatives of m in (x, y), weighted
  const double gx = a * (D + 2 * x * x * D1) + c * (2 * x * y * D1);
  const double gy = a * (2 * x * y * D1) + c * (D + 2 * y * y * D1);
  // second derivatives of m in (x, y), weighted
  const double mxx = a * (6 * x * D1 + 4 * x * x * x * D2) + c * (2 * y * D1 + 4 * x * x * y * D2);
  const double mxy = a * (2 * y * D1 + 4 * x * x * y * D2) + c * (2 * x * D1 + 4 * x * y * y * D2);
  const double myy = a * (2 * x * D1 + 4 * x * y * y * D2) + c * (6 * y * D1 + 4 * y * y * y * D2);
  // d(x, y)/dq = s [[1, 0, -x], [0, 1, -y]]
  const double Jx[3] = {s, 0, -x * s}, Jy[3] = {0, s, -y * s};
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      G[i * 3 + j] = Jx[i] * (mxx * Jx[j] + mxy * Jy[j]) + Jy[i] * (mxy * Jx[j] + myy * Jy[j]);
  // curvature of x(q), y(q)
  const double s2 = s * s;
  G[0 * 3 + 2] += -gx * s2;
  G[2 * 3 + 0] += -gx * s2;
  G[1 * 3 + 2] += -gy * s2;
  G[2 * 3 + 1] += -gy * s2;
  G[2 * 3 + 2] += 2 * (gx * x + gy * y) * s2;
}

// robustCost, numerics.shade.ts:184-187: Huber cost of a pixel error.
__host__ __device__ __forceinline__ double robust_cost(double error, double huber) {
  if (error <= huber) return error * error;
  return 2 * huber * error - huber * huber;
}

// solve3, numerics.shade.ts:25-45: diagonal equilibration, then a 3x3 Cholesky with pivots > 1e-14.
__device__ __forceinline__ bool solve3(const double a[9], const double b[3], double x[3]) {
  if (!(fmin(a[0], fmin(a[4], a[8])) > 0)) return false;
  const double dx = 1 / sqrt(a[0]), dy = 1 / sqrt(a[4]), dz = 1 / sqrt(a[8]);
  const double l10 = a[3] * dx * dy, l20 = a[6] * dx * dz;
  const double pivot1 = 1 - l10 * l10;
  if (!(pivot1 > 1e-14)) return false;
  const double l11 = sqrt(pivot1), l21 = (a[7] * dy * dz - l20 * l10) / l11;
  const double pivot2 = 1 - l20 * l20 - l21 * l21;
  if (!(pivot2 > 1e-14)) return false;
  const double l22 = sqrt(pivot2), y0 = b[0] * dx;
  const double y1 = (b[1] * dy - l10 * y0) / l11, y2 = (b[2] * dz - l20 * y0 - l21 * y1) / l22;
  const double x2 = y2 / l22, x1 = (y1 - l21 * x2) / l11, x0 = y0 - l10 * x1 - l20 * x2;
  x[0] = x0 * dx;
  x[1] = x1 * dy;
  x[2] = x2 * dz;
  return true;
}

// The inverse of a symmetric 3x3 block by three solve3 columns, symmetrised
// (bundle.shade.ts:403-420). Returns false when a column fails.
__device__ __forceinline__ bool inverse3(const double V[9], double inv[9]) {
  for (int column = 0; column < 3; column++) {
    double e[3] = {0, 0, 0}, x[3];
    e[column] = 1;
    if (!solve3(V, e, x)) return false;
    inv[column] = x[0];
    inv[3 + column] = x[1];
    inv[6 + column] = x[2];
  }
  for (int j = 0; j < 3; j++)
    for (int k = j + 1; k < 3; k++) {
      const double v = (inv[j * 3 + k] + inv[k * 3 + j]) * 0.5;
      inv[j * 3 + k] = v;
      inv[k * 3 + j] = v;
    }
  return true;
}

// True when a symmetric positive semi-definite 3x3 block (row-major) is singular to rounding: its
// smallest eigenvalue e0 below `tau` times its largest e2. With tr the trace, m2 the sum of the
// principal 2x2 minors and e0 <= e1 <= e2: e1 e2 <= m2 <= 3 e1 e2 and e2 <= tr <= 3 e2, so
// det / (m2 tr) lies in [e0 / (9 e2), e0 / e2] and m2 / tr^2 in [e1 / (9 e2), 3 e1 / e2]; the
// second test catches two weak directions, where det is below its own rounding.
__device__ __forceinline__ bool singular_to_rounding3(const double A[9], double tau) {
  const double tr = A[0] + A[4] + A[8];
  const double m00 = A[4] * A[8] - A[5] * A[7], m11 = A[0] * A[8] - A[2] * A[6],
               m22 = A[0] * A[4] - A[1] * A[3];
  const double m2 = m00 + m11 + m22;
  const double det = A[0] * m00 - A[1] * (A[3] * A[8] - A[5] * A[6]) + A[2] * (A[3] * A[7] - A[4] * A[6]);
  return tr > 0 && (fabs(m2) <= tau * tr * tr || fabs(det) <= tau * fabs(m2) * tr);
}

// Lower Cholesky factor of a symmetric positive definite 3x3 block (row-major V): V = L L^T with
// L packed as l00, l10, l11, l20, l21, l22. The direct route splits V_p^-1 = L L^T so that each
// observation's cross block becomes one factor G = W L of the product W V^-1 W^T = G G^T.
__device__ __forceinline__ void lower_cholesky3(const double V[9], double L[6]) {
  if (!(V[0] > 0)) {
    // a held point's zero block (see singular_to_rounding3): a zero factor
    for (int k = 0; k < 6; k++) L[k] = 0;
    return;
  }
  L[0] = sqrt(V[0]);
  L[1] = V[3] / L[0];
  L[2] = sqrt(V[4] - L[1] * L[1]);
  L[3] = V[6] / L[0];
  L[4] = (V[7] - L[3] * L[1]) / L[2];
  L[5] = sqrt(V[8] - L[3] * L[3] - L[4] * L[4]);
}

// The equilibrated Cholesky factor of numerics.shade.ts:221-253 (solve6), computed once.
struct Chol6 {
  double L[21];  // packed lower triangle, row i holds L[i][0..i]
  double scale[6];
};

__device__ __forceinline__ int tri(int i, int j) { return i * (i + 1) / 2 + j; }

__device__ __forceinline__ bool cholesky6(const double m[36], Chol6& f) {
  for (int i = 0; i < 6; i++) {
    if (!(m[i * 6 + i] > 0)) return false;
    f.scale[i] = 1 / sqrt(m[i * 6 + i]);
  }
  for (int i = 0; i < 6; i++)
    for (int j = 0; j <= i; j++) {
      double value = m[i * 6 + j] * f.scale[i] * f.scale[j];
      for (int k = 0; k < j; k++) value -= f.L[tri(i, k)] * f.L[tri(j, k)];
      if (i == j) {
        if (!(value > 1e-14)) return false;
        f.L[tri(i, j)] = sqrt(value);
      } else {
        f.L[tri(i, j)] = value / f.L[tri(j, j)];
      }
    }
  return true;
}

// solve6's forward and back substitution on a factor from cholesky6.
__device__ __forceinline__ void chol6_solve(const Chol6& f, const double rhs[6], double values[6]) {
  for (int i = 0; i < 6; i++) {
    double v = rhs[i] * f.scale[i];
    for (int j = 0; j < i; j++) v -= f.L[tri(i, j)] * values[j];
    values[i] = v / f.L[tri(i, i)];
  }
  for (int index = 6; index > 0; index--) {
    const int i = index - 1;
    double v = values[i];
    for (int j = i + 1; j < 6; j++) v -= f.L[tri(j, i)] * values[j];
    values[i] = v / f.L[tri(i, i)];
  }
  for (int i = 0; i < 6; i++) values[i] *= f.scale[i];
}

// movePose, numerics.shade.ts:254-281: R' = exp([w]x) R, t' = exp([w]x) t + v.
__host__ __device__ __forceinline__ void move_pose(const double p[12], const double delta[6], double out[12]) {
  const double x = delta[0], y = delta[1], z = delta[2];
  const double theta = sqrt(x * x + y * y + z * z);
  double a = 1 - (theta * theta) / 6, b = 0.5 - (theta * theta) / 24;
  if (theta >= 1e-7) {
    a = sin(theta) / theta;
    b = (1 - cos(theta)) / (theta * theta);
  }
  const double W[9] = {0, -z, y, z, 0, -x, -y, x, 0};
  double dR[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      dR[i * 3 + j] += a * W[i * 3 + j];
      for (int k = 0; k < 3; k++) dR[i * 3 + j] += b * W[i * 3 + k] * W[k * 3 + j];
    }
  for (int i = 0; i < 12; i++) out[i] = 0;
  for (int i = 0; i < 3; i++) {
    out[9 + i] = delta[3 + i];
    for (int k = 0; k < 3; k++) {
      out[9 + i] += dR[i * 3 + k] * p[9 + k];
      for (int j = 0; j < 3; j++) out[i * 3 + j] += dR[i * 3 + k] * p[k * 3 + j];
    }
  }
}

// focalScale, joint-focal.shade.ts:110-120: exp(x) by halving |x| below 0.25, a 20-term Taylor
// series, and repeated squaring.
__host__ __device__ __forceinline__ double focal_scale(double log_scale) {
  double x = log_scale;
  unsigned squarings = 0;
  for (int j = 0; j < 12; j++)
    if (fabs(x) > 0.25) {
      x *= 0.5;
      squarings++;
    }
  double value = 1, term = 1;
  for (int j = 1; j <= 20; j++) {
    term *= x / (double)(float)j;
    value += term;
  }
  for (unsigned j = 0; j < squarings; j++) value *= value;
  return value;
}

}  // namespace tcvba
