/*
 * ba_oracle.c — TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * CPU restatement, in IEEE binary64, of the reference's bundle-adjustment operator
 * `bundleAdjustGpu` as its CPU tier executes it (resident run, one window, serial kernels):
 *
 *   src/sfm/bundle/gpu/bundle.shade.ts       fixed-intrinsics route (`bundleAdjustGpuFixed`)
 *   src/sfm/bundle/gpu/joint-focal.shade.ts  shared log-focal route (`bundleAdjustJointFocalGpu`)
 *   src/device/gpu/numerics.shade.ts         project / poseRow / robustCost / solve3 / solve6 /
 *                                            movePose / radialLimit
 *   src/sfm/bundle/gpu/bundle.ts:314-772, joint-focal.ts:340-900  the order of the steps
 *
 * Only tests, the smoke check and the CPU-baseline leg of the benchmark may load it. It is the
 * checker the HIP engine is compared with, never a fallback for it.
 *
 * The arithmetic follows the reference expression by expression (left-to-right products, the same
 * accumulation orders, the CPU-tier `…Oracle` cost partials and `bundleCameraBlocks` camera order).
 * Build with -ffp-contract=off so that no product is fused, as in JavaScript. OpenMP spreads the
 * independent per-point / per-camera / per-block loops over threads without changing any sum order.
 *
 * Pins (tests/test_oracle_*.py): projection and pose/point Jacobians against central differences,
 * solve3/solve6 against numpy, the Schur product against a dense Hessian built from the oracle's own
 * Jacobians, PCG against a dense solve, the LM loop against the reference tests' invariants and the
 * independent dense reduced-camera solver `tcv_oracle_dense_solve` (bundle-adjustment.ts).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../include/typecv_ba.h"

#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
  double fx, fy, cx, cy, k1, k2;
} Model;

typedef struct {
  double px, py;
  double du[3], dv[3];
  int ok;
} Projection;

/* ---------------------------------------------------------------- numerics.shade.ts */

/* radialLimit, numerics.shade.ts:79-99 */
static double radial_limit(const Model* k) {
  if (k->k2 == 0) {
    if (k->k1 < 0) return -1 / (3 * k->k1);
    return 1e30;
  }
  double a = 5 * k->k2, b = 3 * k->k1;
  double d = b * b - 4 * a;
  if (d < 0) return 1e30;
  double sign = 1;
  if (b < 0) sign = -1;
  double q = -0.5 * (b + sign * sqrt(d));
  double limit = 1e30;
  double first = q / a;
  if (first > 0) limit = fmin(limit, first);
  if (q != 0) {
    double second = 1 / q;
    if (second > 0) limit = fmin(limit, second);
  }
  return limit;
}

/* project, numerics.shade.ts:141-173 */
static Projection project(const Model* k, const double q[3]) {
  Projection bad;
  memset(&bad, 0, sizeof bad);
  if (!(q[2] > 1e-12)) return bad;
  double x = q[0] / q[2], y = q[1] / q[2], r2 = x * x + y * y;
  if (!(r2 < radial_limit(k) && r2 < 1e8)) return bad;
  double d = 1 + k->k1 * r2 + k->k2 * r2 * r2, g = 2 * (k->k1 + 2 * k->k2 * r2);
  if (!(d > 0 && 1 + 3 * k->k1 * r2 + 5 * k->k2 * r2 * r2 > 0)) return bad;
  double xx = d + g * x * x, xy = g * x * y, yy = d + g * y * y;
  Projection out;
  out.px = k->fx * x * d + k->cx;
  out.py = k->fy * y * d + k->cy;
  out.du[0] = (k->fx * xx) / q[2];
  out.du[1] = (k->fx * xy) / q[2];
  out.du[2] = (-k->fx * (xx * x + xy * y)) / q[2];
  out.dv[0] = (k->fy * xy) / q[2];
  out.dv[1] = (k->fy * yy) / q[2];
  out.dv[2] = (-k->fy * (xy * x + yy * y)) / q[2];
  out.ok = 1;
  double mp = fmax(fabs(out.px), fabs(out.py));
  double mu = fmax(fmax(fabs(out.du[0]), fabs(out.du[1])), fabs(out.du[2]));
  double mv = fmax(fmax(fabs(out.dv[0]), fabs(out.dv[1])), fabs(out.dv[2]));
  if (!(mp < 1e12 && mu < 1e12 && mv < 1e12)) return bad;
  return out;
}

/* poseRow, numerics.shade.ts:181-183 */
static void pose_row(const double q[3], const double j[3], double out[6]) {
  out[0] = j[2] * q[1] - j[1] * q[2];
  out[1] = j[0] * q[2] - j[2] * q[0];
  out[2] = j[1] * q[0] - j[0] * q[1];
  out[3] = j[0];
  out[4] = j[1];
  out[5] = j[2];
}

/* robustCost, numerics.shade.ts:184-187 */
static double robust_cost(double error, double huber) {
  if (error <= huber) return error * error;
  return 2 * huber * error - huber * huber;
}

/* solve3, numerics.shade.ts:25-45: equilibrated 3x3 Cholesky. */
static int solve3(const double a[9], const double b[3], double out[3]) {
  out[0] = out[1] = out[2] = 0;
  if (!(fmin(a[0], fmin(a[4], a[8])) > 0)) return 0;
  double dx = 1 / sqrt(a[0]), dy = 1 / sqrt(a[4]), dz = 1 / sqrt(a[8]);
  double l10 = a[3] * dx * dy, l20 = a[6] * dx * dz;
  double pivot1 = 1 - l10 * l10;
  if (!(pivot1 > 1e-14)) return 0;
  double l11 = sqrt(pivot1), l21 = (a[7] * dy * dz - l20 * l10) / l11;
  double pivot2 = 1 - l20 * l20 - l21 * l21;
  if (!(pivot2 > 1e-14)) return 0;
  double l22 = sqrt(pivot2), y0 = b[0] * dx;
  double y1 = (b[1] * dy - l10 * y0) / l11, y2 = (b[2] * dz - l20 * y0 - l21 * y1) / l22;
  double x2 = y2 / l22, x1 = (y1 - l21 * x2) / l11, x0 = y0 - l10 * x1 - l20 * x2;
  out[0] = x0 * dx;
  out[1] = x1 * dy;
  out[2] = x2 * dz;
  return 1;
}

/* solve6, numerics.shade.ts:221-253: equilibrated 6x6 Cholesky, pivot > 1e-14. */
static int solve6(const double m[36], const double rhs[6], double values[6]) {
  double L[36], scale[6];
  memset(L, 0, sizeof L);
  for (int i = 0; i < 6; i++) values[i] = 0;
  for (int i = 0; i < 6; i++) {
    if (!(m[i * 6 + i] > 0)) return 0;
    scale[i] = 1 / sqrt(m[i * 6 + i]);
  }
  for (int i = 0; i < 6; i++) {
    for (int j = 0; j <= i; j++) {
      double value = m[i * 6 + j] * scale[i] * scale[j];
      for (int k = 0; k < j; k++) value -= L[i * 6 + k] * L[j * 6 + k];
      if (i == j) {
        if (!(value > 1e-14)) {
          for (int r = 0; r < 6; r++) values[r] = 0;
          return 0;
        }
        L[i * 6 + j] = sqrt(value);
      } else
        L[i * 6 + j] = value / L[j * 6 + j];
    }
    double v = rhs[i] * scale[i];
    for (int j = 0; j < i; j++) v -= L[i * 6 + j] * values[j];
    values[i] = v / L[i * 6 + i];
  }
  for (int index = 6; index > 0; index--) {
    int i = index - 1;
    double v = values[i];
    for (int j = i + 1; j < 6; j++) v -= L[j * 6 + i] * values[j];
    values[i] = v / L[i * 6 + i];
  }
  for (int i = 0; i < 6; i++) values[i] *= scale[i];
  return 1;
}

/* movePose, numerics.shade.ts:254-281: left SO(3) exponential, additive translation. */
static void move_pose(const double p[12], const double delta[6], double out[12]) {
  double x = delta[0], y = delta[1], z = delta[2];
  double theta = sqrt(x * x + y * y + z * z);
  double a = 1 - (theta * theta) / 6, b = 0.5 - (theta * theta) / 24;
  if (theta >= 1e-7) {
    a = sin(theta) / theta;
    b = (1 - cos(theta)) / (theta * theta);
  }
  double W[9] = {0, -z, y, z, 0, -x, -y, x, 0};
  double dR[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  for (int i = 0; i < 12; i++) out[i] = 0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      dR[i * 3 + j] += a * W[i * 3 + j];
      for (int k = 0; k < 3; k++) dR[i * 3 + j] += b * W[i * 3 + k] * W[k * 3 + j];
    }
  for (int i = 0; i < 3; i++) {
    out[9 + i] = delta[3 + i];
    for (int k = 0; k < 3; k++) {
      out[9 + i] += dR[i * 3 + k] * p[9 + k];
      for (int j = 0; j < 3; j++) out[i * 3 + j] += dR[i * 3 + k] * p[k * 3 + j];
    }
  }
}

/* focalScale, joint-focal.shade.ts:110-120: range-reduced 20-term Taylor series of exp. */
static double focal_scale(double log_scale) {
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

/* ---------------------------------------------------------------- the solver state */

typedef struct {
  tcv_ba_control ctl;
  uint32_t C, P, M, m, K, shared;
  double* scene; /* working copy: [poses 12C][points 3P][models 6C][pixels 2M] */
  double* trial; /* [poses 12C][points 3P] */
  const uint32_t* index;
  double* J;      /* obsRow: 24 per observation */
  double* cross;  /* K = 18 (fixed) or 21 (shared focal) per observation */
  double* pblk;   /* 13 per point: V^-1 (9), RHS (3), valid */
  double* cam;    /* damped camera blocks, 36 per camera */
  double* inv;    /* preconditioner blocks, 36 per camera */
  double* rhs;    /* reduced RHS, 6 per camera */
  double* focal;  /* shared route: H_cf 6C, 6 scalars, Jf 2 per observation */
  double* vec;    /* x r p Sp z (each m), y 3P */
  double* partials;
  double state[27];
} Oracle;

static int threads_used = 1;

#define OBS_CAMERA(o, i) ((o)->index[(i)])
#define OBS_POINT(o, i) ((o)->index[(o)->M + (i)])
#define POINT_START(o, p) ((o)->index[2 * (o)->M + (p)])
#define CAMERA_START(o, c) ((o)->index[2 * (o)->M + (o)->P + 1 + (c)])
#define CAMERA_OBS(o, n) ((o)->index[2 * (o)->M + (o)->P + (o)->C + 2 + (n)])
#define VARIABLE(o, c) ((o)->index[3 * (o)->M + (o)->P + (o)->C + 2 + (c)] != 0)

static int directional(const Oracle* o, uint32_t p) {
  return p >= o->P - o->ctl.directional_points;
}

static const double* pose_at(const Oracle* o, uint32_t c, int candidate) {
  return candidate ? o->trial + (size_t)c * 12 : o->scene + (size_t)c * 12;
}
static const double* point_at(const Oracle* o, uint32_t p, int candidate) {
  size_t at = (size_t)o->C * 12 + (size_t)p * 3;
  return candidate ? o->trial + at : o->scene + at;
}
static void observed_point(const Oracle* o, const double* pose, const double* X, uint32_t p,
                           double q[3]) {
  for (int r = 0; r < 3; r++) {
    q[r] = pose[r * 3] * X[0] + pose[r * 3 + 1] * X[1] + pose[r * 3 + 2] * X[2];
    if (!directional(o, p)) q[r] += pose[9 + r];
  }
}
/* cameraAt: bundle.shade.ts:112-122; joint-focal.shade.ts:121-134 scales fx, fy. */
static Model camera_at(const Oracle* o, uint32_t c, int candidate) {
  const double* s = o->scene + (size_t)o->C * 12 + (size_t)o->P * 3 + (size_t)c * 6;
  Model k = {s[0], s[1], s[2], s[3], s[4], s[5]};
  if (o->shared) {
    double log_scale = o->focal[o->C * 6 + (candidate ? 4 : 3)];
    double scale = focal_scale(log_scale);
    k.fx = s[0] * scale;
    k.fy = s[1] * scale;
  }
  return k;
}
static void camera_center(const Oracle* o, uint32_t c, int candidate, double out[3]) {
  const double* p = pose_at(o, c, candidate);
  out[0] = -p[0] * p[9] - p[3] * p[10] - p[6] * p[11];
  out[1] = -p[1] * p[9] - p[4] * p[10] - p[7] * p[11];
  out[2] = -p[2] * p[9] - p[5] * p[10] - p[8] * p[11];
}
static const double* pixel_at(const Oracle* o, uint32_t i) {
  return o->scene + (size_t)o->C * 18 + (size_t)o->P * 3 + (size_t)i * 2;
}
static double length3(const double v[3]) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

/* initializeBundle, bundle.shade.ts:235-250 / joint-focal.shade.ts:269-286 */
static void initialize_bundle(Oracle* o) {
  for (int j = 0; j < 26; j++) o->state[j] = 0;
  o->state[0] = 1e-3;
  if (o->shared) {
    o->focal[o->C * 6 + 3] = 0;
    o->focal[o->C * 6 + 4] = 0;
  }
  o->state[1] = 1;
  double origin[3], scale[3], d[3];
  camera_center(o, o->ctl.anchor, 0, origin);
  camera_center(o, o->ctl.scale_view, 0, scale);
  for (int r = 0; r < 3; r++) d[r] = origin[r] - scale[r];
  double baseline = length3(d);
  o->state[2] = baseline;
  o->state[3] = origin[0];
  o->state[4] = origin[1];
  o->state[5] = origin[2];
  if (!(baseline > 1e-10)) {
    o->state[14] = 1;
    o->state[1] = 0;
  }
}

/* metricLane, bundle.shade.ts:301-314 */
static void metric_lane(const Oracle* o, uint32_t i, double out[3]) {
  out[0] = out[1] = out[2] = 0;
  if (i >= o->M) return;
  uint32_t c = OBS_CAMERA(o, i), p = OBS_POINT(o, i);
  double q[3];
  observed_point(o, pose_at(o, c, 0), point_at(o, p, 0), p, q);
  Model k = camera_at(o, c, 0);
  Projection pr = project(&k, q);
  if (!pr.ok) {
    out[2] = 1;
    return;
  }
  const double* px = pixel_at(o, i);
  double ru = pr.px - px[0], rv = pr.py - px[1], squared = ru * ru + rv * rv;
  double error = sqrt(squared);
  out[0] = robust_cost(error, o->ctl.huber);
  out[1] = squared;
}

static uint32_t observation_blocks(const Oracle* o) { return (o->M + 63) / 64; }

/* bundleMetricPartialsOracle b:346-355 then finishBundleMetrics b:356-377 */
static void initial_metrics(Oracle* o) {
  uint32_t blocks = observation_blocks(o);
#pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < (int64_t)blocks; b++) {
    double t[3] = {0, 0, 0}, v[3];
    for (uint32_t lane = 0; lane < 64; lane++) {
      metric_lane(o, (uint32_t)b * 64 + lane, v);
      t[0] += v[0];
      t[1] += v[1];
      t[2] += v[2];
    }
    o->partials[b * 3] = t[0];
    o->partials[b * 3 + 1] = t[1];
    o->partials[b * 3 + 2] = t[2];
  }
  if (o->state[14] != 0) return;
  double cost = 0, squares = 0, invalid = 0;
  for (uint32_t b = 0; b < blocks; b++) {
    cost += o->partials[b * 3];
    squares += o->partials[b * 3 + 1];
    invalid += o->partials[b * 3 + 2];
  }
  if (invalid > 0) {
    o->state[14] = 2;
    o->state[1] = 0;
    return;
  }
  o->state[6] = cost;
  o->state[7] = squares;
  o->state[8] = cost;
  o->state[9] = squares;
  if (cost < 1e-20) o->state[1] = 0;
}

/* linearizeBundle, bundle.shade.ts:251-297 / joint-focal.shade.ts:287-336 */
static void linearize(Oracle* o) {
  if (o->state[1] == 0) return;
#pragma omp parallel for schedule(static)
  for (int64_t ii = 0; ii < (int64_t)o->M; ii++) {
    uint32_t i = (uint32_t)ii;
    uint32_t c = OBS_CAMERA(o, i), p = OBS_POINT(o, i);
    const double* pose = pose_at(o, c, 0);
    double q[3];
    observed_point(o, pose, point_at(o, p, 0), p, q);
    Model k = camera_at(o, c, 0);
    Projection pr = project(&k, q);
    double* b = o->J + (size_t)i * 24;
    for (int j = 0; j < 24; j++) b[j] = 0;
    if (!pr.ok) {
      b[23] = 1;
      continue;
    }
    const double* px = pixel_at(o, i);
    double ru = pr.px - px[0], rv = pr.py - px[1], squared = ru * ru + rv * rv;
    double error = sqrt(squared);
    double weight = 1;
    if (error > o->ctl.huber) weight = o->ctl.huber / error;
    if (VARIABLE(o, c)) {
      double ju[6], jv[6];
      pose_row(q, pr.du, ju);
      pose_row(q, pr.dv, jv);
      for (int kk = 0; kk < 6; kk++)
        if (kk < 3 || !directional(o, p)) {
          b[kk] = ju[kk];
          b[6 + kk] = jv[kk];
        }
    }
    for (int kk = 0; kk < 3; kk++) {
      b[12 + kk] = pr.du[0] * pose[kk] + pr.du[1] * pose[3 + kk] + pr.du[2] * pose[6 + kk];
      b[15 + kk] = pr.dv[0] * pose[kk] + pr.dv[1] * pose[3 + kk] + pr.dv[2] * pose[6 + kk];
    }
    if (o->shared) {
      double* f = o->focal + (size_t)o->C * 6 + 6 + (size_t)i * 2;
      f[0] = pr.px - k.cx;
      f[1] = pr.py - k.cy;
    }
    b[18] = ru;
    b[19] = rv;
    b[20] = weight;
    b[21] = robust_cost(error, o->ctl.huber);
    b[22] = squared;
  }
}

/* bundlePointBlocks, bundle.shade.ts:378-422 / joint-focal.shade.ts:417-462 */
static void point_blocks(Oracle* o) {
  if (o->state[1] == 0) return;
  const uint32_t K = o->K;
#pragma omp parallel for schedule(static)
  for (int64_t pp = 0; pp < (int64_t)o->P; pp++) {
    uint32_t p = (uint32_t)pp;
    double V[9] = {0}, rhs[3] = {0};
    double* at = o->pblk + (size_t)p * 13;
    for (int j = 0; j < 13; j++) at[j] = 0;
    int failed = 0;
    for (uint32_t i = POINT_START(o, p); i < POINT_START(o, p + 1); i++) {
      const double* base = o->J + (size_t)i * 24;
      double w = base[20];
      double* cross = o->cross + (size_t)i * K;
      int active = VARIABLE(o, OBS_CAMERA(o, i));
      if (base[23] != 0) {
        failed = 1;
        break;
      }
      for (uint32_t j = 0; j < K; j++) cross[j] = 0;
      const double* f = o->shared ? o->focal + (size_t)o->C * 6 + 6 + (size_t)i * 2 : NULL;
      for (int row = 0; row < 2; row++)
        for (int j = 0; j < 3; j++) {
          double g = base[12 + row * 3 + j];
          rhs[j] -= w * g * base[18 + row];
          for (int k = 0; k < 3; k++) V[j * 3 + k] += w * g * base[12 + row * 3 + k];
          if (o->shared) cross[18 + j] += w * f[row] * g;
          if (active)
            for (int k = 0; k < 6; k++) cross[k * 3 + j] += w * base[row * 6 + k] * g;
        }
    }
    if (failed) continue;
    for (int j = 0; j < 3; j++) V[j * 3 + j] += o->state[0] * fmax(V[j * 3 + j], 1e-10);
    for (int column = 0; column < 3; column++) {
      double e[3] = {0, 0, 0}, x[3];
      e[column] = 1;
      if (!solve3(V, e, x)) {
        failed = 1;
        break;
      }
      at[column] = x[0];
      at[3 + column] = x[1];
      at[6 + column] = x[2];
      at[9 + column] = rhs[column];
    }
    if (failed) continue;
    for (int j = 0; j < 3; j++)
      for (int k = j + 1; k < 3; k++) {
        double v = (at[j * 3 + k] + at[k * 3 + j]) * 0.5;
        at[j * 3 + k] = v;
        at[k * 3 + j] = v;
      }
    at[12] = 1;
  }
}

/* bundleCameraBlocks, bundle.shade.ts:423-489 (the CPU tier's resident camera assembly) */
static void camera_blocks(Oracle* o) {
  if (o->state[1] == 0) return;
  const uint32_t K = o->K;
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t cc = 0; cc < (int64_t)o->C; cc++) {
    uint32_t c = (uint32_t)cc;
    double* out = o->cam + (size_t)c * 36;
    double* inv = o->inv + (size_t)c * 36;
    double* bAt = o->rhs + (size_t)c * 6;
    double A[36] = {0}, S[36] = {0}, rhs[6] = {0};
    for (int j = 0; j < 36; j++) out[j] = inv[j] = 0;
    for (int j = 0; j < 6; j++) bAt[j] = 0;
    if (!VARIABLE(o, c)) {
      for (int j = 0; j < 6; j++) {
        out[j * 6 + j] = 1;
        inv[j * 6 + j] = 1;
      }
      continue;
    }
    for (uint32_t kk = CAMERA_START(o, c); kk < CAMERA_START(o, c + 1); kk++) {
      uint32_t i = CAMERA_OBS(o, kk), p = OBS_POINT(o, i);
      const double* point = o->pblk + (size_t)p * 13;
      const double* base = o->J + (size_t)i * 24;
      const double* cross = o->cross + (size_t)i * K;
      double w = base[20];
      double T[18] = {0};
      for (int j = 0; j < 6; j++) {
        for (int row = 0; row < 2; row++) {
          rhs[j] -= w * base[row * 6 + j] * base[18 + row];
          for (int l = 0; l < 6; l++) A[j * 6 + l] += w * base[row * 6 + j] * base[row * 6 + l];
        }
        for (int l = 0; l < 3; l++)
          for (int r = 0; r < 3; r++) T[j * 3 + l] += cross[j * 3 + r] * point[r * 3 + l];
        for (int l = 0; l < 3; l++) rhs[j] -= T[j * 3 + l] * point[9 + l];
        for (int l = 0; l < 6; l++)
          for (int r = 0; r < 3; r++) S[j * 6 + l] -= T[j * 3 + r] * cross[l * 3 + r];
      }
    }
    for (int j = 0; j < 6; j++) A[j * 6 + j] += o->state[0] * fmax(A[j * 6 + j], 1e-10);
    for (int j = 0; j < 36; j++) {
      out[j] = A[j];
      S[j] += A[j];
    }
    for (int j = 0; j < 6; j++) bAt[j] = rhs[j];
    int failed = 0;
    for (int column = 0; column < 6; column++) {
      double e[6] = {0, 0, 0, 0, 0, 0}, x[6];
      e[column] = 1;
      if (!solve6(S, e, x)) {
        inv[0] = -1;
        failed = 1;
        break;
      }
      for (int j = 0; j < 6; j++) inv[j * 6 + column] = x[j];
    }
    if (failed) continue;
    for (int j = 0; j < 6; j++)
      for (int k = j + 1; k < 6; k++) {
        double v = (inv[j * 6 + k] + inv[k * 6 + j]) * 0.5;
        inv[j * 6 + k] = v;
        inv[k * 6 + j] = v;
      }
  }
}

/* The shared-focal assembly: bundleFocalCameraRows j:1138-1151, bundleFocalObservations
 * j:1155-1171, dampBundleFocal j:1172-1180, bundleFocalPoints j:1181-1204. */
static void focal_blocks(Oracle* o) {
  if (o->state[1] == 0) return;
  const uint32_t n = o->C * 6;
  double* focal = o->focal;
#pragma omp parallel for schedule(static)
  for (int64_t rr = 0; rr < (int64_t)n; rr++) {
    uint32_t row = (uint32_t)rr, c = row / 6, j = row % 6;
    double value = 0;
    if (VARIABLE(o, c))
      for (uint32_t k = CAMERA_START(o, c); k < CAMERA_START(o, c + 1); k++) {
        uint32_t i = CAMERA_OBS(o, k);
        const double* base = o->J + (size_t)i * 24;
        const double* f = focal + n + 6 + (size_t)i * 2;
        value += base[20] * (base[j] * f[0] + base[6 + j] * f[1]);
      }
    focal[row] = value;
  }
  double diagonal = 0, rhs = 0;
  for (uint32_t i = 0; i < o->M; i++) {
    const double* b = o->J + (size_t)i * 24;
    const double* f = focal + n + 6 + (size_t)i * 2;
    double u = f[0], v = f[1], w = b[20];
    diagonal += w * (u * u + v * v);
    rhs -= w * (u * b[18] + v * b[19]);
  }
  focal[n + 5] = diagonal;
  focal[n + 2] = rhs;
  /* dampBundleFocal */
  diagonal = focal[n + 5];
  diagonal += o->state[0] * fmax(diagonal, 1e-10);
  focal[n] = diagonal;
  focal[n + 5] = diagonal;
  /* bundleFocalPoints */
  diagonal = focal[n + 5];
  rhs = focal[n + 2];
  for (uint32_t p = 0; p < o->P; p++) {
    const double* at = o->pblk + (size_t)p * 13;
    double cross[3] = {0, 0, 0};
    for (uint32_t i = POINT_START(o, p); i < POINT_START(o, p + 1); i++)
      for (int j = 0; j < 3; j++) cross[j] += o->cross[(size_t)i * o->K + 18 + j];
    for (int j = 0; j < 3; j++) {
      double t = 0;
      for (int k = 0; k < 3; k++) t += cross[k] * at[k * 3 + j];
      diagonal -= t * cross[j];
      rhs -= t * at[9 + j];
    }
  }
  focal[n + 5] = diagonal;
  focal[n + 2] = rhs;
  focal[n + 1] = -1;
  if (diagonal > 0) focal[n + 1] = 1 / diagonal;
}

#define VX(o) ((o)->vec)
#define VR(o) ((o)->vec + (o)->m)
#define VP(o) ((o)->vec + 2 * (size_t)(o)->m)
#define VS(o) ((o)->vec + 3 * (size_t)(o)->m)
#define VZ(o) ((o)->vec + 4 * (size_t)(o)->m)
#define VY(o) ((o)->vec + 5 * (size_t)(o)->m)

/* initializeBundleCg, bundle.shade.ts:695-741 / joint-focal.shade.ts:646-700 */
static void initialize_cg(Oracle* o) {
  for (int j = 18; j < 26; j++) o->state[j] = 0;
  if (o->state[1] == 0) return;
  for (uint32_t p = 0; p < o->P; p++)
    if (o->pblk[(size_t)p * 13 + 12] == 0) {
      o->state[20] = 1;
      return;
    }
  const uint32_t m = o->m;
  double rho = 0, residual = 0;
  for (uint32_t c = 0; c < o->C; c++) {
    const double* inv = o->inv + (size_t)c * 36;
    const double* rhs = o->rhs + (size_t)c * 6;
    if (!(inv[0] > 0)) {
      o->state[20] = 1;
      return;
    }
    for (int j = 0; j < 6; j++) {
      uint32_t i = c * 6 + j;
      double r = rhs[j], z = 0;
      for (int k = 0; k < 6; k++) z += inv[j * 6 + k] * rhs[k];
      o->vec[i] = 0;
      o->vec[m + i] = r;
      o->vec[2 * m + i] = z;
      o->vec[4 * m + i] = z;
      rho += r * z;
      residual += r * r;
    }
  }
  if (o->shared) {
    uint32_t f = m - 1;
    double r = o->focal[f + 2], z = o->focal[f + 1] * r;
    if (!(o->focal[f + 1] > 0)) {
      o->state[20] = 1;
      return;
    }
    o->vec[f] = 0;
    o->vec[m + f] = r;
    o->vec[2 * m + f] = z;
    o->vec[4 * m + f] = z;
    rho += r * z;
    residual += r * r;
  }
  o->state[18] = rho;
  o->state[21] = residual;
  o->state[22] = residual;
  if (residual > 0 && rho > 0)
    o->state[19] = 1;
  else if (residual > 1e-24)
    o->state[20] = 1;
}

/* bundleSchurPointAt b:185-203 / j:198-218: y_p = V_p^-1 sum W_i^T p (+ focal). */
static void schur_points(Oracle* o, const double* pv, double* y) {
  const uint32_t K = o->K, m = o->m;
#pragma omp parallel for schedule(static)
  for (int64_t pp = 0; pp < (int64_t)o->P; pp++) {
    uint32_t p = (uint32_t)pp;
    const double* at = o->pblk + (size_t)p * 13;
    double rhs[3] = {0, 0, 0};
    for (uint32_t i = POINT_START(o, p); i < POINT_START(o, p + 1); i++) {
      uint32_t c = OBS_CAMERA(o, i);
      const double* cross = o->cross + (size_t)i * K;
      for (int j = 0; j < 3; j++) {
        if (o->shared) rhs[j] += cross[18 + j] * pv[m - 1];
        for (int k = 0; k < 6; k++) rhs[j] += cross[k * 3 + j] * pv[c * 6 + k];
      }
    }
    for (int j = 0; j < 3; j++) {
      double v = 0;
      for (int k = 0; k < 3; k++) v += at[j * 3 + k] * rhs[k];
      y[(size_t)p * 3 + j] = v;
    }
  }
}

/* bundleSchurCameraRow b:204-225 / j:219-258 */
static void schur_cameras(Oracle* o, const double* pv, const double* y, double* sp) {
  const uint32_t K = o->K, n = o->C * 6, m = o->m;
#pragma omp parallel for schedule(static)
  for (int64_t rr = 0; rr < (int64_t)n; rr++) {
    uint32_t row = (uint32_t)rr, c = row / 6, j = row % 6;
    const double* A = o->cam + (size_t)c * 36;
    double v = 0;
    if (o->shared) v = o->focal[row] * pv[m - 1];
    for (int k = 0; k < 6; k++) v += A[j * 6 + k] * pv[c * 6 + k];
    for (uint32_t kk = CAMERA_START(o, c); kk < CAMERA_START(o, c + 1); kk++) {
      uint32_t i = CAMERA_OBS(o, kk), p = OBS_POINT(o, i);
      const double* cross = o->cross + (size_t)i * K;
      for (int k = 0; k < 3; k++) v -= cross[j * 3 + k] * y[(size_t)p * 3 + k];
    }
    sp[row] = v;
  }
  if (o->shared) {
    double v = o->focal[n] * pv[n];
    for (uint32_t k = 0; k < n; k++) v += o->focal[k] * pv[k];
    for (uint32_t i = 0; i < o->M; i++) {
      uint32_t p = OBS_POINT(o, i);
      const double* cross = o->cross + (size_t)i * K;
      for (int k = 0; k < 3; k++) v -= cross[18 + k] * y[(size_t)p * 3 + k];
    }
    sp[n] = v;
  }
}

/* bundlePreconditionedResidualAt b:226-233 / j:259-267 */
static double preconditioned(const Oracle* o, uint32_t i, const double* r) {
  if (o->shared && i == o->m - 1) return o->focal[o->C * 6 + 1] * r[i];
  uint32_t c = i / 6, j = i % 6;
  const double* inv = o->inv + (size_t)c * 36;
  double z = 0;
  for (int k = 0; k < 6; k++) z += inv[j * 6 + k] * r[c * 6 + k];
  return z;
}

/* advanceBundleCg b:883-929 / advancePcg(false) j:874-924 */
static void advance_cg(Oracle* o) {
  if (o->state[19] == 0) return;
  const uint32_t m = o->m;
  double* x = VX(o);
  double* r = VR(o);
  double* p = VP(o);
  double* s = VS(o);
  double* z = VZ(o);
  double denominator = 0;
  for (uint32_t i = 0; i < m; i++) denominator += p[i] * s[i];
  if (!(denominator > 0)) {
    o->state[19] = 0;
    o->state[20] = 1;
    return;
  }
  double alpha = o->state[18] / denominator;
  if (!(fabs(alpha) < 1e20)) {
    o->state[19] = 0;
    o->state[20] = 1;
    return;
  }
  double residual = 0, rho = 0;
  for (uint32_t i = 0; i < m; i++) {
    x[i] += alpha * p[i];
    r[i] -= alpha * s[i];
    residual += r[i] * r[i];
  }
  for (uint32_t i = 0; i < m; i++) {
    double zi = preconditioned(o, i, r);
    z[i] = zi;
    rho += zi * r[i];
  }
  o->state[16]++;
  o->state[23]++;
  o->state[22] = residual;
  double tol = o->ctl.relative_tolerance;
  if (residual <= tol * tol * o->state[21]) {
    o->state[19] = 0;
    o->state[24] = 1;
    return;
  }
  if (!(rho > 0)) {
    o->state[19] = 0;
    o->state[20] = 1;
    return;
  }
  double beta = rho / o->state[18];
  o->state[18] = rho;
  for (uint32_t i = 0; i < m; i++) p[i] = z[i] + beta * p[i];
}

/* updateBundleCandidate b:930-961 / j:925-960 */
static void update_candidate(Oracle* o) {
  if (o->state[1] == 0 || o->state[20] != 0) return;
  const uint32_t K = o->K, m = o->m;
  const double* x = VX(o);
  if (o->shared) o->focal[o->C * 6 + 4] = o->focal[o->C * 6 + 3] + x[m - 1];
  for (uint32_t h = 0; h < o->C; h++) {
    const double* p = pose_at(o, h, 0);
    double moved[12];
    if (VARIABLE(o, h))
      move_pose(p, x + (size_t)h * 6, moved);
    else
      memcpy(moved, p, sizeof moved);
    memcpy(o->trial + (size_t)h * 12, moved, sizeof moved);
  }
#pragma omp parallel for schedule(static)
  for (int64_t pp = 0; pp < (int64_t)o->P; pp++) {
    uint32_t point = (uint32_t)pp;
    const double* at = o->pblk + (size_t)point * 13;
    size_t out = (size_t)o->C * 12 + (size_t)point * 3;
    double rhs[3] = {at[9], at[10], at[11]};
    for (uint32_t i = POINT_START(o, point); i < POINT_START(o, point + 1); i++) {
      uint32_t c = OBS_CAMERA(o, i);
      const double* cross = o->cross + (size_t)i * K;
      for (int j = 0; j < 3; j++) {
        if (o->shared) rhs[j] -= cross[18 + j] * x[m - 1];
        for (int k = 0; k < 6; k++) rhs[j] -= cross[k * 3 + j] * x[c * 6 + k];
      }
    }
    for (int j = 0; j < 3; j++) {
      double v = o->scene[out + j];
      for (int k = 0; k < 3; k++) v += at[j * 3 + k] * rhs[k];
      o->trial[out + j] = v;
    }
  }
}

/* bundleCandidateGauge b:962-993 / j:961-1000 */
static void candidate_gauge(Oracle* o) {
  o->state[15] = 0;
  if (o->state[1] == 0 || o->state[20] != 0) return;
  double tol = o->ctl.relative_tolerance;
  if (o->state[21] > 0 && o->state[22] > tol * tol * o->state[21]) {
    o->state[20] = 1;
    return;
  }
  if (o->shared) {
    double log_scale = o->focal[o->C * 6 + 4];
    if (!(fabs(log_scale) < 700)) return;
    double fs = focal_scale(log_scale);
    for (uint32_t c = 0; c < o->C; c++) {
      const double* s = o->scene + (size_t)o->C * 12 + (size_t)o->P * 3 + (size_t)c * 6;
      if (!(s[0] * fs >= 1e-6 && s[1] * fs >= 1e-6 && s[0] * fs <= 1e6 && s[1] * fs <= 1e6))
        return;
    }
  }
  double center[3], d[3];
  camera_center(o, o->ctl.scale_view, 1, center);
  for (int r = 0; r < 3; r++) d[r] = o->state[3 + r] - center[r];
  double baseline = length3(d);
  if (!(baseline > 1e-12)) return;
  double scale = 1;
  if (o->ctl.normalize_scale != 0) scale = o->state[2] / baseline;
  if (!(scale > 1e-9 && scale < 1e9)) return;
  for (uint32_t c = 0; c < o->C; c++) {
    const double* p = pose_at(o, c, 1);
    for (int j = 0; j < 12; j++)
      if (!(fabs(p[j]) <= 1e9)) return;
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++) {
        double dd = 0;
        for (int r = 0; r < 3; r++) dd += p[j * 3 + r] * p[k * 3 + r];
        if (j == k) dd -= 1;
        if (!(fabs(dd) < 1e-6)) return;
      }
  }
  o->state[13] = scale;
  o->state[15] = 1;
}

/* normalizeBundleCandidate b:994-1023 / j:1001-1030 */
static void normalize_candidate(Oracle* o) {
  if (o->state[1] == 0 || o->state[15] == 0) return;
  uint32_t n = o->C > o->P ? o->C : o->P;
  double s = o->state[13];
  const double orig[3] = {o->state[3], o->state[4], o->state[5]};
  for (uint32_t h = 0; h < n; h++) {
    if (h < o->P && directional(o, h)) {
      double* d = o->trial + (size_t)o->C * 12 + (size_t)h * 3;
      double norm = length3(d);
      if (norm > 1e-12) {
        d[0] = d[0] / norm;
        d[1] = d[1] / norm;
        d[2] = d[2] / norm;
      }
    }
    if (o->ctl.normalize_scale == 0) continue;
    if (h < o->P && !directional(o, h)) {
      double* X = o->trial + (size_t)o->C * 12 + (size_t)h * 3;
      for (int j = 0; j < 3; j++) X[j] = orig[j] + s * (X[j] - orig[j]);
    }
    if (h < o->C && h != o->ctl.anchor) {
      double* p = o->trial + (size_t)h * 12;
      double rotated[3];
      for (int r = 0; r < 3; r++)
        rotated[r] = p[r * 3] * orig[0] + p[r * 3 + 1] * orig[1] + p[r * 3 + 2] * orig[2] + p[9 + r] -
                     p[9 + r];
      double t0 = s * p[9] + (s - 1) * rotated[0];
      double t1 = s * p[10] + (s - 1) * rotated[1];
      double t2 = s * p[11] + (s - 1) * rotated[2];
      p[9] = t0;
      p[10] = t1;
      p[11] = t2;
    }
  }
}

/* candidateLane b:1024-1037 / j:1031-1044 */
static void candidate_lane(const Oracle* o, uint32_t i, double out[2]) {
  out[0] = out[1] = 0;
  if (i >= o->M) return;
  uint32_t c = OBS_CAMERA(o, i), p = OBS_POINT(o, i);
  const double* X = point_at(o, p, 1);
  if (!(fmax(fabs(X[0]), fmax(fabs(X[1]), fabs(X[2]))) <= 1e9)) {
    out[0] = out[1] = 1e30;
    return;
  }
  double q[3];
  observed_point(o, pose_at(o, c, 1), X, p, q);
  Model k = camera_at(o, c, 1);
  Projection pr = project(&k, q);
  if (!pr.ok) {
    out[0] = out[1] = 1e30;
    return;
  }
  const double* px = pixel_at(o, i);
  double dx = pr.px - px[0], dy = pr.py - px[1], squares = dx * dx + dy * dy;
  out[0] = robust_cost(sqrt(squares), o->ctl.huber);
  out[1] = squares;
}

/* bundleCandidatePartialsOracle b:1062-1070, acceptBundleCandidate b:1071-1096,
 * commitBundleCandidate b:1109-1114 (joint-focal j:1069-1122) */
static void accept_candidate(Oracle* o) {
  uint32_t blocks = observation_blocks(o);
#pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < (int64_t)blocks; b++) {
    double t[2] = {0, 0}, v[2];
    for (uint32_t lane = 0; lane < 64; lane++) {
      candidate_lane(o, (uint32_t)b * 64 + lane, v);
      t[0] += v[0];
      t[1] += v[1];
    }
    o->partials[b * 2] = t[0];
    o->partials[b * 2 + 1] = t[1];
  }
  o->state[12] = 0;
  if (o->state[1] == 0) return;
  double cost = 0, squares = 0;
  if (o->state[15] != 0)
    for (uint32_t b = 0; b < blocks; b++) {
      cost += o->partials[b * 2];
      squares += o->partials[b * 2 + 1];
    }
  if (o->state[15] != 0 && cost < o->state[6]) {
    double change = o->state[6] - cost;
    o->state[6] = cost;
    o->state[7] = squares;
    o->state[10]++;
    o->state[12] = 1;
    o->state[0] = fmax(1e-10, o->state[0] / 3);
    if (change < 1e-9 * fmax(1, cost)) o->state[1] = 0;
  } else {
    o->state[11]++;
    o->state[0] *= 10;
    if (o->state[20] != 0) o->state[17]++;
    if (o->state[0] > 1e12) o->state[1] = 0;
  }
  if (o->state[12] != 0) {
    memcpy(o->scene, o->trial, ((size_t)o->C * 12 + (size_t)o->P * 3) * sizeof(double));
    if (o->shared) o->focal[o->C * 6 + 3] = o->focal[o->C * 6 + 4];
  }
}

/* ---------------------------------------------------------------- allocation */

static void oracle_free(Oracle* o) {
  free(o->scene);
  free(o->trial);
  free(o->J);
  free(o->cross);
  free(o->pblk);
  free(o->cam);
  free(o->inv);
  free(o->rhs);
  free(o->focal);
  free(o->vec);
  free(o->partials);
}

static int oracle_init(Oracle* o, const tcv_ba_control* ctl, const double* scene,
                       const uint32_t* index) {
  memset(o, 0, sizeof *o);
  o->ctl = *ctl;
  o->C = ctl->cameras;
  o->P = ctl->points;
  o->M = ctl->observations;
  o->shared = ctl->shared_focal ? 1 : 0;
  o->m = o->C * 6 + o->shared;
  o->K = o->shared ? 21 : 18;
  size_t scene_size = (size_t)o->C * 18 + (size_t)o->P * 3 + (size_t)o->M * 2;
  size_t trial_size = (size_t)o->C * 12 + (size_t)o->P * 3;
  o->scene = malloc(scene_size * sizeof(double));
  o->trial = calloc(trial_size, sizeof(double));
  o->index = index;
  o->J = calloc((size_t)o->M * 24 + 1, sizeof(double));
  o->cross = calloc((size_t)o->M * o->K + 1, sizeof(double));
  o->pblk = calloc((size_t)o->P * 13 + 1, sizeof(double));
  o->cam = calloc((size_t)o->C * 36, sizeof(double));
  o->inv = calloc((size_t)o->C * 36, sizeof(double));
  o->rhs = calloc((size_t)o->C * 6, sizeof(double));
  o->focal = calloc((size_t)o->C * 6 + 6 + (size_t)o->M * 2, sizeof(double));
  o->vec = calloc((size_t)o->m * 5 + (size_t)o->P * 3, sizeof(double));
  o->partials = calloc(3 * (size_t)((o->M + 63) / 64 + 1), sizeof(double));
  if (!o->scene || !o->trial || !o->J || !o->cross || !o->pblk || !o->cam || !o->inv || !o->rhs ||
      !o->focal || !o->vec || !o->partials) {
    oracle_free(o);
    return 0;
  }
  memcpy(o->scene, scene, scene_size * sizeof(double));
  return 1;
}

static void assemble(Oracle* o) {
  linearize(o);
  point_blocks(o);
  camera_blocks(o);
  if (o->shared) focal_blocks(o);
}

static double now_ms(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return t.tv_sec * 1e3 + t.tv_nsec * 1e-6;
}

/* ---------------------------------------------------------------- exported entry points */

/* Threads used by the OpenMP loops (0: the runtime default). Returns the count in effect. */
int tcv_oracle_set_threads(int threads) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
  threads_used = omp_get_max_threads();
#else
  (void)threads;
  threads_used = 1;
#endif
  return threads_used;
}

/*
 * The whole operator (bundle.ts:314-772 / joint-focal.ts:340-900 in CPU-tier order), the stall
 * monitor of bundle-shared.ts:83-123 included. `out` has the packBundleResult layout.
 * `stats` (optional, 4 doubles): LM iterations executed, CG iterations, wall ms, LM loop ms.
 * `max_lm` > 0 stops after that many LM iterations (a bounded timing sample); 0 runs them all.
 */
int tcv_oracle_solve(const tcv_ba_control* ctl, const double* scene, const uint32_t* index,
                     double* out, double* stats, uint32_t max_lm) {
  double t0 = now_ms();
  Oracle o;
  if (!oracle_init(&o, ctl, scene, index)) return TCV_BA_ERR_MEMORY;
  initialize_bundle(&o);
  initial_metrics(&o);
  double t1 = now_ms();
  int lm_active = 1;
  double monitor_cost = o.state[6];
  const int monitor = ctl->stall_tolerance > 1e-9;
  uint32_t executed = 0;
  for (uint32_t iteration = 0; iteration < ctl->iterations && lm_active; iteration++) {
    if (max_lm && executed >= max_lm) break;
    if (o.state[1] != 0) executed++;
    assemble(&o);
    initialize_cg(&o);
    for (uint32_t k = 0; k < ctl->max_cg && o.state[19] != 0; k++) {
      schur_points(&o, VP(&o), VY(&o));
      schur_cameras(&o, VP(&o), VY(&o), VS(&o));
      advance_cg(&o);
    }
    update_candidate(&o);
    candidate_gauge(&o);
    normalize_candidate(&o);
    accept_candidate(&o);
    lm_active = o.state[1] != 0;
    if (monitor && lm_active) {
      double next = o.state[6], change = monitor_cost - next;
      int accepted = o.state[12] == 1;
      monitor_cost = next;
      if (accepted && change < ctl->stall_tolerance * fmax(1, next)) lm_active = 0;
    }
  }
  double t2 = now_ms();
  size_t total = (size_t)o.C * 12 + (size_t)o.P * 3;
  memcpy(out, o.scene, total * sizeof(double));
  for (int i = 0; i < 26; i++) out[total + i] = o.state[i];
  out[total + 24] = sqrt(o.state[9] / (double)(float)o.M);
  out[total + 25] = sqrt(o.state[7] / (double)(float)o.M);
  if (o.shared) out[total + 26] = o.focal[o.C * 6 + 3];
  if (stats) {
    stats[0] = executed;
    stats[1] = o.state[16];
    stats[2] = t2 - t0;
    stats[3] = t2 - t1;
  }
  oracle_free(&o);
  return TCV_BA_OK;
}

/*
 * One linear system: linearise at damping `lambda` and log focal `log_focal`, assemble, and return
 * the reduced RHS, the Schur product S·p, the preconditioned vector M^-1·p and the validity flag.
 * Optional block outputs (NULL to skip): J (24M), cross (K·M), point blocks (13P), camera blocks
 * (36C), preconditioner blocks (36C), focal array (6C + 6 + 2M).
 */
int tcv_oracle_schur_probe(const tcv_ba_control* ctl, const double* scene, const uint32_t* index,
                           double lambda, double log_focal, const double* p, double* rhs,
                           double* sp, double* z, int32_t* status, double* J_out,
                           double* cross_out, double* pblk_out, double* cam_out, double* inv_out,
                           double* focal_out) {
  Oracle o;
  if (!oracle_init(&o, ctl, scene, index)) return TCV_BA_ERR_MEMORY;
  for (int j = 0; j < 26; j++) o.state[j] = 0;
  o.state[0] = lambda;
  o.state[1] = 1;
  if (o.shared) o.focal[o.C * 6 + 3] = o.focal[o.C * 6 + 4] = log_focal;
  assemble(&o);
  int32_t bad = 0;
  for (uint32_t q = 0; q < o.P; q++)
    if (o.pblk[(size_t)q * 13 + 12] == 0) bad = 1;
  for (uint32_t c = 0; c < o.C; c++)
    if (!(o.inv[(size_t)c * 36] > 0)) bad = 1;
  if (o.shared && !(o.focal[o.C * 6 + 1] > 0)) bad = 1;
  *status = bad;
  for (uint32_t c = 0; c < o.C; c++)
    for (int j = 0; j < 6; j++) rhs[c * 6 + j] = o.rhs[(size_t)c * 6 + j];
  if (o.shared) rhs[o.m - 1] = o.focal[o.C * 6 + 2];
  schur_points(&o, p, VY(&o));
  schur_cameras(&o, p, VY(&o), sp);
  for (uint32_t i = 0; i < o.m; i++) z[i] = preconditioned(&o, i, p);
  if (J_out) memcpy(J_out, o.J, (size_t)o.M * 24 * sizeof(double));
  if (cross_out) memcpy(cross_out, o.cross, (size_t)o.M * o.K * sizeof(double));
  if (pblk_out) memcpy(pblk_out, o.pblk, (size_t)o.P * 13 * sizeof(double));
  if (cam_out) memcpy(cam_out, o.cam, (size_t)o.C * 36 * sizeof(double));
  if (inv_out) memcpy(inv_out, o.inv, (size_t)o.C * 36 * sizeof(double));
  if (focal_out)
    memcpy(focal_out, o.focal, ((size_t)o.C * 6 + 6 + (size_t)o.M * 2) * sizeof(double));
  oracle_free(&o);
  return TCV_BA_OK;
}

/*
 * One LM iteration's linear solve: assemble at `lambda` / `log_focal`, initializeBundleCg, then
 * PCG (advanceBundleCg) for at most control->max_cg iterations. Returns x (m) and
 * stats[0..3] = {PCG steps, converged, failed, final |r|^2 / |r0|^2}.
 */
int tcv_oracle_pcg(const tcv_ba_control* ctl, const double* scene, const uint32_t* index,
                   double lambda, double log_focal, double* x, double* stats) {
  Oracle o;
  if (!oracle_init(&o, ctl, scene, index)) return TCV_BA_ERR_MEMORY;
  for (int j = 0; j < 26; j++) o.state[j] = 0;
  o.state[0] = lambda;
  o.state[1] = 1;
  if (o.shared) o.focal[o.C * 6 + 3] = o.focal[o.C * 6 + 4] = log_focal;
  assemble(&o);
  initialize_cg(&o);
  for (uint32_t k = 0; k < ctl->max_cg && o.state[19] != 0; k++) {
    schur_points(&o, VP(&o), VY(&o));
    schur_cameras(&o, VP(&o), VY(&o), VS(&o));
    advance_cg(&o);
  }
  memcpy(x, VX(&o), (size_t)o.m * sizeof(double));
  stats[0] = o.state[23];
  stats[1] = o.state[24];
  stats[2] = o.state[20];
  stats[3] = o.state[21] > 0 ? o.state[22] / o.state[21] : 0;
  oracle_free(&o);
  return TCV_BA_OK;
}

/* Scalar helpers, exported for the pins. */
int tcv_oracle_project(const double* model, const double* q, double* out8) {
  Model k = {model[0], model[1], model[2], model[3], model[4], model[5]};
  Projection p = project(&k, q);
  out8[0] = p.px;
  out8[1] = p.py;
  for (int j = 0; j < 3; j++) {
    out8[2 + j] = p.du[j];
    out8[5 + j] = p.dv[j];
  }
  return p.ok;
}
double tcv_oracle_radial_limit(const double* model) {
  Model k = {model[0], model[1], model[2], model[3], model[4], model[5]};
  return radial_limit(&k);
}
void tcv_oracle_pose_row(const double* q, const double* j, double* out6) { pose_row(q, j, out6); }
double tcv_oracle_robust_cost(double error, double huber) { return robust_cost(error, huber); }
int tcv_oracle_solve3(const double* a, const double* b, double* x) { return solve3(a, b, x); }
int tcv_oracle_solve6(const double* a, const double* b, double* x) { return solve6(a, b, x); }
void tcv_oracle_move_pose(const double* p, const double* d, double* out) { move_pose(p, d, out); }
double tcv_oracle_focal_scale(double x) { return focal_scale(x); }

/* ------------------------------------------------ fixed-scene observation evaluation
 *
 * evaluateObservationsGpu (src/geometry/gpu/observations.ts:34-127) as its CPU tier runs the
 * kernels of observation-evaluation.shade.ts, in their order:
 *   projectIndexedObservations      :44-54   flag and error of every observation
 *   filterIndexedObservationDepth   :85-97   (guarded) tracks whose depth is not stable lose all
 *                                            their flags; flag count + h marks track h rejected
 *   indexedObservationPartialsOracle :128-137 per 64 observations, errors and counts summed in lane
 *                                            order (the CPU tier's partials kernel)
 *   observationSummary  triangulation.shade.ts:134-143  the partials summed in block order
 * cameras: 18 per view (pose record R row-major, t, then fx fy cx cy k1 k2); points: 3 per
 * landmark; pixels: 2 per observation; indices: (view, landmark) per observation; offsets: the
 * tracks' first observations (tracks + 1), guarded runs only. flags: count (+ tracks) words;
 * errors: count; summary: [error sum, inlier count].
 */
static void obs_transform(const double* p, const double* X, double q[3]) {
  q[0] = p[0] * X[0] + p[1] * X[1] + p[2] * X[2] + p[9];
  q[1] = p[3] * X[0] + p[4] * X[1] + p[5] * X[2] + p[10];
  q[2] = p[6] * X[0] + p[7] * X[1] + p[8] * X[2] + p[11];
}

static Projection obs_project(const double* cameras, const uint32_t* indices, const double* points,
                              uint32_t i, const double* X) {
  const double* rec = cameras + (size_t)indices[2 * i] * 18;
  const Model k = {rec[12], rec[13], rec[14], rec[15], rec[16], rec[17]};
  double q[3];
  obs_transform(rec, X ? X : points + 3 * (size_t)indices[2 * i + 1], q);
  return project(&k, q);
}

/* stableDepth, observation-evaluation.shade.ts:57-83 */
static int obs_stable_depth(const double* cameras, const uint32_t* indices, const uint32_t* flags,
                            uint32_t start, uint32_t end, const double* X, double limit,
                            double noise) {
  double A[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  uint32_t count = 0;
  for (uint32_t i = start; i < end; i++) {
    if (flags[i] == 0) continue;
    count++;
    const double* p = cameras + (size_t)indices[2 * i] * 18;
    Projection pr = obs_project(cameras, indices, NULL, i, X);
    if (!pr.ok) return 0;
    for (int row = 0; row < 2; row++) {
      const double* g = row == 0 ? pr.du : pr.dv;
      const double J[3] = {g[0] * p[0] + g[1] * p[3] + g[2] * p[6],
                           g[0] * p[1] + g[1] * p[4] + g[2] * p[7],
                           g[0] * p[2] + g[1] * p[5] + g[2] * p[8]};
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) A[3 * r + c] += J[r] * J[c];
    }
  }
  if (count < 2) return 0;
  for (uint32_t i = start; i < end; i++) {
    if (flags[i] == 0) continue;
    const double* p = cameras + (size_t)indices[2 * i] * 18;
    const double axis[3] = {p[6], p[7], p[8]};
    double q[3], solved[3];
    obs_transform(p, X, q);
    const int ok = solve3(A, axis, solved);
    if (!ok || !(q[2] > 0)) return 0;
    const double variance = axis[0] * solved[0] + axis[1] * solved[1] + axis[2] * solved[2];
    if (!(variance >= 0)) return 0;
    const double relative = noise * sqrt(variance) / q[2];
    if (!(relative <= limit)) return 0;
  }
  return 1;
}

void tcv_oracle_observations(const double* cameras, const double* points, const double* pixels,
                             const uint32_t* indices, uint32_t count, const uint32_t* offsets,
                             uint32_t tracks, double max_error, double max_relative_depth_std,
                             double pixel_noise_std, uint32_t* flags, double* errors,
                             double* summary) {
  for (uint32_t i = 0; i < count; i++) {
    flags[i] = 0;
    errors[i] = 0;
    Projection pr = obs_project(cameras, indices, points, i, NULL);
    if (!pr.ok) continue;
    const double x = pr.px - pixels[2 * (size_t)i], y = pr.py - pixels[2 * (size_t)i + 1];
    const double e = sqrt(x * x + y * y);
    if (e <= max_error) {
      flags[i] = 1;
      errors[i] = e;
    }
  }
  if (offsets)
    for (uint32_t h = 0; h < tracks; h++) {
      const uint32_t start = offsets[h], end = offsets[h + 1];
      flags[count + h] = 0;
      if (end <= start) continue;
      uint32_t accepted = 0;
      for (uint32_t i = start; i < end; i++) accepted += flags[i] != 0;
      if (accepted < 2) continue;
      const double* X = points + 3 * (size_t)indices[2 * start + 1];
      if (!obs_stable_depth(cameras, indices, flags, start, end, X, max_relative_depth_std,
                            pixel_noise_std)) {
        flags[count + h] = 1;
        for (uint32_t i = start; i < end; i++) {
          flags[i] = 0;
          errors[i] = 0;
        }
      }
    }
  const uint32_t blocks = (count > 0 ? count : 1) / 64 + ((count > 0 ? count : 1) % 64 != 0);
  double error = 0, inliers = 0;
  for (uint32_t b = 0; b < blocks; b++) {
    double e = 0, n = 0;
    for (uint32_t lane = 0; lane < 64; lane++) {
      const uint32_t i = b * 64 + lane;
      if (i >= count || flags[i] == 0) continue;
      e += errors[i];
      n += 1;
    }
    error += e;
    inliers += n;
  }
  summary[0] = error;
  summary[1] = inliers;
}

/*
 * The gradient of the half robust cost E = 1/2 sum_i robustCost(|r_i|) in the movePose chart at
 * log focal scale `log_focal`: 6 per camera (omega, v; zero for a fixed camera), 3 per point, then
 * the log focal scale on the shared route: g = sum_i J_i^T w_i r_i with the forward's Jacobians
 * and Huber weights (linearizeBundle). `half_cost` (optional) receives E. Observations that do not
 * project are skipped; returns 1 when there was one, else 0. A negative value: out of memory.
 */
int tcv_oracle_gradient(const tcv_ba_control* ctl, const double* scene, const uint32_t* index,
                        double log_focal, double* grad, double* half_cost) {
  Oracle o;
  if (!oracle_init(&o, ctl, scene, index)) return -TCV_BA_ERR_MEMORY;
  for (int j = 0; j < 26; j++) o.state[j] = 0;
  o.state[1] = 1;
  if (o.shared) o.focal[o.C * 6 + 3] = o.focal[o.C * 6 + 4] = log_focal;
  linearize(&o);
  const size_t n = (size_t)o.C * 6 + (size_t)o.P * 3 + o.shared;
  for (size_t k = 0; k < n; k++) grad[k] = 0;
  double cost = 0;
  int skipped = 0;
  for (uint32_t i = 0; i < o.M; i++) {
    const double* b = o.J + (size_t)i * 24;
    if (b[23] != 0) {
      skipped = 1;
      continue;
    }
    const uint32_t c = OBS_CAMERA(&o, i), p = OBS_POINT(&o, i);
    const double wu = b[20] * b[18], wv = b[20] * b[19];
    for (int k = 0; k < 6; k++) grad[(size_t)c * 6 + k] += b[k] * wu + b[6 + k] * wv;
    for (int k = 0; k < 3; k++)
      grad[(size_t)o.C * 6 + (size_t)p * 3 + k] += b[12 + k] * wu + b[15 + k] * wv;
    if (o.shared) {
      const double* f = o.focal + (size_t)o.C * 6 + 6 + (size_t)i * 2;
      grad[n - 1] += f[0] * wu + f[1] * wv;
    }
    cost += 0.5 * b[21];
  }
  if (half_cost) *half_cost = cost;
  oracle_free(&o);
  return skipped;
}

/* ---------------------------------------------------------------- the dense CPU reference */

/* inverse3, numerics.ts:100-131: closed-form adjugate inverse; 0 for a singular block. */
static int adjugate_inverse3(const double a[9], double out[9]) {
  const double a00 = a[0], a01 = a[1], a02 = a[2], a10 = a[3], a11 = a[4], a12 = a[5], a20 = a[6],
               a21 = a[7], a22 = a[8];
  const double c00 = a11 * a22 - a12 * a21, c01 = a02 * a21 - a01 * a22, c02 = a01 * a12 - a02 * a11;
  const double det = a00 * c00 + a10 * c01 + a20 * c02;
  if (!isfinite(det) || det == 0) return 0;
  out[0] = c00 / det;
  out[1] = c01 / det;
  out[2] = c02 / det;
  out[3] = (a12 * a20 - a10 * a22) / det;
  out[4] = (a00 * a22 - a02 * a20) / det;
  out[5] = (a02 * a10 - a00 * a12) / det;
  out[6] = (a10 * a21 - a11 * a20) / det;
  out[7] = (a01 * a20 - a00 * a21) / det;
  out[8] = (a00 * a11 - a01 * a10) / det;
  for (int k = 0; k < 9; k++)
    if (!isfinite(out[k])) return 0;
  return 1;
}

/* solveSystem, numerics.ts:70-98: Gaussian elimination with partial pivoting on [A | b]; 0 when a
 * pivot falls below 1e-14 times the largest |a_ij| or the system is not finite. */
static int solve_system(const double* A, const double* b, uint32_t n, double* out) {
  double* a = malloc((size_t)n * (n + 1) * sizeof(double));
  if (!a) return 0;
  double scale = 0;
  for (uint32_t i = 0; i < n; i++) {
    for (uint32_t j = 0; j < n; j++) {
      const double v = A[(size_t)i * n + j];
      if (!isfinite(v)) {
        free(a);
        return 0;
      }
      a[(size_t)i * (n + 1) + j] = v;
      if (fabs(v) > scale) scale = fabs(v);
    }
    a[(size_t)i * (n + 1) + n] = b[i];
    if (!isfinite(b[i])) {
      free(a);
      return 0;
    }
  }
  if (!(scale > 0)) {
    free(a);
    return n == 0;
  }
  for (uint32_t j = 0; j < n; j++) {
    uint32_t p = j;
    for (uint32_t i = j + 1; i < n; i++)
      if (fabs(a[(size_t)i * (n + 1) + j]) > fabs(a[(size_t)p * (n + 1) + j])) p = i;
    if (fabs(a[(size_t)p * (n + 1) + j]) < 1e-14 * scale) {
      free(a);
      return 0;
    }
    if (p != j)
      for (uint32_t k = 0; k <= n; k++) {
        const double t = a[(size_t)p * (n + 1) + k];
        a[(size_t)p * (n + 1) + k] = a[(size_t)j * (n + 1) + k];
        a[(size_t)j * (n + 1) + k] = t;
      }
    const double* rj = a + (size_t)j * (n + 1);
#pragma omp parallel for schedule(static)
    for (int64_t ii = j + 1; ii < (int64_t)n; ii++) {
      double* ri = a + (size_t)ii * (n + 1);
      const double f = ri[j] / rj[j];
      for (uint32_t k = j; k <= n; k++) ri[k] -= f * rj[k];
    }
  }
  for (int64_t i = (int64_t)n - 1; i >= 0; i--) {
    double v = a[(size_t)i * (n + 1) + n];
    for (uint32_t j = (uint32_t)i + 1; j < n; j++) v -= a[(size_t)i * (n + 1) + j] * out[j];
    out[i] = v / a[(size_t)i * (n + 1) + i];
  }
  free(a);
  for (uint32_t i = 0; i < n; i++)
    if (!isfinite(out[i])) return 0;
  return 1;
}

/* bundleMetrics (bundle-shared.ts:32-54) of a scene at focal scale `fscale`: robust cost and sum of
 * squares; 0 when an observation does not project. */
static int dense_metrics(const tcv_ba_control* ctl, const double* scene, const uint32_t* index,
                         double fscale, double* cost, double* squares) {
  const uint32_t C = ctl->cameras, P = ctl->points, M = ctl->observations;
  const double* pixels = scene + (size_t)C * 18 + (size_t)P * 3;
  double c = 0, s = 0;
  for (uint32_t i = 0; i < M; i++) {
    const uint32_t cam = index[i], p = index[M + i];
    const double* pose = scene + (size_t)cam * 12;
    const double* X = scene + (size_t)C * 12 + (size_t)p * 3;
    const double* mo = scene + (size_t)C * 12 + (size_t)P * 3 + (size_t)cam * 6;
    Model k = {mo[0] * fscale, mo[1] * fscale, mo[2], mo[3], mo[4], mo[5]};
    const int dir = p >= P - ctl->directional_points;
    double q[3];
    for (int r = 0; r < 3; r++)
      q[r] = pose[r * 3] * X[0] + pose[r * 3 + 1] * X[1] + pose[r * 3 + 2] * X[2] + (dir ? 0 : pose[9 + r]);
    Projection pr = project(&k, q);
    if (!pr.ok) return 0;
    const double e = hypot(pr.px - pixels[2 * (size_t)i], pr.py - pixels[2 * (size_t)i + 1]);
    c += e <= ctl->huber ? e * e : 2 * ctl->huber * e - ctl->huber * ctl->huber;
    s += e * e;
  }
  *cost = c;
  *squares = s;
  return 1;
}

/*
 * The reference's dense CPU bundle adjustment, `bundleAdjust` (bundle-adjustment.ts:80-291, fixed
 * intrinsics; joint-focal.ts:45-287 with one shared log focal scale): LM with point-block Schur
 * elimination into a dense reduced system over the variable cameras' 6 rows (and the focal row),
 * solved by solveSystem; damping lambda * max(diag, 1e-10), adjugate point inverses, the
 * one-fixed-camera similarity normalisation, accept on a lower robust cost, stop on the stall rule
 * (bundleStalled) or lambda > 1e12. `out` has the packBundleResult layout with words 6 (cost),
 * 7 (sum of squares), 8/9 (initial), 10/11 (accepted/rejected), 24/25 (RMSE) and, on the shared
 * route, 26 = log of the focal scale. Returns 0, or TCV_BA_ERR_MEMORY / TCV_BA_ERR_ARGUMENT.
 */
int tcv_oracle_dense_solve(const tcv_ba_control* ctl, const double* scene_in, const uint32_t* index,
                           double* out) {
  const uint32_t C = ctl->cameras, P = ctl->points, M = ctl->observations;
  const int shared = ctl->shared_focal ? 1 : 0;
  const uint32_t* point_start = index + 2 * (size_t)M;
  const uint32_t* active = index + 3 * (size_t)M + P + C + 2;
  int32_t* row_of = malloc(C * sizeof(int32_t));
  uint32_t n_active = 0, fixed = 0;
  for (uint32_t c = 0; c < C; c++) {
    row_of[c] = active[c] ? (int32_t)(6 * n_active++) : -1;
    if (!active[c]) fixed++;
  }
  const uint32_t focal_row = 6 * n_active, n = 6 * n_active + shared;
  if ((uint64_t)n * n > 16ull * 1024 * 1024) {
    free(row_of);
    return TCV_BA_ERR_ARGUMENT;
  }
  const size_t scene_size = (size_t)C * 18 + (size_t)P * 3 + (size_t)M * 2, coords = (size_t)C * 12 + (size_t)P * 3;
  double* model = malloc(scene_size * sizeof(double));
  double* cand = malloc(scene_size * sizeof(double));
  double* A = malloc((size_t)n * n * sizeof(double));
  double* rhs = malloc(n * sizeof(double));
  double* step = malloc(n * sizeof(double));
  double* Vinv = malloc((size_t)P * 9 * sizeof(double));
  double* bp = malloc((size_t)P * 3 * sizeof(double));
  double* Wobs = malloc((size_t)M * 18 * sizeof(double)); /* camera rows x point, per observation */
  double* Wf = malloc((size_t)P * 3 * sizeof(double));    /* focal row x point, per point */
  if (!model || !cand || !A || !rhs || !step || !Vinv || !bp || !Wobs || !Wf) {
    free(row_of); free(model); free(cand); free(A); free(rhs); free(step); free(Vinv); free(bp);
    free(Wobs); free(Wf);
    return TCV_BA_ERR_MEMORY;
  }
  memcpy(model, scene_in, scene_size * sizeof(double));
  double origin[3], scale_center[3];
  {
    const double* a = model + (size_t)ctl->anchor * 12;
    const double* s = model + (size_t)ctl->scale_view * 12;
    for (int i = 0; i < 3; i++) {
      origin[i] = -(a[i] * a[9] + a[3 + i] * a[10] + a[6 + i] * a[11]);
      scale_center[i] = -(s[i] * s[9] + s[3 + i] * s[10] + s[6 + i] * s[11]);
    }
  }
  const double baseline = hypot(hypot(origin[0] - scale_center[0], origin[1] - scale_center[1]),
                                origin[2] - scale_center[2]);
  double fscale = 1, log_focal = 0;
  double cost0 = 0, sq0 = 0;
  if (!dense_metrics(ctl, model, index, fscale, &cost0, &sq0)) {
    free(row_of); free(model); free(cand); free(A); free(rhs); free(step); free(Vinv); free(bp);
    free(Wobs); free(Wf);
    return TCV_BA_ERR_ARGUMENT;
  }
  double cost = cost0, sq = sq0, lambda = 1e-3;
  uint32_t accepted = 0, rejected = 0;
  const double* pixels = model + (size_t)C * 18 + (size_t)P * 3;
  for (uint32_t it = 0; it < ctl->iterations; it++) {
    memset(A, 0, (size_t)n * n * sizeof(double));
    memset(rhs, 0, n * sizeof(double));
    int bad_point = 0;
    for (uint32_t p = 0; p < P; p++) {
      double V[9] = {0}, b[3] = {0}, wf[3] = {0};
      const int dir = p >= P - ctl->directional_points;
      const double* X = model + (size_t)C * 12 + (size_t)p * 3;
      for (uint32_t i = point_start[p]; i < point_start[p + 1]; i++) {
        const uint32_t c = index[i];
        const double* pose = model + (size_t)c * 12;
        const double* mo = model + (size_t)C * 12 + (size_t)P * 3 + (size_t)c * 6;
        Model k = {mo[0] * fscale, mo[1] * fscale, mo[2], mo[3], mo[4], mo[5]};
        double q[3];
        for (int r = 0; r < 3; r++)
          q[r] = pose[r * 3] * X[0] + pose[r * 3 + 1] * X[1] + pose[r * 3 + 2] * X[2] + (dir ? 0 : pose[9 + r]);
        Projection pr = project(&k, q);
        double jp[6], jc[12];
        for (int a = 0; a < 3; a++) {
          jp[a] = pr.du[0] * pose[a] + pr.du[1] * pose[3 + a] + pr.du[2] * pose[6 + a];
          jp[3 + a] = pr.dv[0] * pose[a] + pr.dv[1] * pose[3 + a] + pr.dv[2] * pose[6 + a];
        }
        pose_row(q, pr.du, jc);
        pose_row(q, pr.dv, jc + 6);
        if (dir)
          for (int a = 3; a < 6; a++) jc[a] = jc[6 + a] = 0;
        const double r[2] = {pr.px - pixels[2 * (size_t)i], pr.py - pixels[2 * (size_t)i + 1]};
        const double e = hypot(r[0], r[1]), w = e <= ctl->huber ? 1 : ctl->huber / e;
        const double jf[2] = {pr.px - k.cx, pr.py - k.cy};
        const int base = row_of[c];
        double* W = Wobs + (size_t)i * 18;
        for (int a = 0; a < 18; a++) W[a] = 0;
        for (int axis = 0; axis < 2; axis++) {
          for (int a = 0; a < 3; a++) {
            b[a] -= w * jp[axis * 3 + a] * r[axis];
            for (int l = 0; l < 3; l++) V[a * 3 + l] += w * jp[axis * 3 + a] * jp[axis * 3 + l];
          }
          /* the observation's rows: its camera's 6 (when variable) and the focal row */
          int rows[7], nr = 0;
          double J[7];
          if (base >= 0)
            for (int a = 0; a < 6; a++) {
              rows[nr] = base + a;
              J[nr++] = jc[axis * 6 + a];
            }
          if (shared) {
            rows[nr] = (int)focal_row;
            J[nr++] = jf[axis];
          }
          for (int a = 0; a < nr; a++) {
            rhs[rows[a]] -= w * J[a] * r[axis];
            for (int l = 0; l < nr; l++) A[(size_t)rows[a] * n + rows[l]] += w * J[a] * J[l];
            for (int l = 0; l < 3; l++) {
              const double v = w * J[a] * jp[axis * 3 + l];
              if (rows[a] == (int)focal_row && shared) wf[l] += v;
              else W[(rows[a] - base) * 3 + l] += v;
            }
          }
        }
      }
      for (int a = 0; a < 3; a++) V[a * 3 + a] += lambda * fmax(V[a * 3 + a], 1e-10);
      if (!adjugate_inverse3(V, Vinv + (size_t)p * 9)) bad_point = 1;
      for (int a = 0; a < 3; a++) {
        bp[(size_t)p * 3 + a] = b[a];
        Wf[(size_t)p * 3 + a] = wf[a];
      }
    }
    if (bad_point) {
      lambda *= 10;
      rejected++;
      continue;
    }
    for (uint32_t a = 0; a < n; a++) A[(size_t)a * n + a] += lambda * fmax(A[(size_t)a * n + a], 1e-10);
    /* S = A - W V^-1 W^T, reduced rhs = b - W V^-1 b_p, point by point */
    for (uint32_t p = 0; p < P; p++) {
      const double* inv = Vinv + (size_t)p * 9;
      const uint32_t i0 = point_start[p], i1 = point_start[p + 1];
      /* the point's rows and their cross vectors */
      const uint32_t nrows = 6 * (i1 - i0) + (uint32_t)shared;
      int* rows = malloc(nrows * sizeof(int));
      double* Wr = malloc((size_t)nrows * 3 * sizeof(double));
      double* Tr = malloc((size_t)nrows * 3 * sizeof(double));
      uint32_t cnt = 0;
      for (uint32_t i = i0; i < i1; i++) {
        const int base = row_of[index[i]];
        if (base < 0) continue;
        for (int a = 0; a < 6; a++) {
          rows[cnt] = base + a;
          for (int l = 0; l < 3; l++) Wr[cnt * 3 + l] = Wobs[(size_t)i * 18 + a * 3 + l];
          cnt++;
        }
      }
      if (shared) {
        rows[cnt] = (int)focal_row;
        for (int l = 0; l < 3; l++) Wr[cnt * 3 + l] = Wf[(size_t)p * 3 + l];
        cnt++;
      }
      for (uint32_t a = 0; a < cnt; a++) {
        for (int l = 0; l < 3; l++) {
          double t = 0;
          for (int k = 0; k < 3; k++) t += Wr[a * 3 + k] * inv[k * 3 + l];
          Tr[a * 3 + l] = t;
        }
        for (int l = 0; l < 3; l++) rhs[rows[a]] -= Tr[a * 3 + l] * bp[(size_t)p * 3 + l];
      }
      for (uint32_t a = 0; a < cnt; a++)
        for (uint32_t l = 0; l < cnt; l++) {
          double v = 0;
          for (int k = 0; k < 3; k++) v += Tr[a * 3 + k] * Wr[l * 3 + k];
          A[(size_t)rows[a] * n + rows[l]] -= v;
        }
      free(rows);
      free(Wr);
      free(Tr);
    }
    if (!solve_system(A, rhs, n, step)) {
      lambda *= 10;
      rejected++;
      continue;
    }
    const double new_scale = shared ? exp(step[focal_row]) : 1.0;
    if (!(isfinite(new_scale) && new_scale > 0)) {
      lambda *= 10;
      rejected++;
      continue;
    }
    memcpy(cand, model, scene_size * sizeof(double));
    for (uint32_t c = 0; c < C; c++)
      if (row_of[c] >= 0) move_pose(model + (size_t)c * 12, step + row_of[c], cand + (size_t)c * 12);
    for (uint32_t p = 0; p < P; p++) {
      double b[3] = {bp[(size_t)p * 3], bp[(size_t)p * 3 + 1], bp[(size_t)p * 3 + 2]};
      for (uint32_t i = point_start[p]; i < point_start[p + 1]; i++) {
        const int base = row_of[index[i]];
        if (base < 0) continue;
        for (int j = 0; j < 3; j++)
          for (int a = 0; a < 6; a++) b[j] -= Wobs[(size_t)i * 18 + a * 3 + j] * step[base + a];
      }
      if (shared)
        for (int j = 0; j < 3; j++) b[j] -= Wf[(size_t)p * 3 + j] * step[focal_row];
      const double* inv = Vinv + (size_t)p * 9;
      for (int j = 0; j < 3; j++)
        cand[(size_t)C * 12 + (size_t)p * 3 + j] +=
            inv[j * 3] * b[0] + inv[j * 3 + 1] * b[1] + inv[j * 3 + 2] * b[2];
    }
    const double* sv = cand + (size_t)ctl->scale_view * 12;
    double cs[3];
    for (int i = 0; i < 3; i++) cs[i] = -(sv[i] * sv[9] + sv[3 + i] * sv[10] + sv[6 + i] * sv[11]);
    const double nb = hypot(hypot(origin[0] - cs[0], origin[1] - cs[1]), origin[2] - cs[2]);
    if (!(nb > 1e-12)) {
      lambda *= 10;
      rejected++;
      continue;
    }
    const double s = fixed == 1 ? baseline / nb : 1.0;
    for (uint32_t c = 0; c < C; c++) {
      if (c == ctl->anchor) continue;
      double* pc = cand + (size_t)c * 12;
      for (int i = 0; i < 3; i++) {
        const double ro = pc[i * 3] * origin[0] + pc[i * 3 + 1] * origin[1] + pc[i * 3 + 2] * origin[2];
        pc[9 + i] = s * pc[9 + i] + (s - 1) * ro;
      }
    }
    for (uint32_t p = 0; p < P; p++) {
      double* X = cand + (size_t)C * 12 + (size_t)p * 3;
      if (p >= P - ctl->directional_points) {
        const double nx = hypot(hypot(X[0], X[1]), X[2]);
        for (int j = 0; j < 3; j++) X[j] /= nx;
      } else {
        for (int j = 0; j < 3; j++) X[j] = origin[j] + s * (X[j] - origin[j]);
      }
    }
    double next_cost, next_sq;
    const int ok = dense_metrics(ctl, cand, index, fscale * new_scale, &next_cost, &next_sq);
    if (ok && next_cost < cost) {
      const double improvement = cost - next_cost;
      memcpy(model, cand, coords * sizeof(double));
      fscale *= new_scale;
      if (shared) log_focal += step[focal_row];
      cost = next_cost;
      sq = next_sq;
      lambda = fmax(1e-10, lambda / 3);
      accepted++;
      if (improvement < ctl->stall_tolerance * fmax(1, cost)) break;
    } else {
      lambda *= 10;
      rejected++;
      if (lambda > 1e12) break;
    }
  }
  memcpy(out, model, coords * sizeof(double));
  for (int i = 0; i < 26; i++) out[coords + i] = 0;
  out[coords + 2] = baseline;
  out[coords + 6] = cost;
  out[coords + 7] = sq;
  out[coords + 8] = cost0;
  out[coords + 9] = sq0;
  out[coords + 10] = accepted;
  out[coords + 11] = rejected;
  out[coords + 24] = sqrt(sq0 / M);
  out[coords + 25] = sqrt(sq / M);
  if (shared) out[coords + 26] = log_focal;
  free(row_of); free(model); free(cand); free(A); free(rhs); free(step); free(Vinv); free(bp);
  free(Wobs); free(Wf);
  return 0;
}
