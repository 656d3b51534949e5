// obs_eval.hip — fixed-scene observation evaluation: the reference's evaluateObservationsGpu /
// evaluateContinuedObservationsGpu (src/geometry/gpu/observations.ts:34-127) on the device. Every
// observation is projected through its view's pose and camera record, its pixel error taken and
// compared with the threshold; guarded runs then drop every track whose depth is not stable under
// pixel noise; the inliers' errors and count are summed.
//
//   k_obs_project    one thread per observation (projectIndexedObservations,
//                    observation-evaluation.shade.ts:44-54)
//   k_obs_depth      one thread per track (filterIndexedObservationDepth :85-97 with stableDepth
//                    :57-83): the 3x3 depth information of the track's accepted observations and,
//                    per observation, the relative depth standard deviation along its optical axis
//   k_obs_partials   one thread per 64 observations, errors and counts summed in lane order (the
//                    CPU tier's indexedObservationPartialsOracle :128-137)
//   k_obs_summary    one workgroup, the error partials in block order (observationSummary,
//                    triangulation.shade.ts:134-143)
//
// The arithmetic is the reference's expression by expression with no fused multiply-adds (the
// pragma below, as the reference's CPU tier computes in JavaScript), so the flags, errors and sums
// equal the CPU tier's. The projection and the 3x3 solve are restated here, under that pragma,
// rather than taken from ba_math.h, whose BA kernels keep their contractions.
#pragma clang fp contract(off)

#include "ba_math.h"

namespace tcvba {

struct ObsProblem {
  const double* cameras;   // 18 per view: pose record (R row-major, t), fx fy cx cy k1 k2
  const double* points;    // 3 per landmark
  const double* pixels;    // 2 per observation
  const uint32_t* indices; // view, landmark per observation
  const uint32_t* offsets; // tracks + 1 first observations (guarded runs)
  uint32_t* flags;         // count, then tracks (guarded)
  double* errors;          // count
  double* partials;        // 2 per 64 observations
  double* summary;         // error sum, inlier count
  uint32_t count, tracks, blocks;
  double max_error, max_relative_depth_std, pixel_noise_std;
};

// radialLimit, numerics.shade.ts:79-99
__device__ double obs_radial_limit(const Model& k) {
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

// project, numerics.shade.ts:141-173
__device__ bool obs_project(const Model& k, const double q[3], double& px, double& py,
                            double du[3], double dv[3]) {
  if (!(q[2] > 1e-12)) return false;
  const double x = q[0] / q[2], y = q[1] / q[2], r2 = x * x + y * y;
  if (!(r2 < obs_radial_limit(k) && r2 < 1e8)) return false;
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

// solve3, numerics.shade.ts:25-45
__device__ bool obs_solve3(const double a[9], const double b[3], double x[3]) {
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

// transform, numerics.shade.ts:174-180: q = R X + t for a pose record.
__device__ __forceinline__ void obs_transform(const double* p, const double* X, double q[3]) {
  q[0] = p[0] * X[0] + p[1] * X[1] + p[2] * X[2] + p[9];
  q[1] = p[3] * X[0] + p[4] * X[1] + p[5] * X[2] + p[10];
  q[2] = p[6] * X[0] + p[7] * X[1] + p[8] * X[2] + p[11];
}

__device__ __forceinline__ Model obs_model(const double* rec) {
  return Model{rec[12], rec[13], rec[14], rec[15], rec[16], rec[17]};
}

__global__ __launch_bounds__(256) void k_obs_project(ObsProblem o) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= o.count) return;
  uint32_t flag = 0;
  double error = 0;
  const double* rec = o.cameras + (size_t)o.indices[2 * (size_t)i] * 18;
  const double* X = o.points + 3 * (size_t)o.indices[2 * (size_t)i + 1];
  double q[3], px, py, du[3], dv[3];
  obs_transform(rec, X, q);
  if (obs_project(obs_model(rec), q, px, py, du, dv)) {
    const double x = px - o.pixels[2 * (size_t)i], y = py - o.pixels[2 * (size_t)i + 1];
    const double e = sqrt(x * x + y * y);
    if (e <= o.max_error) {
      flag = 1;
      error = e;
    }
  }
  o.flags[i] = flag;
  o.errors[i] = error;
}

// stableDepth (observation-evaluation.shade.ts:57-83) of the track [start, end) at point X.
__device__ bool obs_stable_depth(const ObsProblem& o, uint32_t start, uint32_t end,
                                 const double* X) {
  double A[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  uint32_t count = 0;
  for (uint32_t i = start; i < end; i++) {
    if (o.flags[i] == 0) continue;
    count++;
    const double* p = o.cameras + (size_t)o.indices[2 * (size_t)i] * 18;
    double q[3], px, py, du[3], dv[3];
    obs_transform(p, X, q);
    if (!obs_project(obs_model(p), q, px, py, du, dv)) return false;
    for (int row = 0; row < 2; row++) {
      const double* g = row == 0 ? du : dv;
      const double J[3] = {g[0] * p[0] + g[1] * p[3] + g[2] * p[6],
                           g[0] * p[1] + g[1] * p[4] + g[2] * p[7],
                           g[0] * p[2] + g[1] * p[5] + g[2] * p[8]};
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) A[3 * r + c] += J[r] * J[c];
    }
  }
  if (count < 2) return false;
  for (uint32_t i = start; i < end; i++) {
    if (o.flags[i] == 0) continue;
    const double* p = o.cameras + (size_t)o.indices[2 * (size_t)i] * 18;
    const double axis[3] = {p[6], p[7], p[8]};
    double q[3], solved[3];
    obs_transform(p, X, q);
    if (!obs_solve3(A, axis, solved) || !(q[2] > 0)) return false;
    const double variance = axis[0] * solved[0] + axis[1] * solved[1] + axis[2] * solved[2];
    if (!(variance >= 0)) return false;
    const double relative = o.pixel_noise_std * sqrt(variance) / q[2];
    if (!(relative <= o.max_relative_depth_std)) return false;
  }
  return true;
}

__global__ __launch_bounds__(256) void k_obs_depth(ObsProblem o) {
  const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= o.tracks) return;
  const uint32_t start = o.offsets[h], end = o.offsets[h + 1];
  o.flags[o.count + h] = 0;
  if (end <= start) return;
  uint32_t accepted = 0;
  for (uint32_t i = start; i < end; i++) accepted += o.flags[i] != 0 ? 1u : 0u;
  if (accepted < 2) return;
  const double* X = o.points + 3 * (size_t)o.indices[2 * (size_t)start + 1];
  if (!obs_stable_depth(o, start, end, X)) {
    o.flags[o.count + h] = 1;
    for (uint32_t i = start; i < end; i++) {
      o.flags[i] = 0;
      o.errors[i] = 0;
    }
  }
}

__global__ __launch_bounds__(256) void k_obs_partials(ObsProblem o) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= o.blocks) return;
  double e = 0, n = 0;
  for (uint32_t lane = 0; lane < 64; lane++) {
    const uint32_t i = b * 64 + lane;
    if (i >= o.count || o.flags[i] == 0) continue;
    e += o.errors[i];
    n += 1;
  }
  o.partials[2 * (size_t)b] = e;
  o.partials[2 * (size_t)b + 1] = n;
}

// One workgroup. The error sum is the sequential one, block by block: wave 0 adds the blocks'
// error partials in block order, one dependent addition after another, reading them from LDS 16 at
// a time, while the workgroup stages the next chunk of SUMMARY_CHUNK partials into the other LDS
// buffer. The counts are whole numbers, exact in any order: every thread adds those of the blocks
// it stages, and the workgroup sums them at the end.
constexpr uint32_t SUMMARY_THREADS = 1024;
constexpr uint32_t SUMMARY_CHUNK = 2048;

__global__ __launch_bounds__(SUMMARY_THREADS) void k_obs_summary(ObsProblem o) {
  __shared__ double buf[2][SUMMARY_CHUNK];
  __shared__ double red[SUMMARY_THREADS / 64];
  const uint32_t tid = threadIdx.x;
  const double* p = o.partials;
  double e = 0, n = 0;
  auto stage = [&](uint32_t chunk) {
    const uint32_t first = chunk * SUMMARY_CHUNK;
    double* to = buf[chunk & 1];
    for (uint32_t k = tid; k < SUMMARY_CHUNK && first + k < o.blocks; k += SUMMARY_THREADS) {
      to[k] = p[2 * (size_t)(first + k)];
      n += p[2 * (size_t)(first + k) + 1];
    }
  };
  const uint32_t chunks = (o.blocks + SUMMARY_CHUNK - 1) / SUMMARY_CHUNK;
  if (chunks) stage(0);
  __syncthreads();
  for (uint32_t c = 0; c < chunks; c++) {
    if (c + 1 < chunks) stage(c + 1);
    if (tid < 64) {
      const double* x = buf[c & 1];
      const uint32_t len = min(SUMMARY_CHUNK, o.blocks - c * SUMMARY_CHUNK);
      uint32_t k = 0;
      for (; k + 16 <= len; k += 16) {
        double v[16];
#pragma unroll
        for (uint32_t u = 0; u < 16; u++) v[u] = x[k + u];
#pragma unroll
        for (uint32_t u = 0; u < 16; u++) e += v[u];
      }
      for (; k < len; k++) e += x[k];
    }
    __syncthreads();
  }
  for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off, 64);
  if ((tid & 63) == 0) red[tid >> 6] = n;
  __syncthreads();
  if (tid == 0) {
    double total = 0;
    for (uint32_t w = 0; w < SUMMARY_THREADS / 64; w++) total += red[w];
    o.summary[0] = e;
    o.summary[1] = total;
  }
}

// The index checks of the allocation-free entry (tcv_obs_evaluate_noalloc), whose arrays are
// device arrays: every observation's view and landmark in range, and on guarded runs the track
// offsets ascending from 0 to the observation count. Sets bit 1 (indices) or 2 (offsets) of *bad.
__global__ __launch_bounds__(256) void k_obs_check(ObsProblem o, uint32_t views, uint32_t points,
                                                   uint32_t* bad) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < o.count && (o.indices[2 * (size_t)i] >= views || o.indices[2 * (size_t)i + 1] >= points))
    atomicOr(bad, 1u);
  if (i < o.tracks) {
    const uint32_t a = o.offsets[i], b = o.offsets[i + 1];
    if (b < a || (i == 0 && a != 0) || (i + 1 == o.tracks && b != o.count)) atomicOr(bad, 2u);
  }
}

// n words set to zero (the allocation-free entry clears its scratch words with a kernel, which a
// captured graph replays correctly).
__global__ void k_obs_zero(uint32_t* words, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) words[i] = 0;
}
__global__ void k_obs_zero_f64(double* words, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) words[i] = 0;
}

}  // namespace tcvba

// back to the engine's contraction (-ffp-contract=on: within an expression only)
#pragma clang fp contract(on)
