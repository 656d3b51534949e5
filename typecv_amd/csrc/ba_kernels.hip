// ba_kernels.hip — gfx950 kernels of the bundle-adjustment LM loop (included by ba_engine.hip).
//
// Kernel map (reference entry points in src/sfm/bundle/gpu/bundle.shade.ts "b:" and
// joint-focal.shade.ts "j:"). "Per bundle" kernels run one workgroup per bundle of the batch.
//   k_check_points / k_check_tasks   index validation, camera-major point ids (host packing checks)
//   k_init_bundle (per bundle)       initializeBundle b:235-250 / j:269-286
//   k_cost / k_finish_cost           bundleMetricPartials + finishBundleMetrics b:319-377,
//                                    bundleCandidatePartials + acceptBundleCandidate b:1040-1096
//   k_point_pass                     linearizeBundle + bundlePointBlocks b:251-297, 378-422 and the
//                                    focal point terms of bundleFocalPoints j:1181-1204, fused:
//                                    one wave per tile of whole points, one lane per observation
//   k_camera_pass / k_camera_finish  bundleCameraBlocks / bundleCameraRows b:423-555 with
//                                    invertBundleCameraColumns b:655-694 and the focal camera terms
//                                    bundleFocalCameraRows/Observations j:1138-1171: one workgroup
//                                    per camera-list task, fixed-order task sums, Cholesky inverse
//   k_init_cg (per bundle)           initializeBundleCg b:695-741 / j:646-700 with dampBundleFocal
//   k_point_sweep<SCHUR>             bundleSchurPoint b:742-745 (body b:185-203, j:198-218)
//   k_camera_sweep                   bundleSchurCamera* b:746-882 / j:705-870 (camera partials)
//   k_point_sweep_fused + k_camera_tsum   the same product with the cross blocks read once (a row
//                                    of W y per observation); the same sums, the same bits
//   k_cg_product / k_cg_step / k_cg_direction   finish S·p, then advanceBundleCg b:883-929
//   k_update_cameras / k_point_sweep<UPDATE>   updateBundleCandidate b:930-961 / j:925-960
//   k_gauge (per bundle)             bundleCandidateGauge b:962-993 / j:961-1000
//   k_normalize                      normalizeBundleCandidate b:994-1023
//   k_commit / k_pack                commitBundleCandidate b:1109-1114, packBundleResult b:1115-1125
#include "ba_kernels.h"

namespace tcvba {

// ------------------------------------------------------------------ small helpers

__device__ __forceinline__ double* bstate(const Problem& q, uint32_t b) {
  return q.state + (size_t)b * S_WORDS;
}
__device__ __forceinline__ bool variable_camera(const Problem& q, uint32_t c) {
  return camera_active(q)[c] != 0;
}
// The passes store the bundle's cross blocks, and the stored-block sweeps cover it.
__device__ __forceinline__ bool stores_w(const Bundle& g) { return g.unstored == 0; }
__device__ __forceinline__ bool directional_point(const Bundle& g, uint32_t p) {
  return p >= g.first_directional;
}

// Model of camera c at focal scale `fscale` (= focalScale(log), 1 on the fixed route).
__device__ __forceinline__ Model camera_model(const Problem& q, uint32_t c, double fscale) {
  const double* s = q.scene + (size_t)q.C * 12 + (size_t)q.P * 3 + (size_t)c * 6;
  Model k{s[0], s[1], s[2], s[3], s[4], s[5]};
  if (q.shared) {
    k.fx = s[0] * fscale;
    k.fy = s[1] * fscale;
  }
  return k;
}

// The sum of v over the workgroup, identical in every thread (butterfly within each wave, then
// the wave sums in wave order). `lds` holds one double per wave. Call from uniform control flow.
__device__ __forceinline__ double block_sum(double v, double* lds) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, waves = blockDim.x >> 6;
  __syncthreads();
  if (lane == 0) lds[wave] = v;
  __syncthreads();
  double s = 0;
  for (int w = 0; w < waves; w++) s += lds[w];
  return s;
}

// block_sum of two values with one LDS round: the same sums, in the same order, as two block_sum
// calls. `lds` holds two doubles per wave.
__device__ __forceinline__ void block_sum2(double& a, double& b, double* lds) {
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_xor(a, o, 64);
    b += __shfl_xor(b, o, 64);
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, waves = blockDim.x >> 6;
  __syncthreads();
  if (lane == 0) {
    lds[wave] = a;
    lds[waves + wave] = b;
  }
  __syncthreads();
  double sa = 0, sb = 0;
  for (int w = 0; w < waves; w++) sa += lds[w];
  for (int w = 0; w < waves; w++) sb += lds[waves + w];
  a = sa;
  b = sb;
}

// The workgroup's logical index such that each of the 8 XCDs (which receive workgroups round-robin
// in dispatch order) runs one contiguous range of the grid: workgroups of neighbouring cameras then
// share that XCD's L2, where the vectors they gather (y by point in the camera sweep, a bundle's G
// rows in the dense assembly) stay resident.
__device__ __forceinline__ uint32_t xcd_contiguous(uint32_t i, uint32_t n) {
  const uint32_t q = n / 8, r = n % 8, x = i % 8, k = i / 8;
  return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

// The wave tile of a wave of a point-tile kernel (TILE_WAVES tiles per workgroup).
__device__ __forceinline__ uint32_t point_tile(uint32_t wave) {
  return xcd_contiguous(blockIdx.x, gridDim.x) * TILE_WAVES + wave;
}

// Marks a loaded value as used here: the load is issued (and waited for) before what follows, so a
// kernel's array loads go out together with its state-word loads instead of after the test of
// those words. An empty asm statement: no instruction is emitted.
__device__ __forceinline__ void settle(double& v) { asm volatile("" : "+v"(v)); }
__device__ __forceinline__ void settle(uint32_t& v) { asm volatile("" : "+v"(v)); }
// The same for a value that is uniform across the wave (kept in scalar registers).
__device__ __forceinline__ void settle_uniform(double& v) { asm volatile("" : "+s"(v)); }

// Orders one wave's LDS writes before its later LDS reads. The point-tile kernels give every wave
// its own LDS slice and their loops have per-wave trip counts, so they synchronise per wave, never
// per workgroup.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A cross-block word: streamed once per sweep, so it is loaded non-temporally to leave the caches
// to the gathered vectors.
template <class WT>
__device__ __forceinline__ double wload(const WT* p) {
  return (double)__builtin_nontemporal_load(p);
}

// Linearisation of one observation (linearizeBundle b:251-297 / j:287-336).
struct Lin {
  double Jc[12];  // d(u,v)/d(omega, v): row u then row v (zero for a fixed camera)
  double Jp[6];   // d(u,v)/dX: row u then row v
  double r[2];    // projected - observed pixel
  double w;       // Huber weight
  double Jf[2];   // d(u,v)/d(log f) = pixel - principal point
  // the intermediates the exact Hessian reuses (dead, and removed, where unused)
  double qc[3];   // camera-frame point
  double du[3], dv[3];  // d(u,v)/dq
  double R[9];    // rotation, row-major
  Model k;        // camera model at the focal scale
};

// The camera-frame point q = R X + t of a pose (R row-major, then t) and a point X; R X for a
// directional point. The one expression of q: the FP64 form's passes store factors without it and
// its products recompute it, the same bits (the tests of streamed against resident solves hold
// the two to bit equality).
__device__ __forceinline__ void camera_point(const double* pose, double x0, double x1, double x2,
                                             bool dir, double* qc) {
#pragma unroll
  for (int r = 0; r < 3; r++) {
    qc[r] = pose[r * 3] * x0 + pose[r * 3 + 1] * x1 + pose[r * 3 + 2] * x2;
    if (!dir) qc[r] += pose[9 + r];
  }
}

__device__ __forceinline__ bool linearize(const Problem& q, uint32_t c, uint32_t p, uint32_t i,
                                          double fscale, bool var, bool dir, Lin& L) {
  const double* pose = q.scene + (size_t)c * 12;
  const double* X = q.scene + (size_t)q.C * 12 + (size_t)p * 3;
  double R[12];
#pragma unroll
  for (int j = 0; j < 12; j++) R[j] = pose[j];
  double qc[3];
  camera_point(R, X[0], X[1], X[2], dir, qc);
  const Model k = camera_model(q, c, fscale);
  double px, py, du[3], dv[3];
  if (!project(k, qc, px, py, du, dv)) return false;
  const double* obs = q.scene + (size_t)q.C * 18 + (size_t)q.P * 3 + (size_t)i * 2;
  const double ru = px - obs[0], rv = py - obs[1];
  const double squared = ru * ru + rv * rv, error = sqrt(squared);
  L.w = error > q.huber ? q.huber / error : 1.0;
  L.r[0] = ru;
  L.r[1] = rv;
  if (var) {
    // poseRow, numerics.shade.ts:181-183; directional points have no translation columns
    L.Jc[0] = du[2] * qc[1] - du[1] * qc[2];
    L.Jc[1] = du[0] * qc[2] - du[2] * qc[0];
    L.Jc[2] = du[1] * qc[0] - du[0] * qc[1];
    L.Jc[6] = dv[2] * qc[1] - dv[1] * qc[2];
    L.Jc[7] = dv[0] * qc[2] - dv[2] * qc[0];
    L.Jc[8] = dv[1] * qc[0] - dv[0] * qc[1];
    L.Jc[3] = dir ? 0.0 : du[0];
    L.Jc[4] = dir ? 0.0 : du[1];
    L.Jc[5] = dir ? 0.0 : du[2];
    L.Jc[9] = dir ? 0.0 : dv[0];
    L.Jc[10] = dir ? 0.0 : dv[1];
    L.Jc[11] = dir ? 0.0 : dv[2];
  } else {
#pragma unroll
    for (int j = 0; j < 12; j++) L.Jc[j] = 0;
  }
#pragma unroll
  for (int j = 0; j < 3; j++) {
    L.Jp[j] = du[0] * R[j] + du[1] * R[3 + j] + du[2] * R[6 + j];
    L.Jp[3 + j] = dv[0] * R[j] + dv[1] * R[3 + j] + dv[2] * R[6 + j];
  }
  L.Jf[0] = px - k.cx;
  L.Jf[1] = py - k.cy;
#pragma unroll
  for (int j = 0; j < 3; j++) {
    L.qc[j] = qc[j];
    L.du[j] = du[j];
    L.dv[j] = dv[j];
  }
#pragma unroll
  for (int j = 0; j < 9; j++) L.R[j] = R[j];
  L.k = k;
  return true;
}

// Packed upper triangle of a symmetric 10x10 block over (omega 3, v 3, X 3, log f).
__host__ __device__ constexpr int up10(int j, int l) { return j * 10 - j * (j - 1) / 2 + (l - j); }

// The exact Hessian of one observation's half robust cost minus its Gauss-Newton block w J^T J:
//   C = J^T (D - w I) J + sum_k rho_k d2(pixel_k)/dtheta2,   rho = w r,
// D = d(w r)/dr (I inside the Huber radius, w (I - r r^T / |r|^2) outside), in the left-perturbation
// chart q = exp(omega)(R X + t) + v of movePose (q = exp(omega) R X for a directional point). Its
// parts: -w a a^T with a = J^T r/|r| outside the radius; Q^T G Q with Q = dq/dtheta and G the
// weighted curvature of the projection; g . d2q/dtheta2 with g = rho_u du + rho_v dv, i.e.
// 1/2 (g q^T + q g^T) - (g . q) I on omega-omega and (R_j x g) on omega-X_j; and rho . J on the
// focal row and column (d pixel/d log f = pixel - c). Rows of a fixed camera stay zero.
// `curvature` false (the Gauss-Newton adjoint): the Huber term -w a a^T only.
__host__ __device__ void exact_correction(const Lin& L, bool var, bool dir, double huber, double C[55],
                                          bool curvature = true) {
  for (int k = 0; k < 55; k++) C[k] = 0;
  double J[2][10];
  for (int k = 0; k < 6; k++) {
    J[0][k] = L.Jc[k];
    J[1][k] = L.Jc[6 + k];
  }
  for (int k = 0; k < 3; k++) {
    J[0][6 + k] = L.Jp[k];
    J[1][6 + k] = L.Jp[3 + k];
  }
  J[0][9] = L.Jf[0];
  J[1][9] = L.Jf[1];
  const double rho0 = L.w * L.r[0], rho1 = L.w * L.r[1];
  const double e = sqrt(L.r[0] * L.r[0] + L.r[1] * L.r[1]);
  if (e > huber) {
    const double h0 = L.r[0] / e, h1 = L.r[1] / e;
    double a[10];
    for (int k = 0; k < 10; k++) a[k] = J[0][k] * h0 + J[1][k] * h1;
    for (int j = 0; j < 10; j++)
      for (int l = j; l < 10; l++) C[up10(j, l)] -= L.w * a[j] * a[l];
  }
  if (!curvature) return;
  double G[9];
  projection_curvature(L.k, L.qc, rho0, rho1, G);
  // Q = dq/dtheta over omega, v, X (columns)
  double Q[3][9];
  const double* q = L.qc;
  const double cross[3][3] = {{0, -q[2], q[1]}, {q[2], 0, -q[0]}, {-q[1], q[0], 0}};  // e_a x q
  for (int a = 0; a < 3; a++)
    for (int m = 0; m < 3; m++) {
      Q[m][a] = var ? cross[a][m] : 0.0;
      Q[m][3 + a] = (var && !dir && m == a) ? 1.0 : 0.0;
      Q[m][6 + a] = L.R[m * 3 + a];
    }
  double GQ[3][9];
  for (int m = 0; m < 3; m++)
    for (int j = 0; j < 9; j++) GQ[m][j] = G[m * 3] * Q[0][j] + G[m * 3 + 1] * Q[1][j] + G[m * 3 + 2] * Q[2][j];
  for (int i = 0; i < 9; i++)
    for (int j = i; j < 9; j++) C[up10(i, j)] += Q[0][i] * GQ[0][j] + Q[1][i] * GQ[1][j] + Q[2][i] * GQ[2][j];
  const double g[3] = {rho0 * L.du[0] + rho1 * L.dv[0], rho0 * L.du[1] + rho1 * L.dv[1],
                       rho0 * L.du[2] + rho1 * L.dv[2]};
  if (var) {
    const double gq = g[0] * q[0] + g[1] * q[1] + g[2] * q[2];
    for (int a = 0; a < 3; a++)
      for (int b = a; b < 3; b++)
        C[up10(a, b)] += 0.5 * (g[a] * q[b] + q[a] * g[b]) - (a == b ? gq : 0.0);
    for (int j = 0; j < 3; j++) {
      const double Rj[3] = {L.R[j], L.R[3 + j], L.R[6 + j]};
      const double x[3] = {Rj[1] * g[2] - Rj[2] * g[1], Rj[2] * g[0] - Rj[0] * g[2],
                           Rj[0] * g[1] - Rj[1] * g[0]};
      for (int a = 0; a < 3; a++) C[up10(a, 6 + j)] += x[a];
    }
  }
  for (int k = 0; k < 9; k++) C[up10(k, 9)] += rho0 * J[0][k] + rho1 * J[1][k];
  C[up10(9, 9)] += rho0 * J[0][9] + rho1 * J[1][9];
}

// Cross block W = w Jc^T Jp (18 words, [k*3+j]) and the focal row w Jf^T Jp (3 words).
__device__ __forceinline__ void cross_block(const Lin& L, bool shared, double W[21]) {
#pragma unroll
  for (int k = 0; k < 6; k++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      W[k * 3 + j] = L.w * L.Jc[k] * L.Jp[j] + L.w * L.Jc[6 + k] * L.Jp[3 + j];
  if (shared) {
#pragma unroll
    for (int j = 0; j < 3; j++) W[18 + j] = L.w * L.Jf[0] * L.Jp[j] + L.w * L.Jf[1] * L.Jp[3 + j];
  }
}

// The cross block of observation i (camera c, point p) as the passes store it, relinearised from
// the scene: the words of cross_block rounded to the storage type WT and widened back (zero where
// the point does not project). The streamed route's sweeps use these in place of stored words, in
// the same expressions, so a streamed run gives the stored run's bits.
template <class WT, bool SHARED>
__device__ __forceinline__ void relinearised_w(const Problem& q, const Bundle& g, uint32_t c,
                                               uint32_t p, uint32_t i, double fscale, double* w) {
  constexpr int K = SHARED ? 21 : 18;
  Lin L;
  double W[21];
  if (linearize(q, c, p, i, fscale, variable_camera(q, c), directional_point(g, p), L))
    cross_block(L, SHARED, W);
  else
    for (int k = 0; k < 21; k++) W[k] = 0;
#pragma unroll
  for (int k = 0; k < K; k++) w[k] = (double)(WT)W[k];
}

// Word k of the cross block of observation (or camera slot) i. Blocked layout [M/64][K][64]: the 64
// observations of a wave read K runs of 64 consecutive words that lie next to each other, one
// contiguous stretch of HBM per wave instead of K distant streams.
template <class WT, bool SHARED>
__device__ __forceinline__ WT* wptr(void* base, uint32_t k, uint32_t i, uint32_t M) {
  constexpr uint32_t K = SHARED ? 21 : 18;
  (void)M;
  return static_cast<WT*>(base) + ((size_t)(i >> 6) * K + k) * 64 + (i & 63);
}

// The factored forms store each observation's cross block as its Jacobian factors
// (ba_factored.hip; FT: float in the mixed form, double in the FP64 form): s du (3), s dv (3),
// [FP32: the camera-frame point q (3),] [s Jf (2)]. The FP64 form recomputes q from the pose and
// the point where the factors are used (camera_point) instead of storing it. Stored word k is
// word kf_raw(k) of cross_factors' eleven.
template <class FT>
__host__ __device__ constexpr int kf_words(bool shared) {
  return sizeof(FT) == 8 ? (shared ? 8 : 6) : (shared ? 11 : 9);
}
template <class FT>
__host__ __device__ constexpr int kf_raw(int k) {
  return sizeof(FT) == 8 && k >= 6 ? k + 3 : k;
}
// kf_words of a form by its stored word size (8: FP64, 4: FP32).
constexpr int kf_words_of(size_t ws, bool shared) {
  return ws == 8 ? kf_words<double>(shared) : kf_words<float>(shared);
}

// Word k of observation i in the point-major stored words (Wp, K words per observation: cross
// blocks or factors), i in the tile whose observations are [s, s + n). Each tile's words are K rows
// of n consecutive words, one row per word, tiles back to back in observation order (no padding):
// a wave's load of one word reads one run of n words, and the tile's runs lie next to each other,
// so its loads share their boundary cache lines.
template <class T>
__device__ __forceinline__ T* tile_word(void* base, uint32_t K, uint32_t s, uint32_t n, uint32_t k,
                                        uint32_t i) {
  return static_cast<T*>(base) + (size_t)K * s + (size_t)k * n + (i - s);
}

// The factor words of one observation: s du (3), s dv (3), q (3)[, s Jf (2)], zero where the point
// does not project.
__device__ __forceinline__ void cross_factors(const Lin& L, bool ok, bool shared, double* f) {
  if (!ok) {
    for (int k = 0; k < 11; k++) f[k] = 0;
    return;
  }
  const double s = sqrt(L.w);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    f[k] = s * L.du[k];
    f[3 + k] = s * L.dv[k];
    f[6 + k] = L.qc[k];
  }
  f[9] = shared ? s * L.Jf[0] : 0.0;
  f[10] = shared ? s * L.Jf[1] : 0.0;
}

// Word k of slot i's factors (FT: float or double): blocked [slots/64][KF][64] as the cross blocks
// are.
template <bool SHARED, class FT = float>
__device__ __forceinline__ FT* fptr(void* base, uint32_t k, uint32_t i) {
  constexpr uint32_t KF = kf_words<FT>(SHARED);
  return static_cast<FT*>(base) + ((size_t)(i >> 6) * KF + k) * 64 + (i & 63);
}

__device__ __forceinline__ double* vx(const Problem& q) { return q.vec; }
__device__ __forceinline__ double* vr(const Problem& q) { return q.vec + q.m; }
__device__ __forceinline__ double* vp(const Problem& q) { return q.vec + 2 * (size_t)q.m; }
__device__ __forceinline__ double* vs(const Problem& q) { return q.vec + 3 * (size_t)q.m; }
__device__ __forceinline__ double* vz(const Problem& q) { return q.vec + 4 * (size_t)q.m; }
__device__ __forceinline__ double* vy(const Problem& q) { return q.vec + 5 * (size_t)q.m; }

// ------------------------------------------------------------------ index checks

// Every observation of point p names p and a camera of p's bundle, and lies in p's range.
__global__ void k_check_points(Problem q) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= q.P) return;
  const Bundle& g = q.bundles[q.point_bundle[p]];
  const uint32_t* ps = point_start(q);
  for (uint32_t i = ps[p]; i < ps[p + 1]; i++) {
    const uint32_t c = obs_camera(q)[i];
    if (obs_point(q)[i] != p || c < g.cam0 || c >= g.cam1) atomicOr(q.error, 1u);
  }
}

// Every slot of a camera task names an observation of that camera, ascending within the camera's
// list; writes the camera-major point ids. Flag 4 (not an error) and the bundle's word 40 mark a
// camera that observes one point twice: the direct route's point lookup needs each camera's points
// strictly ascending, so that bundle keeps PCG.
__global__ void k_check_tasks(Problem q) {
  const uint32_t t = blockIdx.x;
  const uint32_t c = q.tasks[4 * t], s = q.tasks[4 * t + 1], e = q.tasks[4 * t + 2];
  const uint32_t list = camera_start(q)[c];
  for (uint32_t n = s + threadIdx.x; n < e; n += blockDim.x) {
    const uint32_t i = camera_obs(q)[n];
    bool bad = i >= q.M;
    if (!bad) bad = obs_camera(q)[i] != c;
    if (!bad && n > list) {
      const uint32_t before = camera_obs(q)[n - 1];
      bad = before >= i;
      if (!bad && obs_point(q)[before] == obs_point(q)[i]) {
        atomicOr(q.error, 4u);
        bstate(q, q.tasks[4 * t + 3])[S_DUPLICATE] = 1;
      }
    }
    if (bad) {
      atomicOr(q.error, 2u);
      q.cam_point[n] = 0;
    } else {
      q.cam_point[n] = obs_point(q)[i];
    }
  }
}

// ------------------------------------------------------------------ LM set-up and cost

__device__ __forceinline__ void center_of(const double* p, double o[3]) {
  o[0] = -p[0] * p[9] - p[3] * p[10] - p[6] * p[11];
  o[1] = -p[1] * p[9] - p[4] * p[10] - p[7] * p[11];
  o[2] = -p[2] * p[9] - p[5] * p[10] - p[8] * p[11];
}

// initializeBundle, per bundle: lambda 1e-3, gauge origin at the anchor centre, baseline to the
// scale view.
__global__ void k_init_bundle(Problem q) {
  if (threadIdx.x != 0) return;
  const uint32_t b = blockIdx.x;
  const Bundle& g = q.bundles[b];
  double* s = bstate(q, b);
  for (int j = 0; j < S_WORDS; j++) s[j] = 0;
  s[S_LAMBDA] = 1e-3;
  s[S_ACTIVE] = 1;
  double o[3], v[3];
  center_of(q.scene + (size_t)g.anchor * 12, o);
  center_of(q.scene + (size_t)g.scale_view * 12, v);
  const double d0 = o[0] - v[0], d1 = o[1] - v[1], d2 = o[2] - v[2];
  const double baseline = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
  s[S_BASELINE] = baseline;
  s[S_ORIGIN] = o[0];
  s[S_ORIGIN + 1] = o[1];
  s[S_ORIGIN + 2] = o[2];
  if (!(baseline > 1e-10)) {
    s[S_INPUT_ERROR] = 1;
    s[S_ACTIVE] = 0;
  }
}

// Robust cost, squared error and invalid count of one cost block (at most 256 observations of
// one bundle): of the scene (metricLane b:301-314) or of the trial (candidateLane b:1024-1037,
// 1e30 for an invalid lane).
template <bool CANDIDATE>
__global__ __launch_bounds__(COST_OBS) void k_cost(Problem q) {
  __shared__ double lds[3][COST_OBS / 64];
  const uint32_t blk = blockIdx.x;
  const uint32_t b = q.cost_blocks[3 * blk], e = q.cost_blocks[3 * blk + 2];
  const uint32_t i = q.cost_blocks[3 * blk + 1] + threadIdx.x;
  const Bundle& g = q.bundles[b];
  double cost = 0, sq = 0, bad = 0;
  if (i < e) {
    const uint32_t c = obs_camera(q)[i], p = obs_point(q)[i];
    const double* base = CANDIDATE ? q.trial : q.scene;
    const double* pose = base + (size_t)c * 12;
    const double* X = base + (size_t)q.C * 12 + (size_t)p * 3;
    const double* s = bstate(q, b);
    const double fscale = q.shared ? focal_scale(s[CANDIDATE ? S_TRIAL_FOCAL : S_LOG_FOCAL]) : 1.0;
    bool ok = true;
    if (CANDIDATE) ok = fmax(fabs(X[0]), fmax(fabs(X[1]), fabs(X[2]))) <= 1e9;
    double px = 0, py = 0, du[3], dv[3];
    if (ok) {
      const bool dir = directional_point(g, p);
      double qc[3];
      for (int r = 0; r < 3; r++) {
        qc[r] = pose[r * 3] * X[0] + pose[r * 3 + 1] * X[1] + pose[r * 3 + 2] * X[2];
        if (!dir) qc[r] += pose[9 + r];
      }
      ok = project(camera_model(q, c, fscale), qc, px, py, du, dv);
    }
    if (ok) {
      const double* obs = q.scene + (size_t)q.C * 18 + (size_t)q.P * 3 + (size_t)i * 2;
      const double x = px - obs[0], y = py - obs[1];
      sq = x * x + y * y;
      cost = robust_cost(sqrt(sq), q.huber);
    } else if (CANDIDATE) {
      cost = sq = 1e30;
    } else {
      bad = 1;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    cost += __shfl_xor(cost, o, 64);
    sq += __shfl_xor(sq, o, 64);
    bad += __shfl_xor(bad, o, 64);
  }
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    lds[0][wave] = cost;
    lds[1][wave] = sq;
    lds[2][wave] = bad;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const double* v = lds[threadIdx.x];
    q.cost_part[(size_t)blk * 3 + threadIdx.x] = (v[0] + v[1]) + (v[2] + v[3]);
  }
}

// Per bundle: finishBundleMetrics (CANDIDATE = false) or acceptBundleCandidate (true).
template <bool CANDIDATE>
__global__ __launch_bounds__(COST_OBS) void k_finish_cost(Problem q) {
  __shared__ double lds[COST_OBS / 64];
  const uint32_t b = blockIdx.x;
  const Bundle& g = q.bundles[b];
  double* s = bstate(q, b);
  double cost = 0, sq = 0, bad = 0;
  for (uint32_t k = g.cost0 + threadIdx.x; k < g.cost1; k += COST_OBS) {
    cost += q.cost_part[(size_t)k * 3];
    sq += q.cost_part[(size_t)k * 3 + 1];
    bad += q.cost_part[(size_t)k * 3 + 2];
  }
  const double active = s[S_ACTIVE], valid = s[S_TRIAL_VALID], current = s[S_COST];
  const double input_error = s[S_INPUT_ERROR];
  cost = block_sum(cost, lds);
  sq = block_sum(sq, lds);
  bad = block_sum(bad, lds);
  if (threadIdx.x != 0) return;
  if (!CANDIDATE) {
    if (input_error != 0) return;
    if (bad > 0) {
      s[S_INPUT_ERROR] = 2;
      s[S_ACTIVE] = 0;
      return;
    }
    s[S_COST] = cost;
    s[S_SSE] = sq;
    s[S_COST0] = cost;
    s[S_SSE0] = sq;
    if (cost < 1e-20) s[S_ACTIVE] = 0;
    return;
  }
  s[S_COMMIT] = 0;
  if (active == 0) return;
  if (valid != 0 && cost < current) {
    const double change = current - cost;
    s[S_COST] = cost;
    s[S_SSE] = sq;
    s[S_ACCEPTED] += 1;
    s[S_COMMIT] = 1;
    s[S_LAMBDA] = fmax(1e-10, s[S_LAMBDA] / 3);
    if (change < q.stop_change * fmax(1.0, cost)) s[S_ACTIVE] = 0;
  } else {
    s[S_REJECTED] += 1;
    s[S_LAMBDA] *= 10;
    if (s[S_PCG_FAILED] != 0) s[S_FAILED] += 1;
    if (s[S_LAMBDA] > 1e12) s[S_ACTIVE] = 0;
  }
}

// ------------------------------------------------------------------ assembly

// One wave per tile of whole points, one lane per observation: linearise, write the cross blocks
// or (FACTORS) the Jacobian factors (point-major), stage the point terms in LDS; then one lane per
// point sums its observations in order, damps V, inverts it and writes the point block. The shared
// route adds the point's share of the reduced focal diagonal and RHS (bundleFocalPoints), summed
// per tile in point order.
template <class WT, bool SHARED, bool EXACT, bool FACTORS = (sizeof(WT) == 4)>
__global__ __launch_bounds__(TILE* TILE_WAVES) void k_point_pass(Problem q) {
  // staged words per observation: V 6, b 3, [fc 3], [exact systems: the Gauss-Newton V 6], bad
  constexpr int GN = SHARED ? 12 : 9;
  constexpr int NS = GN + (EXACT ? 6 : 0) + 1;
  __shared__ double st_all[TILE_WAVES][TILE][NS];
  __shared__ double fsum_all[TILE_WAVES][TILE][2];
  // the wave's index is uniform across its lanes: keep the tile's loads and branches scalar
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / TILE), lane = threadIdx.x % TILE;
  const uint32_t t = point_tile(wave);
  if (t >= q.n_tiles) return;
  const uint32_t b = q.tile_bundle[t];
  const double* sb = bstate(q, b);
  if (sb[S_ACTIVE] == 0) return;
  const Bundle& g = q.bundles[b];
  auto& st = st_all[wave];
  auto& fsum = fsum_all[wave];
  const uint32_t fp = q.tiles[t], ep = q.tiles[t + 1];
  const uint32_t* ps = point_start(q);
  const uint32_t s = ps[fp], e = ps[ep];
  const uint32_t my = fp + lane;
  const bool has = my < ep;
  const uint32_t mys = has ? ps[my] : 0, mye = has ? ps[my + 1] : 0;
  const double lambda = sb[S_LAMBDA];
  const double fscale = SHARED ? focal_scale(sb[S_LOG_FOCAL]) : 1.0;
  double V[6] = {0, 0, 0, 0, 0, 0}, bp[3] = {0, 0, 0}, fc[3] = {0, 0, 0};
  double Vg[EXACT ? 6 : 1] = {0};
  bool bad = false;
  const uint32_t K = SHARED ? 21 : 18;
  for (uint32_t base = s; base < e; base += TILE) {
    const uint32_t i = base + lane;
    if (i < e) {
      const uint32_t c = obs_camera(q)[i], p = obs_point(q)[i];
      const bool var = variable_camera(q, c), dir = directional_point(g, p);
      Lin L;
      const bool ok = linearize(q, c, p, i, fscale, var, dir, L);
      double W[21];
      double Cx[EXACT ? 55 : 1];
      if (ok) {
        cross_block(L, SHARED, W);
        if (EXACT) {
          exact_correction(L, var, dir, q.huber, Cx, !q.gauss_newton);
          for (int k = 0; k < 6; k++)
            for (int j = 0; j < 3; j++) W[k * 3 + j] += Cx[up10(k, 6 + j)];
          if (SHARED)
            for (int j = 0; j < 3; j++) W[18 + j] += Cx[up10(6 + j, 9)];
        }
      } else {
        for (int k = 0; k < 21; k++) W[k] = 0;
      }
      if (stores_w(g) && q.Wp) {
        if constexpr (FACTORS) {
          double fw[11];
          cross_factors(L, ok, SHARED, fw);
          for (int k = 0; k < kf_words<WT>(SHARED); k++)
            *tile_word<WT>(q.Wp, kf_words<WT>(SHARED), s, e - s, k, i) = (WT)fw[kf_raw<WT>(k)];
        } else {
          for (uint32_t k = 0; k < K; k++) *tile_word<WT>(q.Wp, K, s, e - s, k, i) = (WT)W[k];
        }
      }
      double* o = st[lane];
      if (ok) {
        const double w = L.w;
        const double a0 = w * L.Jp[0], a1 = w * L.Jp[1], a2 = w * L.Jp[2];
        const double c0 = w * L.Jp[3], c1 = w * L.Jp[4], c2 = w * L.Jp[5];
        o[0] = a0 * L.Jp[0] + c0 * L.Jp[3];  // V00
        o[1] = a1 * L.Jp[0] + c1 * L.Jp[3];  // V10
        o[2] = a2 * L.Jp[0] + c2 * L.Jp[3];  // V20
        o[3] = a1 * L.Jp[1] + c1 * L.Jp[4];  // V11
        o[4] = a2 * L.Jp[1] + c2 * L.Jp[4];  // V21
        o[5] = a2 * L.Jp[2] + c2 * L.Jp[5];  // V22
        if (EXACT)
          for (int k = 0; k < 6; k++) o[GN + k] = o[k];
        o[6] = -(a0 * L.r[0]) - c0 * L.r[1];
        o[7] = -(a1 * L.r[0]) - c1 * L.r[1];
        o[8] = -(a2 * L.r[0]) - c2 * L.r[1];
        if (EXACT) {
          o[0] += Cx[up10(6, 6)];
          o[1] += Cx[up10(6, 7)];
          o[2] += Cx[up10(6, 8)];
          o[3] += Cx[up10(7, 7)];
          o[4] += Cx[up10(7, 8)];
          o[5] += Cx[up10(8, 8)];
          if (q.rhs_in) o[6] = o[7] = o[8] = 0;
        }
        if (SHARED) {
          o[9] = W[18];
          o[10] = W[19];
          o[11] = W[20];
        }
        o[NS - 1] = 0;
      } else {
        for (int k = 0; k < NS - 1; k++) o[k] = 0;
        o[NS - 1] = 1;
      }
    }
    wave_sync();
    if (has) {
      const uint32_t lo = mys > base ? mys : base, hi = mye < base + TILE ? mye : base + TILE;
      for (uint32_t j = lo; j < hi; j++) {
        const double* o = st[j - base];
        for (int k = 0; k < 6; k++) V[k] += o[k];
        for (int k = 0; k < 3; k++) bp[k] += o[6 + k];
        if (SHARED)
          for (int k = 0; k < 3; k++) fc[k] += o[9 + k];
        if (EXACT)
          for (int k = 0; k < 6; k++) Vg[k] += o[GN + k];
        bad = bad || o[NS - 1] != 0;
      }
    }
    wave_sync();
  }
  bool failed = false, held = false;
  double dF = 0, dR = 0;
  if (has) {
    double inv[9];
    failed = bad;
    if (EXACT && q.rhs_in)
      for (int k = 0; k < 3; k++) bp[k] = q.rhs_in[(size_t)q.m + (size_t)my * 3 + k];
    if (!failed) {
      double A[9] = {V[0], V[1], V[2], V[1], V[3], V[4], V[2], V[4], V[5]};
      A[0] += lambda * fmax(A[0], 1e-10);
      A[4] += lambda * fmax(A[4], 1e-10);
      A[8] += lambda * fmax(A[8], 1e-10);
      double G[9] = {Vg[0], Vg[EXACT ? 1 : 0], Vg[EXACT ? 2 : 0], Vg[EXACT ? 1 : 0], Vg[EXACT ? 3 : 0],
                     Vg[EXACT ? 4 : 0], Vg[EXACT ? 2 : 0], Vg[EXACT ? 4 : 0], Vg[EXACT ? 5 : 0]};
      if (q.exact_gauge && directional_point(g, my)) {
        // the radial direction of a directional point leaves every projection unchanged: a zero
        // row and column of the Hessian. Fix it with mu X X^T (X is unit after normalisation);
        // the right-hand sides carry no radial component, so the solution keeps none.
        const double* X = q.scene + (size_t)q.C * 12 + (size_t)my * 3;
        const double mu = fmax(fmax(A[0], A[4]), fmax(A[8], 1e-10));
        for (int j = 0; j < 3; j++)
          for (int k = 0; k < 3; k++) A[j * 3 + k] += mu * X[j] * X[k];
        if (EXACT) {
          const double mg = fmax(fmax(G[0], G[4]), fmax(G[8], 1e-10));
          for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) G[j * 3 + k] += mg * X[j] * X[k];
        }
      }
      // Undamped exact-Hessian systems (Newton polish, adjoint) hold fixed a point whose
      // Gauss-Newton block sum w Jp^T Jp is singular to rounding: a track that has run off
      // towards infinity, whose observations no longer depend on its position along the ray and
      // which no finite step settles. Its zero block removes it from the reduced system; its
      // observations still act on their cameras.
      if (EXACT) held = q.exact_gauge && lambda == 0 && singular_to_rounding3(G, kHeldPoint);
      if (!held) failed = !inverse3(A, inv);
    }
    double* out = q.pblk + (size_t)my * 12;
    double* pf = (SHARED && q.pfc) ? q.pfc + (size_t)my * 3 : nullptr;
    if (failed || held) {
      for (int k = 0; k < 12; k++) out[k] = 0;
      if (pf) pf[0] = pf[1] = pf[2] = 0;
    } else {
      for (int k = 0; k < 9; k++) out[k] = inv[k];
      for (int k = 0; k < 3; k++) out[9 + k] = bp[k];
      if (SHARED) {
        for (int j = 0; j < 3; j++) {
          double tj = 0;
          for (int k = 0; k < 3; k++) tj += fc[k] * inv[k * 3 + j];
          dF += tj * fc[j];
          dR += tj * bp[j];
        }
        if (pf) {
          double L[6];
          lower_cholesky3(inv, L);
          pf[0] = L[0] * fc[0] + L[1] * fc[1] + L[3] * fc[2];
          pf[1] = L[2] * fc[1] + L[4] * fc[2];
          pf[2] = L[5] * fc[2];
        }
      }
    }
  }
  const unsigned long long any = __ballot(failed);
  const uint32_t n_held = (uint32_t)__popcll(__ballot(held));
  if (SHARED) {
    fsum[lane][0] = dF;
    fsum[lane][1] = dR;
    wave_sync();
  }
  if (lane == 0) {
    // bit 0: a point block failed; above it, the points held fixed
    q.tile_fail[t] = (any != 0ull ? 1u : 0u) | (n_held << 1);
    if (SHARED) {
      double f0 = 0, f1 = 0;
      for (uint32_t j = 0; j < ep - fp && j < (uint32_t)TILE; j++) {
        f0 += fsum[j][0];
        f1 += fsum[j][1];
      }
      q.tile_focal[2 * (size_t)t] = f0;
      q.tile_focal[2 * (size_t)t + 1] = f1;
    }
  }
}

// Packed upper triangle of a symmetric 6x6 block.
__host__ __device__ constexpr int up6(int j, int l) { return j * 6 - j * (j - 1) / 2 + (l - j); }

// One workgroup per camera-list task: relinearise each observation of the task in camera-major
// order, write its cross block (or factors) to the camera-major copy, and sum
//   U = sum w Jc^T Jc (21), Q = sum W V^-1 W^T (21), rhs = -sum w Jc^T r - W V^-1 b_p (6)
//   [shared] H_cf = sum w Jc^T Jf (6), H_ff = sum w Jf^T Jf, b_f = -sum w Jf^T r
// into ASM_WORDS per task, in a fixed order. Two passes over the task's observations, each
// relinearising them: the first sums U and the focal words, the second Q and the RHS and writes
// the stored words, so that a thread holds half of the 56 sums at once (twice the waves a SIMD
// holds); every word takes the same terms in the same order as one pass would give it.
template <int NW>
__device__ __forceinline__ void task_sums(double* acc, double (*red)[29], double* out,
                                          const int* word) {
#pragma unroll
  for (int k = 0; k < NW; k++)
    for (int o = 32; o > 0; o >>= 1) acc[k] += __shfl_xor(acc[k], o, 64);
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < NW; k++) red[wave][k] = acc[k];
  __syncthreads();
  if (threadIdx.x < (uint32_t)NW) {
    double v = 0;
    for (int w = 0; w < TASK_THREADS / 64; w++) v += red[w][threadIdx.x];
    out[word[threadIdx.x]] = v;
  }
  __syncthreads();
}

template <class WT, bool SHARED, bool EXACT>
__global__ __launch_bounds__(TASK_THREADS) void k_camera_pass(Problem q) {
  constexpr int N1 = SHARED ? 29 : 21, N2 = 27;
  __shared__ double red[TASK_THREADS / 64][29];
  const uint32_t t = blockIdx.x;
  const uint32_t c = q.tasks[4 * t], s = q.tasks[4 * t + 1], e = q.tasks[4 * t + 2];
  const uint32_t b = q.camera_bundle[c];
  const double* sb = bstate(q, b);
  if (sb[S_ACTIVE] == 0) return;
  const Bundle& g = q.bundles[b];
  const bool var = variable_camera(q, c);
  const double fscale = SHARED ? focal_scale(sb[S_LOG_FOCAL]) : 1.0;
  const uint32_t K = SHARED ? 21 : 18;
  // with a right-hand side given, the camera's own b_c enters once, in k_camera_finish
  const double rw = (EXACT && q.rhs_in) ? 0.0 : 1.0;
  double* out = q.task_part + (size_t)t * ASM_WORDS;
  // pass 1: U (words 0-20), H_cf (48-53), H_ff (54), b_f (55)
  {
    double acc[N1];
#pragma unroll
    for (int k = 0; k < N1; k++) acc[k] = 0;
    for (uint32_t n = s + threadIdx.x; n < e; n += TASK_THREADS) {
      const uint32_t i = camera_obs(q)[n], p = q.cam_point[n];
      const bool dir = directional_point(g, p);
      Lin L;
      const bool ok = linearize(q, c, p, i, fscale, var, dir, L);
      if (!ok) continue;
      double Cx[EXACT ? 55 : 1];
      if (EXACT) exact_correction(L, var, dir, q.huber, Cx, !q.gauss_newton);
      const double w = L.w;
      if (var) {
#pragma unroll
        for (int j = 0; j < 6; j++) {
          const double wj0 = w * L.Jc[j], wj1 = w * L.Jc[6 + j];
#pragma unroll
          for (int l = j; l < 6; l++) {
            acc[up6(j, l)] += wj0 * L.Jc[l] + wj1 * L.Jc[6 + l];
            if (EXACT) acc[up6(j, l)] += Cx[up10(j, l)];
          }
          if (SHARED) acc[21 + j] += wj0 * L.Jf[0] + wj1 * L.Jf[1];
          if (SHARED && EXACT) acc[21 + j] += Cx[up10(j, 9)];
        }
      }
      if (SHARED) {
        acc[27] += w * (L.Jf[0] * L.Jf[0] + L.Jf[1] * L.Jf[1]);
        if (EXACT) acc[27] += Cx[up10(9, 9)];
        acc[28] += rw * -(w * (L.Jf[0] * L.r[0] + L.Jf[1] * L.r[1]));
      }
    }
    static constexpr int words[29] = {0,  1,  2,  3,  4,  5,  6,  7,  8,  9,  10, 11, 12, 13, 14,
                                      15, 16, 17, 18, 19, 20, 48, 49, 50, 51, 52, 53, 54, 55};
    task_sums<N1>(acc, red, out, words);
  }
  // pass 2: Q (words 21-41) and the RHS (42-47); the stored words
  {
    double acc[N2];
#pragma unroll
    for (int k = 0; k < N2; k++) acc[k] = 0;
    for (uint32_t n = s + threadIdx.x; n < e; n += TASK_THREADS) {
      const uint32_t i = camera_obs(q)[n], p = q.cam_point[n];
      const bool dir = directional_point(g, p);
      Lin L;
      const bool ok = linearize(q, c, p, i, fscale, var, dir, L);
      double W[21];
      if (ok) {
        cross_block(L, SHARED, W);
        if (EXACT) {
          double Cx[55];
          exact_correction(L, var, dir, q.huber, Cx, !q.gauss_newton);
          for (int k = 0; k < 6; k++)
            for (int j = 0; j < 3; j++) W[k * 3 + j] += Cx[up10(k, 6 + j)];
          if (SHARED)
            for (int j = 0; j < 3; j++) W[18 + j] += Cx[up10(6 + j, 9)];
        }
      } else {
        for (int k = 0; k < 21; k++) W[k] = 0;
      }
      if (stores_w(g) && q.Wc) {
        if (sizeof(WT) == 4 || (!EXACT && q.factored64)) {
          double fw[11];
          cross_factors(L, ok, SHARED, fw);
          for (int k = 0; k < kf_words<WT>(SHARED); k++)
            *fptr<SHARED, WT>(q.Wc, k, n) = (WT)fw[kf_raw<WT>(k)];
        } else {
          for (uint32_t k = 0; k < K; k++) *wptr<WT, SHARED>(q.Wc, k, n, q.M) = (WT)W[k];
        }
      }
      if (!ok) continue;
      const double w = L.w;
      if (var) {
        const double* pb = q.pblk + (size_t)p * 12;
        double Vi[9], bpv[3];
        for (int k = 0; k < 9; k++) Vi[k] = pb[k];
        for (int k = 0; k < 3; k++) bpv[k] = pb[9 + k];
        double T[18];
#pragma unroll
        for (int j = 0; j < 6; j++)
#pragma unroll
          for (int l = 0; l < 3; l++)
            T[j * 3 + l] = W[j * 3] * Vi[l] + W[j * 3 + 1] * Vi[3 + l] + W[j * 3 + 2] * Vi[6 + l];
        if (q.camG) {
          // the direct route's factor of this slot's share of W V^-1 W^T
          double Lf[6];
          lower_cholesky3(Vi, Lf);
          double* G = q.camG + (size_t)n * 18;
#pragma unroll
          for (int j = 0; j < 6; j++) {
            G[j * 3] = W[j * 3] * Lf[0] + W[j * 3 + 1] * Lf[1] + W[j * 3 + 2] * Lf[3];
            G[j * 3 + 1] = W[j * 3 + 1] * Lf[2] + W[j * 3 + 2] * Lf[4];
            G[j * 3 + 2] = W[j * 3 + 2] * Lf[5];
          }
        }
#pragma unroll
        for (int j = 0; j < 6; j++) {
          const double wj0 = w * L.Jc[j], wj1 = w * L.Jc[6 + j];
#pragma unroll
          for (int l = j; l < 6; l++)
            acc[up6(j, l)] +=
                T[j * 3] * W[l * 3] + T[j * 3 + 1] * W[l * 3 + 1] + T[j * 3 + 2] * W[l * 3 + 2];
          acc[21 + j] += rw * (-(wj0 * L.r[0]) - wj1 * L.r[1]) -
                         (T[j * 3] * bpv[0] + T[j * 3 + 1] * bpv[1] + T[j * 3 + 2] * bpv[2]);
        }
      }
    }
    static constexpr int words[27] = {21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 31, 32, 33, 34,
                                      35, 36, 37, 38, 39, 40, 41, 42, 43, 44, 45, 46, 47};
    task_sums<N2>(acc, red, out, words);
  }
}

// One thread per camera: sum the camera's tasks in order, damp, form the Schur diagonal block
// S_cc = A - Q and its inverse (block-Jacobi preconditioner, failure marker -1), as
// bundleCameraBlocks b:423-489 does for the whole camera.
template <bool SHARED>
__global__ void k_camera_finish(Problem q) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= q.C) return;
  const double* sb = bstate(q, q.camera_bundle[c]);
  if (sb[S_ACTIVE] == 0) return;
  constexpr int NV = SHARED ? 56 : 48;
  double acc[NV];
  for (int k = 0; k < NV; k++) acc[k] = 0;
  for (uint32_t t = q.task_start[c]; t < q.task_start[c + 1]; t++)
    for (int k = 0; k < NV; k++) acc[k] += q.task_part[(size_t)t * ASM_WORDS + k];
  double* A = q.camA + (size_t)c * 36;
  double* inv = q.camInv + (size_t)c * 36;
  double* rhs = q.camRhs + (size_t)c * 6;
  if (SHARED) {
    q.camFocal[2 * (size_t)c] = acc[54];
    q.camFocal[2 * (size_t)c + 1] = acc[55];
  }
  if (!variable_camera(q, c)) {
    for (int j = 0; j < 36; j++) A[j] = inv[j] = (j % 7 == 0) ? 1.0 : 0.0;
    if (q.camS)
      for (int j = 0; j < 36; j++) q.camS[(size_t)c * 36 + j] = (j % 7 == 0) ? 1.0 : 0.0;
    for (int j = 0; j < 6; j++) {
      rhs[j] = 0;
      if (SHARED) q.camHcf[(size_t)c * 6 + j] = 0;
    }
    return;
  }
  const double lambda = sb[S_LAMBDA];
  double a[36], S[36];
  for (int j = 0; j < 6; j++)
    for (int l = j; l < 6; l++) {
      a[j * 6 + l] = a[l * 6 + j] = acc[up6(j, l)];
      S[j * 6 + l] = S[l * 6 + j] = acc[21 + up6(j, l)];
    }
  for (int j = 0; j < 6; j++) a[j * 6 + j] += lambda * fmax(a[j * 6 + j], 1e-10);
  if (q.exact_gauge) {
    // The exact Hessian of a one-fixed-camera bundle is singular along the similarity about the
    // anchor centre o (the gauge the forward fixes by rescaling the baseline b = |C_s - o|). Adding
    // mu a a^T to the scale view's translation block, a = R_s (C_s - o) / b (unit; the baseline's
    // gradient up to sign), makes it definite without changing the solution: for a right-hand side
    // orthogonal to the gauge direction n (a gradient, or a gauge-projected cotangent) and a.n != 0,
    // the solution of H x = b with a.x = 0 also solves (H + mu a a^T) x = b, for every mu > 0.
    const uint32_t b = q.camera_bundle[c];
    const Bundle& g = q.bundles[b];
    if (g.normalize_scale && c == g.scale_view) {
      const double* pose = q.scene + (size_t)c * 12;
      const double* o = bstate(q, b) + S_ORIGIN;
      double C[3], d[3], av[3];
      center_of(pose, C);
      for (int i = 0; i < 3; i++) d[i] = C[i] - o[i];
      const double base = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      for (int i = 0; i < 3; i++)
        av[i] = (pose[i * 3] * d[0] + pose[i * 3 + 1] * d[1] + pose[i * 3 + 2] * d[2]) / base;
      const double mu = fmax(fmax(a[21], a[28]), fmax(a[35], 1e-10));
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) a[(3 + i) * 6 + 3 + j] += mu * av[i] * av[j];
    }
  }
  for (int j = 0; j < 36; j++) {
    A[j] = a[j];
    S[j] = a[j] - S[j];
  }
  if (q.camS)
    for (int j = 0; j < 36; j++) q.camS[(size_t)c * 36 + j] = S[j];
  for (int j = 0; j < 6; j++) {
    rhs[j] = acc[42 + j];
    if (q.rhs_in) rhs[j] += q.rhs_in[(size_t)c * 6 + j];
    if (SHARED) q.camHcf[(size_t)c * 6 + j] = acc[48 + j];
  }
  Chol6 f;
  if (!cholesky6(S, f)) {
    for (int j = 0; j < 36; j++) inv[j] = 0;
    inv[0] = -1;
    return;
  }
  double Mi[36];
  for (int column = 0; column < 6; column++) {
    double e6[6] = {0, 0, 0, 0, 0, 0}, x[6];
    e6[column] = 1;
    chol6_solve(f, e6, x);
    for (int j = 0; j < 6; j++) Mi[j * 6 + column] = x[j];
  }
  for (int j = 0; j < 6; j++)
    for (int k = j + 1; k < 6; k++) {
      const double v = (Mi[j * 6 + k] + Mi[k * 6 + j]) * 0.5;
      Mi[j * 6 + k] = v;
      Mi[k * 6 + j] = v;
    }
  for (int j = 0; j < 36; j++) inv[j] = Mi[j];
}

// ------------------------------------------------------------------ PCG

// Row `row` = 6c + j of a 6x6-block-diagonal product: sum_k B_c[j][k] v[6c + k]. Consecutive
// rows read consecutive 48-byte block rows, so a wave's loads are coalesced.
__device__ __forceinline__ double block_row(const double* blocks, uint32_t row, const double* v) {
  const uint32_t c = row / 6;
  const double* bk = blocks + (size_t)c * 36 + (size_t)(row - c * 6) * 6;
  const double* vc = v + (size_t)c * 6;
  double a = 0;
#pragma unroll
  for (int k = 0; k < 6; k++) a += bk[k] * vc[k];
  return a;
}

// initializeBundleCg (with dampBundleFocal and the rest of bundleFocalPoints on the shared route),
// one workgroup per bundle: point and camera block checks, x = 0, r = b, z = M^-1 r, p = z, rho,
// |r0|^2.
template <bool SHARED>
__global__ __launch_bounds__(CG_THREADS) void k_init_cg(Problem q) {
  __shared__ double lds[2 * (CG_THREADS / 64)];
  const uint32_t bi = blockIdx.x;
  const Bundle& g = q.bundles[bi];
  double* s = bstate(q, bi);
  const uint32_t tid = threadIdx.x;
  const double active = s[S_ACTIVE], lambda = s[S_LAMBDA];
  __syncthreads();
  if (tid == 0) {
    for (int j = S_RHO; j < 26; j++) s[j] = 0;
    s[S_DIRECT] = 0;
    s[S_STEP_LIMIT] = pcg_budget(q, g);
  }
  if (active == 0) return;
  double fail_tiles = 0, fail_cameras = 0, held = 0;
  for (uint32_t t = g.tile0 + tid; t < g.tile1; t += CG_THREADS) {
    fail_tiles += (q.tile_fail[t] & 1u) ? 1.0 : 0.0;
    held += (double)(q.tile_fail[t] >> 1);
  }
  for (uint32_t c = g.cam0 + tid; c < g.cam1; c += CG_THREADS)
    if (!(q.camInv[(size_t)c * 36] > 0)) fail_cameras += 1;
  block_sum2(fail_tiles, fail_cameras, lds);
  held = block_sum(held, lds);
  double fail = fail_tiles + fail_cameras;
  if (tid == 0) {
    s[S_FAIL_TILES] = fail_tiles;
    s[S_FAIL_CAMERAS] = fail_cameras;
    s[S_FAIL_FOCAL] = 0;
    s[S_HELD] = held;
  }
  double hff = 0, bf = 0, tf = 0, tr = 0;
  if (SHARED) {
    for (uint32_t c = g.cam0 + tid; c < g.cam1; c += CG_THREADS) {
      hff += q.camFocal[2 * (size_t)c];
      bf += q.camFocal[2 * (size_t)c + 1];
    }
    for (uint32_t t = g.tile0 + tid; t < g.tile1; t += CG_THREADS) {
      tf += q.tile_focal[2 * (size_t)t];
      tr += q.tile_focal[2 * (size_t)t + 1];
    }
    hff = block_sum(hff, lds);
    bf = block_sum(bf, lds);
    tf = block_sum(tf, lds);
    tr = block_sum(tr, lds);
  }
  double finv = 0, frhs = 0;
  if (SHARED && q.rhs_in) bf = q.rhs_in[g.focal_row];
  if (SHARED) {
    const double damped = hff + lambda * fmax(hff, 1e-10);
    const double diag = damped - tf;
    frhs = bf - tr;
    finv = diag > 0 ? 1 / diag : -1.0;
    if (tid == 0) {
      s[S_FOCAL_DIAG] = damped;
      s[S_FOCAL_INV] = finv;
      s[S_FOCAL_RHS] = frhs;
      s[S_FOCAL_SCHUR] = diag;
    }
    if (!(finv > 0)) {
      fail += 1;
      if (tid == 0) s[S_FAIL_FOCAL] = 1;
    }
  }
  if (fail > 0) {
    if (tid == 0) s[S_PCG_FAILED] = 1;
    return;
  }
  double rho = 0, res = 0;
  double* x = vx(q);
  double* r = vr(q);
  double* p = vp(q);
  double* z = vz(q);
  for (uint32_t row = 6 * g.cam0 + tid; row < 6 * g.cam1; row += CG_THREADS) {
    const double rv = q.camRhs[row];
    const double zv = block_row(q.camInv, row, q.camRhs);
    x[row] = 0;
    r[row] = rv;
    z[row] = zv;
    p[row] = zv;
    rho += rv * zv;
    res += rv * rv;
  }
  if (SHARED && tid == 0) {
    const uint32_t f = g.focal_row;
    const double zf = finv * frhs;
    x[f] = 0;
    r[f] = frhs;
    z[f] = zf;
    p[f] = zf;
    rho += frhs * zf;
    res += frhs * frhs;
  }
  rho = block_sum(rho, lds);
  res = block_sum(res, lds);
  if (tid == 0) {
    s[S_RHO] = rho;
    s[S_RHO_NEXT] = rho;
    s[S_R0] = res;
    s[S_R] = res;
    s[S_PASS_STOP] = q.tol * q.tol * res;
    if (res > 0 && rho > 0)
      s[S_PCG_ACTIVE] = 1;
    else if (res > 1e-24)
      s[S_PCG_FAILED] = 1;
  }
}

// acc += st[j - base] for j in [lo, hi), in that order: four rows' LDS reads are issued together
// and waited for once, the adds stay one after another, so the sums are the plain loop's.
__device__ __forceinline__ void segment_sum(const double (*st)[3], uint32_t base, uint32_t lo,
                                            uint32_t hi, double acc[3]) {
  uint32_t j = lo;
  for (; j + 4 <= hi; j += 4) {
    double r[4][3];
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int k = 0; k < 3; k++) r[u][k] = st[j + u - base][k];
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int k = 0; k < 3; k++) acc[k] += r[u][k];
  }
  for (; j < hi; j++)
#pragma unroll
    for (int k = 0; k < 3; k++) acc[k] += st[j - base][k];
}

// The point sweep over tiles of whole points, one lane per observation.
//   SCHUR:  y_p = V_p^-1 sum_i (W_i^T v_c(i) + w_fi v_f)                (bundleSchurPointAt)
//   UPDATE: X_p' = X_p + V_p^-1 (b_p - sum_i (W_i^T x_c(i) + w_fi x_f))  (updateBundleCandidate)
//   SOLVE:  y_p = V_p^-1 (b_p - sum_i (W_i^T x_c(i) + w_fi x_f))  (the point part of a solve)
//   CERTIFY: as SCHUR on x, for the bundles with an uncertified direct solution
enum SweepMode { SWEEP_SCHUR = 0, SWEEP_UPDATE = 1, SWEEP_SOLVE = 2, SWEEP_CERTIFY = 3 };

// One tile t of the point sweep by one wave (st: the wave's LDS slice). RECOMPUTE: the streamed
// route, whose cross blocks are relinearised (relinearised_w) instead of read from Wp.
template <class WT, bool SHARED, int MODE, bool RECOMPUTE = false>
__device__ __forceinline__ void sweep_tile(const Problem& q, uint32_t t, uint32_t lane,
                                           double (*st)[3]) {
  // independent loads first: the tile's bounds and bundle, then the bundle's state word
  const uint32_t fp = q.tiles[t], ep = q.tiles[t + 1], b = q.tile_bundle[t];
  const double* sb = bstate(q, b);
  if (MODE == SWEEP_SCHUR && sb[S_PCG_ACTIVE] == 0) return;
  if (MODE == SWEEP_CERTIFY && sb[S_DIRECT] != 1) return;
  if ((MODE == SWEEP_UPDATE || MODE == SWEEP_SOLVE) && (sb[S_ACTIVE] == 0 || sb[S_PCG_FAILED] != 0))
    return;
  if (q.n_unstored != 0 && !stores_w(q.bundles[b])) return;
  const double* v = MODE == SWEEP_SCHUR ? vp(q) : vx(q);
  const uint32_t* ps = point_start(q);
  const uint32_t s = ps[fp], e = ps[ep];
  const uint32_t my = fp + lane;
  const bool has = my < ep;
  const uint32_t mys = has ? ps[my] : 0, mye = has ? ps[my + 1] : 0;
  const double vf = SHARED ? v[6 * q.C + b] : 0.0;  // the bundle's focal row
  const double fscale = RECOMPUTE && SHARED ? focal_scale(sb[S_LOG_FOCAL]) : 1.0;
  constexpr uint32_t KW = SHARED ? 21 : 18;
  double acc[3] = {0, 0, 0};
  for (uint32_t base = s; base < e; base += TILE) {
    const uint32_t i = base + lane;
    if (i < e) {
      const uint32_t c = obs_camera(q)[i];
      double vc[6];
      for (int k = 0; k < 6; k++) vc[k] = v[c * 6 + k];
      double wr[RECOMPUTE ? (SHARED ? 21 : 18) : 1];
      if (RECOMPUTE) relinearised_w<WT, SHARED>(q, q.bundles[b], c, obs_point(q)[i], i, fscale, wr);
      double tj[3];
#pragma unroll
      for (int j = 0; j < 3; j++) {
        double a = 0;
#pragma unroll
        for (int k = 0; k < 6; k++)
          a += (RECOMPUTE ? wr[k * 3 + j]
                          : wload(tile_word<WT>(q.Wp, KW, s, e - s, k * 3 + j, i))) *
               vc[k];
        if (SHARED)
          a += (RECOMPUTE ? wr[(18 + j) % (SHARED ? 21 : 18)]
                          : wload(tile_word<WT>(q.Wp, KW, s, e - s, 18 + j, i))) * vf;
        tj[j] = a;
      }
      st[lane][0] = tj[0];
      st[lane][1] = tj[1];
      st[lane][2] = tj[2];
    }
    wave_sync();
    if (has) {
      const uint32_t lo = mys > base ? mys : base, hi = mye < base + TILE ? mye : base + TILE;
      segment_sum(st, base, lo, hi, acc);
    }
    wave_sync();
  }
  if (!has) return;
  const double* pb = q.pblk + (size_t)my * 12;
  if (MODE == SWEEP_SCHUR || MODE == SWEEP_CERTIFY) {
    double* y = vy(q) + (size_t)my * 3;
    for (int j = 0; j < 3; j++) y[j] = pb[j * 3] * acc[0] + pb[j * 3 + 1] * acc[1] + pb[j * 3 + 2] * acc[2];
  } else if (MODE == SWEEP_UPDATE) {
    const double r0 = pb[9] - acc[0], r1 = pb[10] - acc[1], r2 = pb[11] - acc[2];
    const size_t at = (size_t)q.C * 12 + (size_t)my * 3;
    for (int j = 0; j < 3; j++)
      q.trial[at + j] = q.scene[at + j] + (pb[j * 3] * r0 + pb[j * 3 + 1] * r1 + pb[j * 3 + 2] * r2);
  } else {
    const double r0 = pb[9] - acc[0], r1 = pb[10] - acc[1], r2 = pb[11] - acc[2];
    double* y = vy(q) + (size_t)my * 3;
    for (int j = 0; j < 3; j++) y[j] = pb[j * 3] * r0 + pb[j * 3 + 1] * r1 + pb[j * 3 + 2] * r2;
  }
}

// The Schur product's point sweep with the camera products fused in: a wave that has summed y_p
// for the points of its tile still holds each observation's cross block, so it forms
// t_i = W_i y_p (6 words, 7 with the focal row) at once, in FP64, and writes it as the
// observation's 64-byte row of tprod (point-major, so a wave's rows are one contiguous 4 KiB
// store); k_camera_tsum then gathers the rows of each camera task instead of streaming Wc and
// gathering y. The cross blocks (FP64, or the mixed form's FP32) are read once per product instead
// of twice (t written and read back, from the device's last-level cache where it still holds
// them). Each t_i and each sum is the expression and order of k_point_sweep + k_camera_sweep, so
// the products are the same bits. (FP32 rows, tried for the mixed form, cost the refinement its
// 1e-6 bar on a 40-camera scene.) A tile of more than 64 observations (one point with more) reads
// its cross blocks a second time for its products.
template <class WT, bool SHARED>
__device__ __forceinline__ void sweep_tile_fused(const Problem& q, uint32_t t, uint32_t lane,
                                                 double (*st)[3], double (*ys)[3],
                                                 uint32_t* own) {
  constexpr int K = SHARED ? 21 : 18, NV = SHARED ? 7 : 6;
  const uint32_t fp = q.tiles[t], ep = q.tiles[t + 1], b = q.tile_bundle[t];
  if (bstate(q, b)[S_PCG_ACTIVE] == 0) return;
  if (q.n_unstored != 0 && !stores_w(q.bundles[b])) return;
  const double* v = vp(q);
  const uint32_t* ps = point_start(q);
  const uint32_t s = ps[fp], e = ps[ep];
  const uint32_t my = fp + lane;
  const bool has = my < ep;
  const uint32_t mys = has ? ps[my] : 0, mye = has ? ps[my + 1] : 0;
  const double vf = SHARED ? v[6 * q.C + b] : 0.0;
  const uint32_t n = e - s;
  double acc[3] = {0, 0, 0};
  if (e - s <= TILE) {
    const uint32_t i = s + lane;
    const bool live = i < e;
    double w[K];
    if (live) {
      const uint32_t c = obs_camera(q)[i];
      double vc[6];
      for (int k = 0; k < 6; k++) vc[k] = v[c * 6 + k];
#pragma unroll
      for (int k = 0; k < K; k++) w[k] = wload(tile_word<WT>(q.Wp, K, s, n, k, i));
#pragma unroll
      for (int j = 0; j < 3; j++) {
        double a = 0;
#pragma unroll
        for (int k = 0; k < 6; k++) a += w[k * 3 + j] * vc[k];
        if (SHARED) a += w[18 + j] * vf;
        st[lane][j] = a;
      }
    }
    wave_sync();
    if (has) {
      segment_sum(st, s, mys, mye, acc);
      const double* pb = q.pblk + (size_t)my * 12;
      for (int j = 0; j < 3; j++)
        ys[lane][j] = pb[j * 3] * acc[0] + pb[j * 3 + 1] * acc[1] + pb[j * 3 + 2] * acc[2];
      for (uint32_t j = mys; j < mye; j++) own[j - s] = lane;
    }
    wave_sync();
    if (live) {
      const uint32_t o = own[lane];
      const double y0 = ys[o][0], y1 = ys[o][1], y2 = ys[o][2];
      double* out = q.tprod + (size_t)i * 8;
#pragma unroll
      for (int j = 0; j < NV; j++) out[j] = w[j * 3] * y0 + w[j * 3 + 1] * y1 + w[j * 3 + 2] * y2;
#pragma unroll
      for (int j = NV; j < 8; j++) out[j] = 0.0;  // whole 64-byte rows: full cache lines
    }
    return;
  }
  // one point (lane 0's): its sums chunk by chunk as k_point_sweep's, then its products
  for (uint32_t base = s; base < e; base += TILE) {
    const uint32_t i = base + lane;
    if (i < e) {
      const uint32_t c = obs_camera(q)[i];
      double vc[6];
      for (int k = 0; k < 6; k++) vc[k] = v[c * 6 + k];
#pragma unroll
      for (int j = 0; j < 3; j++) {
        double a = 0;
#pragma unroll
        for (int k = 0; k < 6; k++)
          a += wload(tile_word<WT>(q.Wp, K, s, n, k * 3 + j, i)) * vc[k];
        if (SHARED) a += wload(tile_word<WT>(q.Wp, K, s, n, 18 + j, i)) * vf;
        st[lane][j] = a;
      }
    }
    wave_sync();
    if (has) {
      const uint32_t lo = mys > base ? mys : base, hi = mye < base + TILE ? mye : base + TILE;
      segment_sum(st, base, lo, hi, acc);
    }
    wave_sync();
  }
  if (has) {
    const double* pb = q.pblk + (size_t)my * 12;
    for (int j = 0; j < 3; j++)
      ys[0][j] = pb[j * 3] * acc[0] + pb[j * 3 + 1] * acc[1] + pb[j * 3 + 2] * acc[2];
  }
  wave_sync();
  const double y0 = ys[0][0], y1 = ys[0][1], y2 = ys[0][2];
  for (uint32_t base = s; base < e; base += TILE) {
    const uint32_t i = base + lane;
    if (i >= e) continue;
    double* out = q.tprod + (size_t)i * 8;
#pragma unroll
    for (int j = 0; j < NV; j++)
      out[j] = wload(tile_word<WT>(q.Wp, K, s, n, j * 3, i)) * y0 +
               wload(tile_word<WT>(q.Wp, K, s, n, j * 3 + 1, i)) * y1 +
               wload(tile_word<WT>(q.Wp, K, s, n, j * 3 + 2, i)) * y2;
#pragma unroll
    for (int j = NV; j < 8; j++) out[j] = 0.0;
  }
}

template <class WT, bool SHARED>
__global__ __launch_bounds__(TILE* TILE_WAVES) void k_point_sweep_fused(Problem q) {
  __shared__ double st_all[TILE_WAVES][TILE][3];
  __shared__ double ys_all[TILE_WAVES][TILE][3];
  __shared__ uint32_t own_all[TILE_WAVES][TILE];
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / TILE), lane = threadIdx.x % TILE;
  const uint32_t t = point_tile(wave);
  if (t >= q.n_tiles) return;
  sweep_tile_fused<WT, SHARED>(q, t, lane, st_all[wave], ys_all[wave], own_all[wave]);
}

template <class WT, bool SHARED, int MODE, bool RECOMPUTE = false>
__global__ __launch_bounds__(TILE* TILE_WAVES) void k_point_sweep(Problem q) {
  __shared__ double st_all[TILE_WAVES][TILE][3];
  // the wave's index is uniform across its lanes: keep the tile's loads and branches scalar
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / TILE), lane = threadIdx.x % TILE;
  const uint32_t t = point_tile(wave);
  if (t >= q.n_tiles) return;
  sweep_tile<WT, SHARED, MODE, RECOMPUTE>(q, t, lane, st_all[wave]);
}

// The camera sweep: one workgroup per camera-list task streams the camera-major cross blocks and
// gathers y of each slot's point: part[t] = sum_n W_n y_p(n) (6) and sum_n w_fn . y_p(n) (focal).
// CERTIFY: for the bundles with an uncertified direct solution instead of those running PCG.
// With FP32 cross blocks the sweep is held to exactly five waves per SIMD (81 registers a lane):
// its loads are half as large, and more waves in flight beat more loads per thread (40.0 against
// 42.4 us per sweep at 500 cameras; five as a lower bound only, the compiler takes 96 registers
// and the sweep 51 us). With FP64 ones the compiler's own choice (four waves, 100 registers) is
// the faster (62 against 82 us when held to five).
// RECOMPUTE: the streamed route (each slot's cross block relinearised, relinearised_w, in place of
// the Wc words; the same slot order and sums).
template <class WT, bool SHARED, bool CERTIFY = false, bool RECOMPUTE = false>
__global__ __launch_bounds__(TASK_THREADS)
__attribute__((amdgpu_waves_per_eu(sizeof(WT) == 4 && !RECOMPUTE ? 5 : 1,
                                   sizeof(WT) == 4 && !RECOMPUTE ? 5 : 8)))
void k_camera_sweep(Problem q) {
  constexpr int NV = SHARED ? 7 : 6;
  __shared__ double red[TASK_THREADS / 64][NV];
  const uint32_t t = xcd_contiguous(blockIdx.x, gridDim.x);
  const uint32_t s = q.tasks[4 * t + 1], e = q.tasks[4 * t + 2], b = q.tasks[4 * t + 3];
  // the thread's (at most) 4 slots with no branch around a load: the point ids (loaded with the
  // bundle's state word), then the gathered y, so that each round of loads is in flight at once; a
  // slot past the task's end repeats its last slot with y scaled to zero (an exact zero term). A
  // camera without observations has one empty task.
  constexpr int PER = TASK_OBS / TASK_THREADS;
  double state = bstate(q, b)[CERTIFY ? S_DIRECT : S_PCG_ACTIVE];
  uint32_t slot[PER], pt[PER];
  const uint32_t last = e > s ? e - 1 : 0;
#pragma unroll
  for (int k = 0; k < PER; k++) {
    slot[k] = min(s + threadIdx.x + TASK_THREADS * k, last);
    pt[k] = e > s ? q.cam_point[slot[k]] : 0;
  }
  settle_uniform(state);
#pragma unroll
  for (int k = 0; k < PER; k++) settle(pt[k]);
  if (CERTIFY ? state != 1 : state == 0) return;
  if (q.n_unstored != 0 && !stores_w(q.bundles[b])) return;
  const double* y = vy(q);
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) acc[k] = 0;
  if (e > s) {
    double yv[PER][3];
#pragma unroll
    for (int k = 0; k < PER; k++) {
      const double f = s + threadIdx.x + TASK_THREADS * k < e ? 1.0 : 0.0;
#pragma unroll
      for (int c = 0; c < 3; c++) yv[k][c] = y[(size_t)pt[k] * 3 + c] * f;
    }
    if (RECOMPUTE) {
      const uint32_t c = q.tasks[4 * t];
      const Bundle& g = q.bundles[b];
      const double fscale = SHARED ? focal_scale(bstate(q, b)[S_LOG_FOCAL]) : 1.0;
      for (int k = 0; k < PER; k++) {
        const uint32_t n = slot[k];
        double wr[SHARED ? 21 : 18];
        relinearised_w<WT, SHARED>(q, g, c, pt[k], camera_obs(q)[n], fscale, wr);
#pragma unroll
        for (int j = 0; j < NV; j++)
          acc[j] += wr[j * 3] * yv[k][0] + wr[j * 3 + 1] * yv[k][1] + wr[j * 3 + 2] * yv[k][2];
      }
    } else {
#pragma unroll
      for (int k = 0; k < PER; k++) {
        const uint32_t n = slot[k];
#pragma unroll
        for (int j = 0; j < NV; j++)
          acc[j] += wload(wptr<WT, SHARED>(q.Wc, j * 3, n, q.M)) * yv[k][0] +
                    wload(wptr<WT, SHARED>(q.Wc, j * 3 + 1, n, q.M)) * yv[k][1] +
                    wload(wptr<WT, SHARED>(q.Wc, j * 3 + 2, n, q.M)) * yv[k][2];
      }
    }
  }
#pragma unroll
  for (int k = 0; k < NV; k++)
    for (int o = 32; o > 0; o >>= 1) acc[k] += __shfl_xor(acc[k], o, 64);
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < NV; k++) red[wave][k] = acc[k];
  __syncthreads();
  if (threadIdx.x < NV) {
    double v = 0;
    for (int w = 0; w < TASK_THREADS / 64; w++) v += red[w][threadIdx.x];
    q.task_part[(size_t)t * 8 + threadIdx.x] = v;
  }
}

// The camera side of the fused product: one workgroup per camera task sums the rows of tprod of
// its slots' observations (t_i = W_i y_p(i), written by k_point_sweep_fused) in k_camera_sweep's
// slot order and reduction, so part[t] is k_camera_sweep's.
template <bool SHARED>
__global__ __launch_bounds__(TASK_THREADS) void k_camera_tsum(Problem q) {
  constexpr int NV = SHARED ? 7 : 6;
  __shared__ double red[TASK_THREADS / 64][NV];
  const uint32_t t = xcd_contiguous(blockIdx.x, gridDim.x);
  const uint32_t s = q.tasks[4 * t + 1], e = q.tasks[4 * t + 2], b = q.tasks[4 * t + 3];
  if (bstate(q, b)[S_PCG_ACTIVE] == 0) return;
  if (q.n_unstored != 0 && !stores_w(q.bundles[b])) return;
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) acc[k] = 0;
  constexpr int PER = TASK_OBS / TASK_THREADS;
  if (e > s) {
    double tv[PER][NV];
#pragma unroll
    for (int k = 0; k < PER; k++) {
      const uint32_t n = min(s + threadIdx.x + TASK_THREADS * k, e - 1);
      const double* row = q.tprod + (size_t)camera_obs(q)[n] * 8;
#pragma unroll
      for (int j = 0; j < NV; j++) tv[k][j] = row[j];
    }
#pragma unroll
    for (int k = 0; k < PER; k++) {
      if (s + threadIdx.x + TASK_THREADS * k < e) {
#pragma unroll
        for (int j = 0; j < NV; j++) acc[j] += tv[k][j];
      }
    }
  }
#pragma unroll
  for (int k = 0; k < NV; k++)
    for (int o = 32; o > 0; o >>= 1) acc[k] += __shfl_xor(acc[k], o, 64);
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < NV; k++) red[wave][k] = acc[k];
  __syncthreads();
  if (threadIdx.x < NV) {
    double v = 0;
    for (int w = 0; w < TASK_THREADS / 64; w++) v += red[w][threadIdx.x];
    q.task_part[(size_t)t * 8 + threadIdx.x] = v;
  }
}

// The camera's sums of the camera sweep: sum over its tasks of part[0..NV-1], in task order.
template <int NV>
__device__ __forceinline__ void camera_sums(const Problem& q, uint32_t c, double part[NV]) {
#pragma unroll
  for (int j = 0; j < NV; j++) part[j] = 0;
  for (uint32_t t = q.task_start[c]; t < q.task_start[c + 1]; t++)
#pragma unroll
    for (int j = 0; j < NV; j++) part[j] += q.task_part[(size_t)t * 8 + j];
}

// Finish S·v of bundle g from the camera-task partials, one thread per camera:
//   (S v)_c    = A_c v_c [+ H_cf v_f] - sum of the camera's task partials [0..5]
//   focal row  = H_ff v_f + sum_c (H_cf . v_c - sum of the camera's task partials [6])
// Returns v . S v, the same value in every thread. Uniform control flow. Used outside the PCG loop
// (the refinement residual and the probe).
template <bool SHARED>
__device__ double schur_finish(const Problem& q, const Bundle& g, const double* s,
                               const double* v, double* sv, double* lds) {
  constexpr int NV = SHARED ? 7 : 6;
  const uint32_t tid = threadIdx.x;
  const double vf = SHARED ? v[g.focal_row] : 0.0;
  double dot = 0, fsum = 0;
  for (uint32_t c = g.cam0 + tid; c < g.cam1; c += blockDim.x) {
    const size_t o = 6 * (size_t)c;
    double part[NV];
    camera_sums<NV>(q, c, part);
    const double* A = q.camA + (size_t)c * 36;
    double fsh = 0;
#pragma unroll
    for (int j = 0; j < 6; j++) {
      double a = SHARED ? q.camHcf[o + j] * vf : 0.0, av = 0;
#pragma unroll
      for (int k = 0; k < 6; k++) av += A[j * 6 + k] * v[o + k];
      a += av;
      a -= part[j];
      sv[o + j] = a;
      dot += v[o + j] * a;
      if (SHARED) fsh += q.camHcf[o + j] * v[o + j];
    }
    if (SHARED) fsum += fsh - part[NV - 1];
  }
  if (!SHARED) return block_sum(dot, lds);
  block_sum2(dot, fsum, lds);
  const double sf = s[S_FOCAL_DIAG] * vf + fsum;
  if (tid == 0) sv[g.focal_row] = sf;
  return dot + vf * sf;
}

// One PCG iteration after the camera sweep (the rest of bundleSchurCameraRow b:204-225 and
// advanceBundleCg b:883-929 / advancePcg(false) j:874-924) in three launches over groups of at most
// GROUP_CAMS cameras of one bundle, one thread per row. No workgroup waits for another: each launch
// reduces the previous launch's per-group partials itself, in one fixed order, so every workgroup of
// a bundle derives the same scalars and takes the same branch; the bundle's first group writes the
// state words.
//   k_cg_product    (S p)_c = A_c p_c [+ H_cf p_f] - the camera's task partials; per group p . S p
//                   and the focal shares H_cf . p_c - part_c[6]
//   k_cg_step       focal row, p . S p, alpha and its tests; x += alpha p, r -= alpha S p,
//                   z = M^-1 r (the camera's 6x6 block); per group |r|^2 and r . z
//   k_cg_direction  |r|^2 and r . z, the convergence and breakdown tests, beta, p = z + beta p
// rho of the next iteration goes to word S_RHO_NEXT (the launch still reads word 18); k_cg_step
// reads it there and its first group copies it to word 18 for the next k_cg_direction. After the
// one-pass product, k_onepass_product finishes S p per camera and k_cg_step sums the cameras'
// p . S p (cam_dot) instead of k_cg_product's group partials.
struct Group {
  uint32_t bundle, c0, c1, g0, g1, row, rows;
  bool has, first;
};

// The PCG-step group of this workgroup: its index in the grid, or in the compacted grid's map.
__device__ __forceinline__ uint32_t pcg_group(const Problem& q) {
  return q.group_map ? q.group_map[blockIdx.x] : blockIdx.x;
}

// Group gi: one record of GROUP_WORDS words (bundle, first and end camera, the bundle's first and
// end group), loaded as one scalar block.
__device__ __forceinline__ Group group_at(const Problem& q, uint32_t gi) {
  Group G;
  const uint32_t* w = q.groups + (size_t)GROUP_WORDS * gi;
  G.bundle = w[0];
  G.c0 = w[1];
  G.c1 = w[2];
  G.g0 = w[3];
  G.g1 = w[4];
  G.rows = 6 * (G.c1 - G.c0);
  G.row = 6 * G.c0 + threadIdx.x;
  G.has = threadIdx.x < G.rows;
  G.first = gi == G.g0;
  return G;
}

// This thread's share of the sums over the bundle's groups of words w and w+1 of group_part (4 per
// group); block_sum2 completes them, in group order.
__device__ __forceinline__ void group_loads(const Problem& q, const Group& G, int w, double& a,
                                            double& b) {
  a = 0;
  b = 0;
  for (uint32_t k = G.g0 + threadIdx.x; k < G.g1; k += blockDim.x) {
    a += q.group_part[4 * (size_t)k + w];
    b += q.group_part[4 * (size_t)k + w + 1];
  }
}

// This thread's share of the sums over bundle g's cameras of their p . S p and focal shares
// (cam_dot); block_sum2 completes them, in camera order.
__device__ __forceinline__ void camera_loads(const Problem& q, const Bundle& g, double& a,
                                             double& b) {
  a = 0;
  b = 0;
  // eight cameras' loads issued together, then added in camera order
  constexpr int U = 8;
  for (uint32_t c0 = g.cam0 + threadIdx.x; c0 < g.cam1; c0 += U * blockDim.x) {
    double2 v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const uint32_t c = c0 + u * blockDim.x;
      v[u] = c < g.cam1 ? *reinterpret_cast<const double2*>(q.cam_dot + 2 * (size_t)c)
                        : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (c0 + u * blockDim.x < g.cam1) {
        a += v[u].x;
        b += v[u].y;
      }
    }
  }
}

// Each launch below issues all of its loads (state words, partials, its rows) before it tests the
// bundle's state: the chain of dependent loads from the group record to the first store is then
// two loads deep (three in k_cg_product, whose task partials follow the camera's task range)
// instead of one per test. Threads past the group's rows load row 6 c0 and store nothing.
// The LDS of the PCG-step bodies: a row exchange and the block sums' slots.
struct CgLds {
  double rr[ADV_THREADS];
  double lds[2 * (ADV_THREADS / 64)];
};

template <bool SHARED>
__device__ __forceinline__ void cg_product_group(const Problem& q, uint32_t gi, CgLds& L) {
  constexpr int NV = SHARED ? 7 : 6;
  double* lds = L.lds;
  const Group G = group_at(q, gi);
  double* s = bstate(q, G.bundle);
  const double* p = vp(q);
  const uint32_t row = G.has ? G.row : 6 * G.c0, c = row / 6, j = row - 6 * c;
  double active = s[S_PCG_ACTIVE];
  double pf = SHARED ? p[6 * q.C + G.bundle] : 0.0, hcf = SHARED ? q.camHcf[row] : 0.0;
  double pr = p[row];
  double A[6], pc[6], part[NV];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    A[k] = q.camA[(size_t)c * 36 + j * 6 + k];
    pc[k] = p[6 * (size_t)c + k];
  }
  camera_sums<NV>(q, c, part);
  settle(active);
  settle(pf);
  settle(hcf);
  settle(pr);
#pragma unroll
  for (int k = 0; k < 6; k++) {
    settle(A[k]);
    settle(pc[k]);
  }
#pragma unroll
  for (int k = 0; k < NV; k++) settle(part[k]);
  if (active == 0) return;
  double d = 0, fs = 0;
  if (G.has) {
    double a = SHARED ? hcf * pf : 0.0, av = 0;
#pragma unroll
    for (int k = 0; k < 6; k++) av += A[k] * pc[k];
    a += av;
    a -= part[j];
    vs(q)[G.row] = a;
    d = pr * a;
    if (SHARED) fs = hcf * pr - (j == 0 ? part[6] : 0.0);
  }
  block_sum2(d, fs, lds);
  if (threadIdx.x == 0) {
    q.group_part[4 * (size_t)gi] = d;
    q.group_part[4 * (size_t)gi + 1] = fs;
  }
}

template <bool SHARED>
__global__ __launch_bounds__(ADV_THREADS) void k_cg_product(Problem q) {
  __shared__ CgLds L;
  cg_product_group<SHARED>(q, pcg_group(q), L);
}

template <bool SHARED>
__device__ __forceinline__ void cg_step_group(const Problem& q, uint32_t gi, CgLds& L) {
  double* rr = L.rr;
  double* lds = L.lds;
  const Group G = group_at(q, gi);
  double* s = bstate(q, G.bundle);
  const uint32_t tid = threadIdx.x, f = 6 * q.C + G.bundle;
  const uint32_t row = G.has ? G.row : 6 * G.c0, c = row / 6, j = row - 6 * c;
  double* x = vx(q);
  double* r = vr(q);
  double* z = vz(q);
  double* sv = vs(q);
  const double* p = vp(q);
  double active = s[S_PCG_ACTIVE], rho = s[S_RHO_NEXT];
  double fdiag = SHARED ? s[S_FOCAL_DIAG] : 0.0, finv = SHARED ? s[S_FOCAL_INV] : 0.0;
  double dot, fsum;
  if (q.cam_dot)
    camera_loads(q, q.bundles[G.bundle], dot, fsum);
  else
    group_loads(q, G, 0, dot, fsum);
  double xr = x[row], pr = p[row], rv = r[row], svr = sv[row];
  double pf = SHARED ? p[f] : 0.0, xf = SHARED ? x[f] : 0.0, rf = SHARED ? r[f] : 0.0;
  double inv[6];
#pragma unroll
  for (int k = 0; k < 6; k++) inv[k] = q.camInv[(size_t)c * 36 + j * 6 + k];
  settle(active);
  settle(rho);
  settle(fdiag);
  settle(finv);
  settle(dot);
  settle(fsum);
  settle(xr);
  settle(pr);
  settle(rv);
  settle(svr);
  settle(pf);
  settle(xf);
  settle(rf);
#pragma unroll
  for (int k = 0; k < 6; k++) settle(inv[k]);
  if (active == 0) return;
  block_sum2(dot, fsum, lds);
  double den = dot, sf = 0;
  if (SHARED) {
    sf = fdiag * pf + fsum;
    den += pf * sf;
  }
  const double alpha = rho / den;
  const bool broke = !(den > 0) || !(fabs(alpha) < 1e20);
  if (G.first && tid == 0) {
    s[S_RHO] = rho;
    s[S_MATVECS] += 1;
    if (SHARED) sv[f] = sf;
    if (broke) {
      s[S_PCG_ACTIVE] = 0;
      s[S_PCG_FAILED] = 1;
    }
  }
  if (broke) return;
  double res = 0, rz = 0, ri = 0;
  if (G.has) {
    x[G.row] = xr + alpha * pr;
    ri = rv - alpha * svr;
    r[G.row] = ri;
    rr[tid] = ri;
    res = ri * ri;
  }
  __syncthreads();
  if (G.has) {
    const uint32_t base = tid - j;
    double zi = 0;
#pragma unroll
    for (int k = 0; k < 6; k++) zi += inv[k] * rr[base + k];
    z[G.row] = zi;
    rz = zi * ri;
  }
  if (SHARED && G.first && tid == 0) {
    x[f] = xf + alpha * pf;
    const double rn = rf - alpha * sf;
    r[f] = rn;
    res += rn * rn;
    const double zf = finv * rn;
    z[f] = zf;
    rz += zf * rn;
  }
  block_sum2(res, rz, lds);
  if (tid == 0) {
    q.group_part[4 * (size_t)gi + 2] = res;
    q.group_part[4 * (size_t)gi + 3] = rz;
  }
}

template <bool SHARED>
__global__ __launch_bounds__(ADV_THREADS) void k_cg_step(Problem q) {
  __shared__ CgLds L;
  cg_step_group<SHARED>(q, pcg_group(q), L);
}

template <bool SHARED>
__device__ __forceinline__ void cg_direction_group(const Problem& q, uint32_t gi, CgLds& L) {
  double* lds = L.lds;
  const Group G = group_at(q, gi);
  double* s = bstate(q, G.bundle);
  const uint32_t f = 6 * q.C + G.bundle, row = G.has ? G.row : 6 * G.c0;
  double* p = vp(q);
  const double* z = vz(q);
  double active = s[S_PCG_ACTIVE], rho_old = s[S_RHO], stop = s[S_PASS_STOP];
  double steps = s[S_STEPS], limit = s[S_STEP_LIMIT];
  double res, rho;
  group_loads(q, G, 2, res, rho);
  double pr = p[row], zr = z[row];
  double pf = SHARED ? p[f] : 0.0, zf = SHARED ? z[f] : 0.0;
  settle(active);
  settle(rho_old);
  settle(stop);
  settle(steps);
  settle(limit);
  settle(res);
  settle(rho);
  settle(pr);
  settle(zr);
  settle(pf);
  settle(zf);
  if (active == 0) return;
  block_sum2(res, rho, lds);
  const bool converged = res <= stop;
  const bool broke = !converged && !(rho > 0);
  // the pass's budget: the bundle stops unconverged, as the reference's host loop stops launching
  // at maxCgIterations (the gauge then rejects the step)
  const bool spent = !converged && !broke && steps + 1 >= limit;
  const double beta = rho / rho_old;
  if (G.first && threadIdx.x == 0) {
    s[S_CG_STEPS] += 1;
    s[S_STEPS] = steps + 1;
    s[S_R] = res;
    if (converged) {
      s[S_PCG_ACTIVE] = 0;
      s[S_CONVERGED] = 1;
    } else if (broke) {
      s[S_PCG_ACTIVE] = 0;
      s[S_PCG_FAILED] = 1;
    } else {
      s[S_RHO_NEXT] = rho;
      if (spent) s[S_PCG_ACTIVE] = 0;
    }
  }
  if (converged || broke) return;
  if (G.has) p[G.row] = zr + beta * pr;
  if (SHARED && G.first && threadIdx.x == 0) p[f] = zf + beta * pf;
}

template <bool SHARED>
__global__ __launch_bounds__(ADV_THREADS) void k_cg_direction(Problem q) {
  __shared__ CgLds L;
  cg_direction_group<SHARED>(q, pcg_group(q), L);
}

// The three steps for a bundle of one PCG-step group (every bundle of at most GROUP_CAMS cameras),
// one workgroup per bundle: no other workgroup contributes to its sums, so the launches' boundaries
// are the workgroup's barriers. The same bodies in the same order as the three launches, so the
// same bits; the barriers also order each step's global writes (the state words, group_part, S p,
// r, z) before the next step's reads, as a launch boundary does.
// PRODUCT false: after k_onepass_product, which has finished S p, the step and the direction only.
template <bool SHARED, bool PRODUCT = true>
__global__ __launch_bounds__(ADV_THREADS) void k_cg_single(Problem q) {
  __shared__ CgLds L;
  const uint32_t gi = q.single_map[blockIdx.x];
  if (PRODUCT) {
    cg_product_group<SHARED>(q, gi, L);
    __syncthreads();
  }
  cg_step_group<SHARED>(q, gi, L);
  __syncthreads();
  cg_direction_group<SHARED>(q, gi, L);
}

// The linear-system probe (bundle 0): S v, M^-1 v and the RHS of the assembled system.
template <bool SHARED>
__global__ __launch_bounds__(CG_THREADS) void k_probe(Problem q, const double* v, double* sv,
                                                      double* zv, double* rhs) {
  __shared__ double lds[2 * (CG_THREADS / 64)];
  const Bundle& g = q.bundles[0];
  const double* s = bstate(q, 0);
  (void)schur_finish<SHARED>(q, g, s, v, sv, lds);
  for (uint32_t row = 6 * g.cam0 + threadIdx.x; row < 6 * g.cam1; row += CG_THREADS) {
    zv[row] = block_row(q.camInv, row, v);
    rhs[row] = q.camRhs[row];
  }
  if (SHARED && threadIdx.x == 0) {
    zv[g.focal_row] = s[S_FOCAL_INV] * v[g.focal_row];
    rhs[g.focal_row] = s[S_FOCAL_RHS];
  }
}

// ------------------------------------------------------------------ implicit FP64 operator

// Which bundles a sweep of the implicit operator covers.
// GUARD_PCG_UNSTORED: the PCG sweeps of the bundles whose cross blocks are not stored;
// GUARD_UPDATE (point sweep only): the candidate's points of those bundles, as
// k_point_sweep<SWEEP_UPDATE> does for the others.
enum SweepGuard {
  GUARD_NONE = 0, GUARD_PCG = 1, GUARD_DIRECT = 2, GUARD_PCG_UNSTORED = 3, GUARD_UPDATE = 4
};

__device__ __forceinline__ bool skip_bundle(const Bundle& g, const double* sb, int guard) {
  if (guard == GUARD_PCG) return sb[S_PCG_ACTIVE] == 0;
  if (guard == GUARD_DIRECT) return sb[S_DIRECT] != 1;
  if (guard == GUARD_PCG_UNSTORED) return sb[S_PCG_ACTIVE] == 0 || stores_w(g);
  if (guard == GUARD_UPDATE) return sb[S_ACTIVE] == 0 || sb[S_PCG_FAILED] != 0 || stores_w(g);
  return false;
}

// The two Schur sweeps with every cross block relinearised from the scene in FP64 instead of read
// from storage: (W^T v)_j = w (Jp_u[j] (Jc_u . v_c + Jf_u v_f) + Jp_v[j] (Jc_v . v_c + Jf_v v_f))
// and (W y)_k = w (Jc_u[k] (Jp_u . y) + Jc_v[k] (Jp_v . y)). The mixed form computes its
// refinement residual b - S x with this operator, so FP64 cross blocks never need storing.
// `guard`: GUARD_PCG skips bundles whose PCG is inactive (the PCG use), GUARD_DIRECT those without
// an uncertified direct solution; GUARD_NONE (the refinement) runs every bundle.
template <bool SHARED>
__global__ __launch_bounds__(TILE* TILE_WAVES) void k_point_sweep_implicit(Problem q,
                                                                          const double* v,
                                                                          int guard) {
  __shared__ double st_all[TILE_WAVES][TILE][3];
  // the wave's index is uniform across its lanes: keep the tile's loads and branches scalar
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / TILE), lane = threadIdx.x % TILE;
  const uint32_t t = point_tile(wave);
  if (t >= q.n_tiles) return;
  const uint32_t b = q.tile_bundle[t];
  const double* sb = bstate(q, b);
  const Bundle& g = q.bundles[b];
  if (skip_bundle(g, sb, guard)) return;
  auto& st = st_all[wave];
  const uint32_t fp = q.tiles[t], ep = q.tiles[t + 1];
  const uint32_t* ps = point_start(q);
  const uint32_t s = ps[fp], e = ps[ep];
  const uint32_t my = fp + lane;
  const bool has = my < ep;
  const uint32_t mys = has ? ps[my] : 0, mye = has ? ps[my + 1] : 0;
  const double vf = SHARED ? v[g.focal_row] : 0.0;
  const double fscale = SHARED ? focal_scale(sb[S_LOG_FOCAL]) : 1.0;
  double acc[3] = {0, 0, 0};
  for (uint32_t base = s; base < e; base += TILE) {
    const uint32_t i = base + lane;
    if (i < e) {
      const uint32_t c = obs_camera(q)[i], p = obs_point(q)[i];
      Lin L;
      double tj[3] = {0, 0, 0};
      if (linearize(q, c, p, i, fscale, variable_camera(q, c), directional_point(g, p), L)) {
        double a0 = 0, a1 = 0;
        for (int k = 0; k < 6; k++) {
          a0 += L.Jc[k] * v[c * 6 + k];
          a1 += L.Jc[6 + k] * v[c * 6 + k];
        }
        if (SHARED) {
          a0 += L.Jf[0] * vf;
          a1 += L.Jf[1] * vf;
        }
        for (int j = 0; j < 3; j++) tj[j] = L.w * (L.Jp[j] * a0 + L.Jp[3 + j] * a1);
      }
      st[lane][0] = tj[0];
      st[lane][1] = tj[1];
      st[lane][2] = tj[2];
    }
    wave_sync();
    if (has) {
      const uint32_t lo = mys > base ? mys : base, hi = mye < base + TILE ? mye : base + TILE;
      segment_sum(st, base, lo, hi, acc);
    }
    wave_sync();
  }
  if (!has) return;
  const double* pb = q.pblk + (size_t)my * 12;
  if (guard == GUARD_UPDATE) {
    const double r0 = pb[9] - acc[0], r1 = pb[10] - acc[1], r2 = pb[11] - acc[2];
    const size_t at = (size_t)q.C * 12 + (size_t)my * 3;
    for (int j = 0; j < 3; j++)
      q.trial[at + j] = q.scene[at + j] + (pb[j * 3] * r0 + pb[j * 3 + 1] * r1 + pb[j * 3 + 2] * r2);
    return;
  }
  double* y = vy(q) + (size_t)my * 3;
  for (int j = 0; j < 3; j++) y[j] = pb[j * 3] * acc[0] + pb[j * 3 + 1] * acc[1] + pb[j * 3 + 2] * acc[2];
}

// The implicit camera sweep: relinearises each slot of the task instead of reading its stored
// cross block. `guard` as for the implicit point sweep.
template <bool SHARED>
__global__ __launch_bounds__(TASK_THREADS) void k_camera_sweep_implicit(Problem q, int guard) {
  constexpr int NV = SHARED ? 7 : 6;
  __shared__ double red[TASK_THREADS / 64][NV];
  const uint32_t t = blockIdx.x;
  const uint32_t c = q.tasks[4 * t], s = q.tasks[4 * t + 1], e = q.tasks[4 * t + 2];
  const uint32_t b = q.camera_bundle[c];
  const double* sb = bstate(q, b);
  const Bundle& g = q.bundles[b];
  if (skip_bundle(g, sb, guard)) return;
  const bool var = variable_camera(q, c);
  const double fscale = SHARED ? focal_scale(sb[S_LOG_FOCAL]) : 1.0;
  const double* y = vy(q);
  double acc[NV];
  for (int k = 0; k < NV; k++) acc[k] = 0;
  for (uint32_t n = s + threadIdx.x; n < e; n += TASK_THREADS) {
    const uint32_t i = camera_obs(q)[n], p = q.cam_point[n];
    Lin L;
    if (!linearize(q, c, p, i, fscale, var, directional_point(g, p), L)) continue;
    const double y0 = y[(size_t)p * 3], y1 = y[(size_t)p * 3 + 1], y2 = y[(size_t)p * 3 + 2];
    const double u0 = L.w * (L.Jp[0] * y0 + L.Jp[1] * y1 + L.Jp[2] * y2);
    const double u1 = L.w * (L.Jp[3] * y0 + L.Jp[4] * y1 + L.Jp[5] * y2);
    for (int k = 0; k < 6; k++) acc[k] += L.Jc[k] * u0 + L.Jc[6 + k] * u1;
    if (SHARED) acc[6] += L.Jf[0] * u0 + L.Jf[1] * u1;
  }
  for (int k = 0; k < NV; k++)
    for (int o = 32; o > 0; o >>= 1) acc[k] += __shfl_xor(acc[k], o, 64);
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < NV; k++) red[wave][k] = acc[k];
  __syncthreads();
  if (threadIdx.x < NV) {
    double v = 0;
    for (int w = 0; w < TASK_THREADS / 64; w++) v += red[w][threadIdx.x];
    q.task_part[(size_t)t * 8 + threadIdx.x] = v;
  }
}

// The refinement step of the mixed form, per bundle, after the implicit sweeps of x: the FP64
// residual r = b - S x replaces the recurrence residual. When |r|^2 <= refine_tol^2 |r0|^2 the
// solve is done (words 22, 24); otherwise PCG restarts from x with z = M^-1 r, p = z (18, 19).
// Runs only for a bundle whose PCG pass converged (word 24 set, word 20 clear).
template <bool SHARED>
__global__ __launch_bounds__(CG_THREADS) void k_refine(Problem q) {
  __shared__ double lds[2 * (CG_THREADS / 64)];
  const uint32_t bi = blockIdx.x;
  const Bundle& g = q.bundles[bi];
  double* s = bstate(q, bi);
  const double r0 = s[S_R0];
  const double active = s[S_ACTIVE];
  const bool setup_failed = s[S_FAIL_TILES] != 0 || s[S_FAIL_CAMERAS] != 0 || s[S_FAIL_FOCAL] != 0;
  // a pass that converged on its own recurrence, or broke down (p.Sp or r.z no longer positive,
  // which the FP32 operator's rounding can bring about near convergence), is judged by the FP64
  // residual of its last iterate. A pass that ran out of budget is judged as the FP64 form judges
  // it (the gauge rejects the step), and a failure at the PCG set-up (a block that is not positive
  // definite) stays. A certified direct solution, and a bundle whose PCG ran on the FP64 implicit
  // operator (no stored cross blocks), have an FP64 residual already.
  const bool ended = s[S_CONVERGED] != 0 || (s[S_PCG_FAILED] != 0 && !setup_failed);
  if (active == 0 || !ended || s[S_DIRECT] == 2 || !stores_w(g)) return;
  const uint32_t tid = threadIdx.x;
  double* x = vx(q);
  double* r = vr(q);
  double* p = vp(q);
  double* S = vs(q);
  double* z = vz(q);
  (void)schur_finish<SHARED>(q, g, s, x, S, lds);
  __syncthreads();
  const uint32_t r_lo = 6 * g.cam0, r_hi = 6 * g.cam1;
  double res = 0;
  for (uint32_t i = r_lo + tid; i < r_hi; i += CG_THREADS) {
    const double ri = q.camRhs[i] - S[i];
    r[i] = ri;
    res += ri * ri;
  }
  if (SHARED && tid == 0) {
    const uint32_t f = g.focal_row;
    r[f] = s[S_FOCAL_RHS] - S[f];
    res += r[f] * r[f];
  }
  res = block_sum(res, lds);
  if (res <= q.refine_tol * q.refine_tol * r0) {
    if (tid == 0) {
      s[S_R] = res;
      s[S_PCG_FAILED] = 0;
    }
    return;
  }
  __syncthreads();  // every row of r before the block products
  double rho = 0;
  for (uint32_t i = r_lo + tid; i < r_hi; i += CG_THREADS) {
    const double zi = block_row(q.camInv, i, r);
    z[i] = zi;
    p[i] = zi;
    rho += zi * r[i];
  }
  if (SHARED && tid == 0) {
    const uint32_t f = g.focal_row;
    z[f] = s[S_FOCAL_INV] * r[f];
    p[f] = z[f];
    rho += z[f] * r[f];
  }
  rho = block_sum(rho, lds);
  if (tid == 0) {
    s[S_R] = res;
    s[S_RHO] = rho;
    s[S_RHO_NEXT] = rho;
    s[S_CONVERGED] = 0;
    s[S_REFINEMENTS] += 1;
    s[S_STEP_LIMIT] = s[S_STEPS] + pcg_budget(q, g);  // each restart: the budget again
    // the restart solves S d = r on the FP32 operator: to the PCG tolerance relative to this
    // residual, and no tighter than a tenth of the refinement target (the next FP64 residual
    // then lands below the target unless the operator's own error stops it)
    s[S_PASS_STOP] = fmax(q.tol * q.tol * res, 0.01 * q.refine_tol * q.refine_tol * r0);
    if (rho > 0) {
      s[S_PCG_ACTIVE] = 1;
      s[S_PCG_FAILED] = 0;
    } else {
      s[S_PCG_FAILED] = 1;
    }
  }
}

// ------------------------------------------------------------------ candidate

// Cameras of the candidate: movePose for variable cameras, the trial log focal on the shared route.
__global__ void k_update_cameras(Problem q) {
  const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= q.C) return;
  const uint32_t b = q.camera_bundle[h];
  double* s = bstate(q, b);
  if (s[S_ACTIVE] == 0 || s[S_PCG_FAILED] != 0) return;
  const Bundle& g = q.bundles[b];
  const double* x = vx(q);
  if (h == g.cam0 && q.shared) s[S_TRIAL_FOCAL] = s[S_LOG_FOCAL] + x[g.focal_row];
  double pose[12], moved[12];
  for (int j = 0; j < 12; j++) pose[j] = q.scene[(size_t)h * 12 + j];
  if (variable_camera(q, h)) {
    double d[6];
    for (int j = 0; j < 6; j++) d[j] = x[h * 6 + j];
    move_pose(pose, d, moved);
  } else {
    for (int j = 0; j < 12; j++) moved[j] = pose[j];
  }
  for (int j = 0; j < 12; j++) q.trial[(size_t)h * 12 + j] = moved[j];
}

// bundleCandidateGauge, per bundle: reject an unconverged solve, check the focal proposal, the
// baseline, the scale and every trial pose; on success set the gauge scale and the trial-valid
// word.
__global__ __launch_bounds__(256) void k_gauge(Problem q) {
  __shared__ double lds[4];
  const uint32_t bi = blockIdx.x;
  const Bundle& g = q.bundles[bi];
  double* s = bstate(q, bi);
  const uint32_t tid = threadIdx.x;
  const double active = s[S_ACTIVE], failed = s[S_PCG_FAILED], r0 = s[S_R0], r = s[S_R];
  const double trial_focal = s[S_TRIAL_FOCAL];
  __syncthreads();
  if (tid == 0) s[S_TRIAL_VALID] = 0;
  if (active == 0 || failed != 0) return;
  if (r0 > 0 && r > q.gauge_tol * q.gauge_tol * r0) {
    if (tid == 0) s[S_PCG_FAILED] = 1;
    return;
  }
  double bad = 0;
  if (q.shared) {
    if (!(fabs(trial_focal) < 700)) return;
    const double fs = focal_scale(trial_focal);
    for (uint32_t c = g.cam0 + tid; c < g.cam1; c += 256) {
      const double* m = q.scene + (size_t)q.C * 12 + (size_t)q.P * 3 + (size_t)c * 6;
      if (!(m[0] * fs >= 1e-6 && m[1] * fs >= 1e-6 && m[0] * fs <= 1e6 && m[1] * fs <= 1e6)) bad += 1;
    }
  }
  for (uint32_t c = g.cam0 + tid; c < g.cam1; c += 256) {
    const double* p = q.trial + (size_t)c * 12;
    for (int j = 0; j < 12; j++)
      if (!(fabs(p[j]) <= 1e9)) bad += 1;
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++) {
        double d = 0;
        for (int l = 0; l < 3; l++) d += p[j * 3 + l] * p[k * 3 + l];
        if (j == k) d -= 1;
        if (!(fabs(d) < 1e-6)) bad += 1;
      }
  }
  bad = block_sum(bad, lds);
  if (tid != 0) return;
  double c[3];
  center_of(q.trial + (size_t)g.scale_view * 12, c);
  const double d0 = s[S_ORIGIN] - c[0], d1 = s[S_ORIGIN + 1] - c[1], d2 = s[S_ORIGIN + 2] - c[2];
  const double baseline = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
  if (!(baseline > 1e-12)) return;
  double scale = 1;
  if (g.normalize_scale) scale = s[S_BASELINE] / baseline;
  if (!(scale > 1e-9 && scale < 1e9)) return;
  if (bad > 0) return;
  s[S_GAUGE] = scale;
  s[S_TRIAL_VALID] = 1;
}

// normalizeBundleCandidate: unit directional points; the similarity X' = o + s (X - o),
// t' = s t + (s - 1) R o about the gauge origin when one camera is fixed. Thread h handles point h
// and camera h.
__global__ void k_normalize(Problem q) {
  const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h < q.P) {
    const uint32_t b = q.point_bundle[h];
    const Bundle& g = q.bundles[b];
    const double* s = bstate(q, b);
    if (s[S_ACTIVE] != 0 && s[S_TRIAL_VALID] != 0) {
      double* X = q.trial + (size_t)q.C * 12 + (size_t)h * 3;
      if (directional_point(g, h)) {
        const double norm = sqrt(X[0] * X[0] + X[1] * X[1] + X[2] * X[2]);
        if (norm > 1e-12) {
          X[0] = X[0] / norm;
          X[1] = X[1] / norm;
          X[2] = X[2] / norm;
        }
      } else if (g.normalize_scale) {
        const double sc = s[S_GAUGE];
        for (int j = 0; j < 3; j++) X[j] = s[S_ORIGIN + j] + sc * (X[j] - s[S_ORIGIN + j]);
      }
    }
  }
  if (h < q.C) {
    const uint32_t b = q.camera_bundle[h];
    const Bundle& g = q.bundles[b];
    const double* s = bstate(q, b);
    if (s[S_ACTIVE] != 0 && s[S_TRIAL_VALID] != 0 && g.normalize_scale && h != g.anchor) {
      const double sc = s[S_GAUGE];
      const double o[3] = {s[S_ORIGIN], s[S_ORIGIN + 1], s[S_ORIGIN + 2]};
      double* p = q.trial + (size_t)h * 12;
      double rot[3];
      for (int r = 0; r < 3; r++)
        rot[r] = (p[r * 3] * o[0] + p[r * 3 + 1] * o[1] + p[r * 3 + 2] * o[2] + p[9 + r]) - p[9 + r];
      const double t0 = sc * p[9] + (sc - 1) * rot[0];
      const double t1 = sc * p[10] + (sc - 1) * rot[1];
      const double t2 = sc * p[11] + (sc - 1) * rot[2];
      p[9] = t0;
      p[10] = t1;
      p[11] = t2;
    }
  }
}

// commitBundleCandidate: the trial becomes the scene for every bundle that accepted it.
__global__ void k_commit(Problem q) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t cams = (size_t)q.C * 12, total = cams + (size_t)q.P * 3;
  if (i >= total) return;
  const uint32_t b = i < cams ? q.camera_bundle[i / 12] : q.point_bundle[(i - cams) / 3];
  double* s = bstate(q, b);
  if (s[S_COMMIT] == 0) return;
  q.scene[i] = q.trial[i];
  if (q.shared && i == (size_t)q.bundles[b].cam0 * 12) s[S_LOG_FOCAL] = s[S_TRIAL_FOCAL];
}

// ------------------------------------------------------------------ Newton polish and adjoint

// A Newton iteration of the polish, per bundle: undamped (lambda 0) on the exact Hessian.
__global__ void k_newton_begin(Problem q) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= q.B) return;
  double* s = bstate(q, b);
  s[S_LAMBDA] = 0;
  s[S_ACTIVE] = s[S_INPUT_ERROR] == 0 ? 1.0 : 0.0;
  s[S_COMMIT] = 0;
}

// The Newton step is kept whenever the gauge admitted it: there is no cost test, since near the
// solution the cost change falls below the rounding of the cost itself.
__global__ void k_newton_accept(Problem q) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= q.B) return;
  double* s = bstate(q, b);
  const bool take = s[S_ACTIVE] != 0 && s[S_TRIAL_VALID] != 0;
  s[S_COMMIT] = take ? 1.0 : 0.0;
  if (take) s[S_ACCEPTED] += 1;
}

// The pixel cotangent of the adjoint, one thread per observation: with lambda the solution of
// H lambda = v (cameras in x, the focal row in x, points in y),
//   d<v, theta*>/d u_i = D_i J_i lambda,
// D_i = d(w r)/dr at the observation (the implicit function theorem on sum_i J_i^T w_i r_i = 0 with
// r_i = pixel_i(theta) - u_i). An observation that does not project gets 0. `focal` (optional, M):
// the observation's share of the input-focal cotangent, -(D_i J_i lambda) . Jf_i - (J_i lambda) .
// (w_i r_i) (the second term in the exact mode; d J_i / d log f = J_i). Jf_i = pixel_i - (cx, cy)
// is the reference's d(pixel)/d(log scale) (src/sfm/bundle/joint-focal.ts:150-154).
template <bool SHARED>
__global__ void k_pixel_cotangent(Problem q, double* out, double* focal) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= q.M) return;
  const uint32_t c = obs_camera(q)[i], p = obs_point(q)[i];
  const uint32_t b = q.camera_bundle[c];
  const double* sb = bstate(q, b);
  const Bundle& g = q.bundles[b];
  const double fscale = SHARED ? focal_scale(sb[S_LOG_FOCAL]) : 1.0;
  Lin L;
  double t0 = 0, t1 = 0, share = 0;
  if (linearize(q, c, p, i, fscale, variable_camera(q, c), directional_point(g, p), L)) {
    const double* x = vx(q);
    const double* y = vy(q) + (size_t)p * 3;
    for (int k = 0; k < 6; k++) {
      t0 += L.Jc[k] * x[(size_t)c * 6 + k];
      t1 += L.Jc[6 + k] * x[(size_t)c * 6 + k];
    }
    for (int k = 0; k < 3; k++) {
      t0 += L.Jp[k] * y[k];
      t1 += L.Jp[3 + k] * y[k];
    }
    if (SHARED) {
      t0 += L.Jf[0] * x[g.focal_row];
      t1 += L.Jf[1] * x[g.focal_row];
    }
    const double j0 = t0, j1 = t1;  // J_i lambda
    const double e = sqrt(L.r[0] * L.r[0] + L.r[1] * L.r[1]);
    if (e > q.huber) {
      const double h0 = L.r[0] / e, h1 = L.r[1] / e, d = h0 * t0 + h1 * t1;
      t0 = L.w * (t0 - h0 * d);
      t1 = L.w * (t1 - h1 * d);
    }
    share = -(t0 * L.Jf[0] + t1 * L.Jf[1]);
    if (!q.gauss_newton) share -= j0 * (L.w * L.r[0]) + j1 * (L.w * L.r[1]);
  }
  out[2 * (size_t)i] = t0;
  out[2 * (size_t)i + 1] = t1;
  if (focal) focal[i] = share;
}

// The input-focal cotangent of each camera: the shares of its observations summed in its list's
// order (one workgroup per camera; strided partial sums, then the workgroup's fixed-order sum).
__global__ __launch_bounds__(256) void k_focal_cotangent(Problem q, const double* share,
                                                         double* out) {
  __shared__ double lds[4];
  const uint32_t c = blockIdx.x;
  const uint32_t s = camera_start(q)[c], e = camera_start(q)[c + 1];
  double sum = 0;
  for (uint32_t n = s + threadIdx.x; n < e; n += blockDim.x) sum += share[camera_obs(q)[n]];
  sum = block_sum(sum, lds);
  if (threadIdx.x == 0) out[c] = sum;
}

// The per-observation linearisation (the Jacobian kernels' output): 23 words per observation,
// r (2), d(u,v)/d(omega, v) (12, row u then row v), d(u,v)/dX (6), d(u,v)/d log f (2), w; all
// zero, w = -1, for an observation that does not project.
template <bool SHARED>
__global__ void k_jacobian(Problem q, double* out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= q.M) return;
  const uint32_t c = obs_camera(q)[i], p = obs_point(q)[i];
  const uint32_t b = q.camera_bundle[c];
  const double* sb = bstate(q, b);
  const Bundle& g = q.bundles[b];
  const double fscale = SHARED ? focal_scale(sb[S_LOG_FOCAL]) : 1.0;
  Lin L;
  double* o = out + (size_t)i * 23;
  if (!linearize(q, c, p, i, fscale, variable_camera(q, c), directional_point(g, p), L)) {
    for (int k = 0; k < 22; k++) o[k] = 0;
    o[22] = -1;
    return;
  }
  o[0] = L.r[0];
  o[1] = L.r[1];
  for (int k = 0; k < 12; k++) o[2 + k] = L.Jc[k];
  for (int k = 0; k < 6; k++) o[14 + k] = L.Jp[k];
  o[20] = L.Jf[0];
  o[21] = L.Jf[1];
  o[22] = L.w;
}

// packBundleResult: coordinates of the whole batch, then per bundle its state words with the
// RMSEs (and the log focal scale).
__global__ void k_pack(Problem q, double* out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t total = (size_t)q.C * 12 + (size_t)q.P * 3;
  const uint32_t words = 26 + q.shared;
  if (i < total) out[i] = q.scene[i];
  if (i < (size_t)q.B * words) {
    const uint32_t b = (uint32_t)(i / words), w = (uint32_t)(i % words);
    const Bundle& g = q.bundles[b];
    const double* s = bstate(q, b);
    const double M = (double)(float)(g.obs1 - g.obs0);
    double v = w < 26 ? s[w] : s[S_LOG_FOCAL];
    if (w == 24) v = sqrt(s[S_SSE0] / M);
    if (w == 25) v = sqrt(s[S_SSE] / M);
    out[total + i] = v;
  }
}

}  // namespace tcvba
