// ba_kernels.h — the gfx950 kernels of one LM iteration of the bundle adjustment.
//
// Data layout in HBM (C cameras, P points, M observations, K = 18 or 21 cross words):
//   scene, trial, index    the reference's packed arrays (include/typecv_ba.h)
//   Wp[K][M]               cross blocks W_i = w Jc^T Jp (+ the focal row w Jf^T Jp), one component
//                          per row, observations in point-major (input) order: the point sweep
//   Wc[K][M]               the same blocks in camera-major order (the order of the reference's
//                          camera lists): the camera sweep streams them
//   camPoint[M]            the point of each camera-major slot
//   pblk[P][12]            V^-1 (9, symmetrised) and the point RHS b_p (3)
//   tiles[T+1]             first point of each wave tile: whole points, at most 64 observations
//                          (a point with more observations is a tile of its own)
//   tasks[3·N], taskStart  camera-list tasks (camera, first slot, end slot), at most `task_obs`
//                          slots each, and the first task of each camera
// Every reduction has a fixed order, so a run is bitwise reproducible; no float atomics.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ba_math.h"

namespace tcvba {

// State words: the reference's 26 (BUNDLE_STATE) and the device's own focal words.
enum : int {
  S_LAMBDA = 0, S_ACTIVE = 1, S_BASELINE = 2, S_ORIGIN = 3, S_COST = 6, S_SSE = 7,
  S_COST0 = 8, S_SSE0 = 9, S_ACCEPTED = 10, S_REJECTED = 11, S_COMMIT = 12, S_GAUGE = 13,
  S_INPUT_ERROR = 14, S_TRIAL_VALID = 15, S_CG_STEPS = 16, S_FAILED = 17, S_RHO = 18,
  S_PCG_ACTIVE = 19, S_PCG_FAILED = 20, S_R0 = 21, S_R = 22, S_STEPS = 23, S_CONVERGED = 24,
  // device-only words (joint-focal.shade.ts keeps them in `focal[6C + 0..5]`)
  S_LOG_FOCAL = 26,      // accepted log focal scale (focal[6C+3])
  S_TRIAL_FOCAL = 27,    // trial log focal scale (focal[6C+4])
  S_FOCAL_DIAG = 28,     // damped focal diagonal H_ff (focal[6C])
  S_FOCAL_INV = 29,      // focal preconditioner 1 / reduced diagonal, or -1 (focal[6C+1])
  S_FOCAL_RHS = 30,      // reduced focal RHS (focal[6C+2])
  S_MATVECS = 31,        // Schur products that ran while PCG was active (whole solve)
  S_REFINEMENTS = 32,    // mixed form: refinement restarts taken (whole solve)
  S_WORDS = 40
};

constexpr int TILE = 64;           // observations per wave tile
#ifndef TCV_TILE_WAVES
#define TCV_TILE_WAVES 4
#endif
#ifndef TCV_W_BLOCKED
#define TCV_W_BLOCKED 1
#endif
#ifndef TCV_W_NONTEMPORAL
#define TCV_W_NONTEMPORAL 1
#endif
constexpr int TILE_WAVES = TCV_TILE_WAVES;  // wave tiles per workgroup of the point-tile kernels
constexpr int TASK_THREADS = 256;  // threads of a camera-task workgroup
constexpr int CG_THREADS = 1024;   // threads of the single-workgroup PCG kernels
constexpr int ASM_WORDS = 56;      // per-task camera assembly sums (see camera_pass)

struct Problem {
  uint32_t C, P, M, m, K, shared, first_directional, anchor, scale_view, normalize_scale;
  uint32_t n_tiles, n_tasks;
  double huber, tol;
  double refine_tol;  // mixed form: relative FP64 residual of the outer refinement
  double* scene;
  double* trial;
  const uint32_t* index;
  const uint32_t* tiles;
  const uint32_t* tasks;
  const uint32_t* task_start;
  uint32_t* cam_point;
  void* Wp;  // WT[K][M]
  void* Wc;  // WT[K][M]
  double* pblk;
  uint32_t* tile_fail;
  double* tile_focal;  // 2 per tile
  double* task_part;   // ASM_WORDS per task
  double* camA;        // 36 per camera
  double* camInv;      // 36 per camera
  double* camRhs;      // 6 per camera
  double* camHcf;      // 6 per camera
  double* camFocal;    // 2 per camera: H_ff and b_f shares
  double* vec;         // x r p s z (m each), y (3P), fpart (C)
  double* state;
  double* cost_part;   // 3 per cost block
  uint32_t* error;     // index validation flag
  uint32_t n_cost_blocks;
};

__host__ __device__ inline const uint32_t* obs_camera(const Problem& q) { return q.index; }
__host__ __device__ inline const uint32_t* obs_point(const Problem& q) { return q.index + q.M; }
__host__ __device__ inline const uint32_t* point_start(const Problem& q) {
  return q.index + 2 * (size_t)q.M;
}
__host__ __device__ inline const uint32_t* camera_start(const Problem& q) {
  return q.index + 2 * (size_t)q.M + q.P + 1;
}
__host__ __device__ inline const uint32_t* camera_obs(const Problem& q) {
  return q.index + 2 * (size_t)q.M + q.P + q.C + 2;
}
__host__ __device__ inline const uint32_t* camera_active(const Problem& q) {
  return q.index + 3 * (size_t)q.M + q.P + q.C + 2;
}

}  // namespace tcvba
