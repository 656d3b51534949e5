// ba_kernels.h — the gfx950 kernels of one LM iteration of the bundle adjustment.
//
// A launch covers a batch of B independent bundles concatenated into one block-diagonal problem:
// global cameras, points and observations are the bundles' own in order, so every per-observation,
// per-point, per-camera, per-tile and per-task kernel runs over all bundles at once, and each
// bundle keeps its own LM and PCG state words (its own damping, accept/reject and convergence). A
// single solve is a batch of one.
//
// Data layout in HBM (C cameras, P points, M observations in all, K = 18 or 21 cross words):
//   scene, trial, index    the reference's packed arrays (include/typecv_ba.h) of the whole batch
//   Wp                     cross blocks W_i = w Jc^T Jp (+ the focal row w Jf^T Jp), observations in
//                          point-major (input) order, each tile's as K rows of its observations'
//                          words (tile_word): the point sweep and the one-pass product stream it
//   Wc                     the same blocks in camera-major order (each camera's list): the camera
//                          sweep streams it. Both blocked [slots/64][K][64] (see wptr)
//   cam_point[M]           the point of each camera-major slot
//   pblk[P][12]            V^-1 (9, symmetrised) and the point RHS b_p (3)
//   tiles[T+1]             first point of each wave tile: whole points of one bundle, at most 64
//                          observations (a point with more observations is a tile of its own)
//   tasks[4·N], task_start camera-list tasks (camera, first slot, end slot, bundle), at most 1024
//                          slots, and the first task of each camera
//   rows                   camera c owns unknowns 6c..6c+5; bundle b's focal unknown is 6C + b
// Every reduction has a fixed order, so a run is bitwise reproducible: no global float atomics,
// and the one-pass product's LDS adds reach each row one lane at a time in a fixed order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ba_math.h"

namespace tcvba {

// State words of each bundle: the reference's 26 (BUNDLE_STATE) and the device's own.
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
  S_RHO_NEXT = 33,       // rho of the next PCG iteration (becomes word 18 at its first launch)
  S_FAIL_TILES = 34,     // point tiles with a failed point block at the last PCG set-up
  S_FAIL_CAMERAS = 35,   // cameras with a failed preconditioner block at the last PCG set-up
  S_FAIL_FOCAL = 36,     // 1 when the reduced focal diagonal was not positive
  S_FOCAL_SCHUR = 37,    // reduced focal diagonal H_ff(damped) - sum_p fc_p^T V_p^-1 fc_p
  S_DIRECT = 38,         // direct route this LM iteration: 0 none, 1 solved and not yet
                         // certified, 2 certified by the matrix-free residual
  S_DIRECT_SOLVES = 39,  // certified direct solves (whole solve)
  S_DUPLICATE = 40,      // set at set-up when a camera of the bundle observes one point twice
  S_DENSE_BAD = 41,      // a pivot of the multi-launch factorisation failed this LM iteration
  S_STEP_LIMIT = 42,     // word 23 at which the running PCG pass stops (its start + the budget)
  S_PASS_STOP = 43,      // |r|^2 at which the running PCG pass has converged: tol^2 |r0|^2 for the
                         // first pass; a refinement restart's own (k_refine)
  S_HELD = 44,           // exact-Hessian systems: points held fixed at the last PCG set-up
  S_WORDS = 48
};

constexpr int TILE = 64;        // observations per wave tile
constexpr int TILE_WAVES = 4;   // wave tiles per workgroup of the point-tile kernels
constexpr int TASK_THREADS = 256;  // threads of a camera-task workgroup
constexpr int TASK_OBS = 1024;     // most slots of a camera-list task (4 per thread)
constexpr int ADV_THREADS = 256;   // threads of a PCG-step workgroup
constexpr int GROUP_CAMS = 42;     // cameras of a PCG-step workgroup (6 rows each)
constexpr int GROUP_WORDS = 8;     // words of a PCG-step group record
constexpr int CG_THREADS = 1024;   // threads of the per-bundle PCG kernels
constexpr int COST_OBS = 256;      // observations per cost block
constexpr int ASM_WORDS = 56;      // per-task camera assembly sums (see k_camera_pass)
// exact-Hessian systems hold fixed a point whose block's smallest eigenvalue is below this fraction
// of its largest (singular_to_rounding3); a well-triangulated point of a long trajectory sits
// near 1e-5, one whose track has run off to 1e8 times the scene's depth near 1e-17
constexpr double kHeldPoint = 1e-12;
constexpr int DENSE_THREADS = 256; // threads of the direct route's row and factor workgroups
constexpr int DENSE_TABLE = 8192;  // point ids of one pass of a row workgroup's lookup table
constexpr int DENSE_NB = 32;       // block size of the dense Cholesky factorisation
constexpr uint32_t DENSE_MAX_CAMERAS = 128;  // largest bundle factored by one workgroup
constexpr uint32_t DENSE_LARGE_CAMERAS = 2048;  // largest bundle the direct route factors (the
                                                // larger ones by a sequence of launches)
constexpr int DENSE_LNB = 64;          // block size of the multi-launch factorisation
constexpr uint32_t DENSE_CHUNK_MIN = 256;  // fewest slots a table-indexed chunk of a camera's list
                                           // holds (sparser lists are searched by bisection)

// One bundle of the batch: its global ranges and gauge.
struct Bundle {
  uint32_t cam0, cam1, pt0, pt1, obs0, obs1;
  uint32_t tile0, tile1, task0, task1, cost0, cost1;
  uint32_t first_directional;  // global index of the bundle's first directional point
  uint32_t anchor, scale_view;  // global camera indices
  uint32_t normalize_scale;
  uint32_t focal_row;  // 6C + b
  uint32_t group0, group1;  // the bundle's PCG-step groups
  uint32_t direct;          // 1: the reduced camera system is assembled and factored
  uint32_t dense_n;         // rows of the bundle's dense array: m_b = 6 C_b [+ 1] unknowns + 1
  uint32_t unstored;        // 1: the passes store no cross blocks of this bundle (a direct bundle
                            // outside the exact systems); its sweeps and update relinearise
  uint32_t large;           // 1: a direct bundle of more than DENSE_MAX_CAMERAS cameras
  uint32_t covis_off;       // first word of the bundle's covisibility bitmap (direct bundles)
  uint32_t max_cg;          // PCG iterations of one pass (the bundle's maxCgIterations)
  uint32_t pad_;
  uint64_t dense_off;       // first double of the bundle's dense array
};

struct Problem {
  uint32_t C, P, M, B, m, K, shared;
  uint32_t n_tiles, n_tasks, n_cost_blocks, n_groups;
  double huber, tol;
  double refine_tol;  // mixed form: relative FP64 residual of the outer refinement
  double gauge_tol;   // relative residual a step's solution must meet to be proposed: the larger of
                      // tol and (mixed form) refine_tol, the tolerance it was certified at
  uint32_t cg_budget; // PCG iterations of one pass for every bundle (the exact-Hessian solves of the
                      // polish and the adjoint); 0: each bundle's own Bundle::max_cg
  uint32_t cg_max;    // the largest PCG budget of the batch's bundles (the host's launch budget)
  double stop_change; // an accepted step whose relative cost change is below this ends the LM loop
                      // (1e-9, acceptBundleCandidate b:1082-1095)
  double* scene;
  double* trial;
  const uint32_t* index;
  const Bundle* bundles;
  const uint32_t* tiles;          // n_tiles + 1 first points
  const uint32_t* tile_bundle;
  // n_tiles + 1: the first slot of each tile's code words (op_code, 64 slots per pass of the tile),
  // a multiple of 64
  const uint32_t* tile_slot;
  const uint32_t* tasks;          // 4 per task: camera, first slot, end slot, bundle
  const uint32_t* task_start;     // C + 1
  const uint32_t* cost_blocks;    // 3 per block: bundle, first and end observation
  const uint32_t* camera_bundle;  // C
  const uint32_t* point_bundle;   // P
  const uint32_t* groups;         // GROUP_WORDS per PCG-step group: bundle, first and end camera,
                                  // the bundle's first and end group, 3 unused
  uint32_t* cam_point;
  void* Wp;
  void* Wc;
  double* pblk;
  uint32_t* tile_fail;
  double* tile_focal;  // 2 per tile
  double* task_part;   // ASM_WORDS per task
  double* camA;        // 36 per camera
  double* camInv;      // 36 per camera
  double* camRhs;      // 6 per camera
  double* camHcf;      // 6 per camera
  double* camFocal;    // 2 per camera: H_ff and b_f shares
  double* group_part;  // 4 per PCG-step group: p . S p, focal share, |r|^2, r . z
  double* vec;         // x r p s z (m each), y (3P)
  double* state;       // S_WORDS per bundle
  double* cost_part;   // 3 per cost block
  uint32_t* error;     // index validation flag
  // a given right-hand side (the adjoint solve): b_c (6C), b_f (one per bundle), then b_p (3P);
  // null: the Gauss-Newton gradient -J^T w r
  const double* rhs_in;
  // exact-Hessian systems (polish, adjoint): regularise the one-fixed-camera similarity gauge
  uint32_t exact_gauge;
  // the Gauss-Newton adjoint: the exact passes keep only the Huber curvature of each observation
  uint32_t gauss_newton;
  // the stored cross blocks hold the exact Hessian's (the implicit operator relinearises the
  // Gauss-Newton blocks only, so it cannot certify such a system)
  uint32_t exact;
  uint32_t n_unstored;  // bundles whose cross blocks are not stored (Bundle::unstored)
  // the streamed route (the working budget holds no cross blocks): Wp and Wc are null and the
  // sweeps relinearise every block in registers, in the stored sweeps' expressions
  uint32_t streamed;
  // the direct route (null / 0 when no bundle takes it)
  uint32_t n_direct, n_direct_cams;
  const uint32_t* direct_bundles;  // n_direct bundle ids
  const uint32_t* direct_cams;     // n_direct_cams camera ids, the cameras of those bundles
  double* camS;    // 36 per camera: the Schur diagonal block S_cc = A_c - sum W V^-1 W^T
  double* camG;    // 18 per camera-major slot, contiguous: G = W L_p with V_p^-1 = L_p L_p^T
  double* pfc;     // 3 per point (shared focal): L_p^T fc_p, fc_p = sum_i w Jp_i^T Jf_i
  double* dense;   // per direct bundle dense_n x dense_n row-major; the lower triangle holds S,
                   // then its Cholesky factor; the last row holds the reduced right-hand side
  uint32_t* covis; // per direct bundle C_b rows of ceil(C_b / 32) words: bit d of row c set when
                   // cameras c > d observe a common point
  uint32_t n_small;                // direct bundles factored by one workgroup each
  const uint32_t* small_bundles;
  // the fused FP64 product (k_point_sweep_fused / k_camera_tsum; null: the two plain sweeps):
  // 8 per observation, t_i = W_i y_p(i) (6 or 7 words used, the rest zero)
  double* tprod;
  // launch schedule, read by the host only: dynamic LDS bytes reserved per camera-sweep workgroup,
  // which caps the workgroups a CU holds at once (0: no reservation)
  uint32_t camera_lds;
  // the one-pass Schur product of the stored-block forms (ba_onepass.hip; n_op = 0: the two sweeps):
  // n_op workgroups, workgroup w taking tiles op_tile[w] .. op_tile[w + 1] (at most OP_TILES, of
  // one bundle); its camera window starts at op_lo[w] and its partial rows at op_off[w] of op_part
  // (8 words per row, one row per camera of the window); camera c's rows are
  // op_row[op_cam[c] .. op_cam[c + 1]), in workgroup order
  uint32_t n_op;
  const uint32_t* op_tile;
  const uint32_t* op_lo;
  const uint32_t* op_off;
  const uint32_t* op_cam;
  const uint32_t* op_row;
  double* op_part;
  // per stored slot a code word (the row, rank and tile point of its observation, the segment of
  // the tile's point numbered as its lane) and per tile OP_RECORD words (ba_onepass.hip), written
  // at set-up
  const uint32_t* op_code;
  const uint32_t* op_trec;
  // the FP64 form (but for the exact systems): the passes store each observation's FP64 Jacobian
  // factors (kf_words doubles, the mixed form's words unrounded) in place of its cross block, and
  // the Schur products and the point update form every product from them (0: the cross blocks)
  uint32_t factored64;
  // after the one-pass product, p . S p and the focal share of each camera (2 per camera), written
  // by k_onepass_product with the camera's rows of S p and summed by k_cg_step over the bundle's
  // cameras (null: the group partials of k_cg_product)
  double* cam_dot;
  // the PCG loop's grids restricted to the bundles whose PCG may be active (null: every bundle's):
  // n_op_map one-pass workgroups, n_cam_map cameras of the one-pass reduction, n_group_map
  // PCG-step groups of bundles with several groups and n_single_map groups of bundles with one
  // (whose three PCG steps run as one launch, k_cg_single), each list in ascending order; the
  // host builds them between PCG chunks from the state words it reads
  const uint32_t* op_map;
  const uint32_t* cam_map;
  const uint32_t* group_map;
  const uint32_t* single_map;
  uint32_t n_op_map, n_cam_map, n_group_map, n_single_map;
};

// One-pass product: tiles per workgroup (OP_WAVE_TILES per wave) and the widest camera window.
constexpr int OP_WAVE_TILES = 8;
constexpr int OP_TILES = OP_WAVE_TILES * TILE_WAVES;
constexpr int OP_SPAN = 64;
constexpr int OP_RECORD = 8;  // words of a tile's record

// PCG iterations of one pass of bundle g.
__host__ __device__ inline double pcg_budget(const Problem& q, const Bundle& g) {
  return q.cg_budget ? q.cg_budget : g.max_cg;
}

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
