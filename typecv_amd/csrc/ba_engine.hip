// ba_engine.hip — the host side of the MI355X bundle-adjustment engine and its C ABI
// (include/typecv_ba.h). One context per HIP device owns the device arrays (grown, never shrunk,
// until tcv_ba_release), a stream, and pinned words for the state the LM loop reads.
//
// Host loop (reference bundleAdjustGpuFixed, src/sfm/bundle/gpu/bundle.ts:178-799, and
// bundleAdjustJointFocalGpu, joint-focal.ts:340-900): initialise and measure the cost; per LM
// iteration assemble, initialise PCG, run PCG in chunks with one read of the state words after each
// chunk, build and score the candidate, accept or reject, read the state; pack the result. Every
// decision is taken on the device from each bundle's state words; the host only reads them to
// stop launching work. A batch of bundles runs in lockstep: every launch covers all bundles and
// the kernels skip the bundles whose LM or PCG has ended.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/typecv_ba.h"
#include "ba_kernels.hip"
#include "ba_factored.hip"
#include "ba_onepass.hip"
#include "ba_dense.hip"
#include "obs_eval.hip"
#include "ba_order.hip"

using namespace tcvba;

namespace {

thread_local std::string g_last_error;

void set_error(const std::string& s) { g_last_error = s; }

struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
};

double now_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

}  // namespace

struct tcv_ba_context {
  // held by every C ABI entry for the whole call: calls on one context run one after another
  mutable std::mutex lock;
  int device = 0;
  hipStream_t stream = nullptr;
  double* h_state = nullptr;  // pinned copy of the state words of every bundle
  size_t h_state_words = 0;
  uint32_t* h_error = nullptr;
  std::vector<DevBuf*> all;
  DevBuf scene, trial, index, out, bundles, tiles, tile_bundle, tasks, task_start, cost_blocks,
      camera_bundle, point_bundle, cam_point, Wp, Wc, pblk, tile_fail, tile_focal, task_part, camA,
      tile_slot,
      camInv, camRhs, camHcf, camFocal, vec, state, cost_part, error, probe, groups, group_part,
      rhs, grad, camS, camG, pfc, dense, direct_bundles, direct_cams, covis, small_bundles,
      tprod,
      obs_cameras, obs_points, obs_pixels, obs_indices, obs_offsets, obs_flags, obs_errors,
      obs_partials, obs_summary, op_tile, op_lo, op_off, op_cam, op_row, op_part, op_code, op_trec,
      pcg_map, cam_dot, ord_scene, ord_index, ord_key, ord_perm, ord_sort;
  // per bundle (B + 1 prefix counts) its first one-pass workgroup, camera and PCG-step group, and
  // the host side of pcg_map (run_pcg's compacted grids)
  std::vector<uint32_t> bundle_ops, bundle_cams, bundle_groups, h_map;
  // the bundles whose lists pcg_map holds (empty: none uploaded since the last set-up)
  std::vector<char> map_on;
  // the direct bundles factored by a sequence of launches, and their dense array sizes
  std::vector<uint32_t> large_bundles, large_n;
  std::vector<uint64_t> large_off;
  std::vector<hipEvent_t> events;
  // the schedule tuner (run_pcg): the shape of the prepared problem, the shape the camera sweep's
  // residency was last tuned for and the workgroups per CU picked for it, and its events
  std::vector<uint64_t> shape, tuned_shape;
  uint32_t camera_cap = 0;
  bool tuned_fused = false;
  std::vector<hipEvent_t> tune_events;
  // the working budget of a solve in bytes (0: none; tcv_ba_set_working_bytes) and the read
  // failure injected by tcv_ba_fail_reads (the n-th next read of the state words fails; 0: none)
  uint64_t budget = 0;
  uint32_t fail_read_in = 0;
  tcv_ba_context() {
    all = {&scene,      &trial,        &index,     &out,       &bundles,    &tiles,
           &tile_bundle, &tasks,       &task_start, &cost_blocks, &camera_bundle,
           &point_bundle, &cam_point,  &Wp,        &Wc,        &pblk,       &tile_fail,
           &tile_focal, &task_part,    &camA,      &camInv,    &camRhs,     &camHcf,
           &camFocal,   &vec,          &state,     &cost_part, &error,      &probe,
           &groups,     &group_part,   &rhs,       &grad,      &camS,       &camG,   &pfc,
           &dense,      &direct_bundles, &direct_cams, &covis, &small_bundles, &tprod,
           &obs_cameras, &obs_points, &obs_pixels, &obs_indices, &obs_offsets, &obs_flags,
           &obs_errors, &obs_partials, &obs_summary, &tile_slot, &op_tile, &op_lo, &op_off, &op_cam, &op_row,
           &op_part,    &op_code,      &op_trec,   &pcg_map,   &cam_dot,   &ord_scene,
           &ord_index,  &ord_key,      &ord_perm,  &ord_sort};
  }
};

#define HIP_TRY(call)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess) {                                                              \
      set_error(std::string(#call) + ": " + hipGetErrorName(e_) + " (" +                \
                hipGetErrorString(e_) + ")");                                            \
      return e_ == hipErrorOutOfMemory ? TCV_BA_ERR_MEMORY : TCV_BA_ERR_DEVICE;          \
    }                                                                                    \
  } while (0)

#define TRY(expr)                   \
  do {                              \
    int s_ = (expr);                \
    if (s_ != TCV_BA_OK) return s_; \
  } while (0)

namespace {

int ensure(DevBuf& b, size_t bytes) {
  if (bytes == 0) bytes = 8;
  if (b.bytes >= bytes) return TCV_BA_OK;
  if (b.ptr) HIP_TRY(hipFree(b.ptr));
  b.ptr = nullptr;
  b.bytes = 0;
  HIP_TRY(hipMalloc(&b.ptr, bytes));
  b.bytes = bytes;
  return TCV_BA_OK;
}

int ensure_host_state(tcv_ba_context* ctx, size_t words) {
  if (ctx->h_state_words >= words) return TCV_BA_OK;
  if (ctx->h_state) HIP_TRY(hipHostFree(ctx->h_state));
  ctx->h_state = nullptr;
  ctx->h_state_words = 0;
  HIP_TRY(hipHostMalloc((void**)&ctx->h_state, words * 8, hipHostMallocDefault));
  ctx->h_state_words = words;
  return TCV_BA_OK;
}

// The checks of bundleProblem (src/sfm/bundle/gpu/bundle-host.ts:43-70) that the packed arrays
// still carry, and the u32 bounds of bundle.ts:118-127.
int check_control(const tcv_ba_control* c) {
  if (!c) return TCV_BA_ERR_ARGUMENT;
  const uint64_t C = c->cameras, P = c->points, M = c->observations;
  const bool ok = C >= 2 && P >= 1 && c->directional_points < P && M >= 1 &&
                  c->iterations <= 100 && c->max_cg >= 1 && c->max_cg <= 256 &&
                  c->huber >= 1e-12 && c->huber <= 1e6 && c->relative_tolerance >= 1e-12 &&
                  c->relative_tolerance <= 1e-2 && c->anchor < C && c->scale_view < C &&
                  c->anchor != c->scale_view && c->normalize_scale <= 1 && c->shared_focal <= 1 &&
                  (c->precision == TCV_BA_FP64 || c->precision == TCV_BA_MIXED) &&
                  3 * M + P + 2 * C + 2 <= 0xffffffffull && 18 * C + 3 * P + 2 * M <= 0xffffffffull &&
                  c->stall_tolerance >= 0 && c->stall_tolerance < 1 &&
                  (c->refinement_tolerance == 0 ||
                   (c->refinement_tolerance >= 1e-14 && c->refinement_tolerance <= 1e-2)) &&
                  (c->flags & ~(TCV_BA_FLAG_TIMING | TCV_BA_FLAG_IMPLICIT |
                                TCV_BA_FLAG_SCHUR_CACHE | TCV_BA_FLAG_DIRECT)) == 0;
  if (!ok) {
    set_error("control outside the admitted ranges (bundleProblem)");
    return TCV_BA_ERR_ARGUMENT;
  }
  return TCV_BA_OK;
}

// The options every bundle of a batch shares (each keeps its own sizes, gauge and PCG budget).
bool same_options(const tcv_ba_control& a, const tcv_ba_control& b) {
  return a.shared_focal == b.shared_focal && a.iterations == b.iterations &&
         a.precision == b.precision && a.flags == b.flags &&
         a.huber == b.huber && a.relative_tolerance == b.relative_tolerance &&
         a.stall_tolerance == b.stall_tolerance &&
         a.refinement_tolerance == b.refinement_tolerance;
}

struct Plan {
  std::vector<Bundle> bundles;
  std::vector<uint32_t> tiles, tile_bundle, tile_slot, tasks, task_start, cost_blocks,
      camera_bundle, point_bundle, groups, direct_bundles, direct_cams, small_bundles,
      large_bundles;
  uint64_t dense_words = 0;
  uint32_t covis_words = 0;
};

// The reference's guard of `schurCache` (joint-focal.ts:102-108): the shared-focal route, at most
// 16 cameras and M (6C + 1) within its work limit of 262,144.
bool schur_cache_admits(uint32_t shared, uint32_t cams, uint32_t obs) {
  return shared != 0 && cams <= 16 && (uint64_t)obs * (6 * cams + 1) <= 262144;
}

// The direct route for the bundles that ask for it (TCV_BA_FLAG_DIRECT: every bundle of at most
// DENSE_LARGE_CAMERAS cameras; TCV_BA_FLAG_SCHUR_CACHE: those inside the reference's guard), except
// those marked in `excluded`; all others stay on PCG. Outside the exact systems a direct bundle
// stores no cross blocks.
void plan_direct(Plan& plan, uint32_t flags, uint32_t shared, bool exact,
                 const std::vector<char>& excluded = {}) {
  plan.direct_bundles.clear();
  plan.direct_cams.clear();
  plan.small_bundles.clear();
  plan.large_bundles.clear();
  plan.dense_words = 0;
  plan.covis_words = 0;
  for (uint32_t b = 0; b < plan.bundles.size(); b++) {
    Bundle& g = plan.bundles[b];
    const uint32_t cams = g.cam1 - g.cam0;
    const bool out = b < excluded.size() && excluded[b];
    const bool want = (flags & TCV_BA_FLAG_DIRECT) != 0 ||
                      ((flags & TCV_BA_FLAG_SCHUR_CACHE) != 0 &&
                       schur_cache_admits(shared, cams, g.obs1 - g.obs0));
    g.direct = want && cams <= DENSE_LARGE_CAMERAS && !out ? 1u : 0u;
    g.large = g.direct && cams > DENSE_MAX_CAMERAS ? 1u : 0u;
    g.unstored = g.direct && !exact ? 1u : 0u;
    // m_b unknowns padded to a multiple of the factorisation's block, plus the right-hand-side
    // row; the multi-launch route keeps a vector after the array
    const uint32_t nb = g.large ? DENSE_LNB : DENSE_NB;
    g.dense_n = g.direct ? (6 * cams + shared + nb - 1) / nb * nb + 1 : 0;
    g.dense_off = plan.dense_words;
    g.covis_off = plan.covis_words;
    if (!g.direct) continue;
    plan.dense_words += (uint64_t)g.dense_n * g.dense_n + g.dense_n;
    if (g.large) plan.dense_words += (uint64_t)(g.dense_n - 1) * DENSE_LNB;  // the X_kk blocks
    plan.covis_words += cams * ((cams + 31) / 32);
    plan.direct_bundles.push_back(b);
    (g.large ? plan.large_bundles : plan.small_bundles).push_back(b);
    for (uint32_t c = g.cam0; c < g.cam1; c++) plan.direct_cams.push_back(c);
  }
}

constexpr uint32_t kTaskObs = TASK_OBS;
constexpr uint32_t kChunk = 8;  // PCG iterations between reads of the state words
constexpr uint32_t kTimingStride = 8;  // TCV_BA_FLAG_TIMING times every 8th Schur product
constexpr uint32_t kMaxRefinements = 4;  // mixed form: refinement restarts per LM iteration
constexpr uint32_t kDirectPasses = 3;  // certificates of a direct solution per LM iteration
                                       // (corrections in between)

// Bundle ranges, wave tiles of whole points of one bundle (at most 64 observations and 64 points;
// a longer point alone), camera-list tasks of at most kTaskObs slots (each camera has at least
// one), cost blocks of at most COST_OBS observations of one bundle.
int build_plan(const std::vector<uint32_t>& ps, const std::vector<uint32_t>& cs,
               const tcv_ba_control* controls, uint32_t B, uint32_t C, uint32_t P, uint32_t M,
               Plan& plan) {
  if (ps[0] != 0 || ps[P] != M || cs[0] != 0 || cs[C] != M) {
    set_error("index offsets do not cover the observations");
    return TCV_BA_ERR_INDEX;
  }
  for (uint32_t p = 0; p < P; p++)
    if (ps[p + 1] < ps[p]) {
      set_error("point offsets decrease");
      return TCV_BA_ERR_INDEX;
    }
  for (uint32_t c = 0; c < C; c++)
    if (cs[c + 1] < cs[c]) {
      set_error("camera offsets decrease");
      return TCV_BA_ERR_INDEX;
    }
  plan = Plan{};
  plan.tiles.push_back(0);
  plan.task_start.assign(C + 1, 0);
  plan.camera_bundle.resize(C);
  plan.point_bundle.resize(P);
  uint32_t cam = 0, pt = 0, obs = 0;
  for (uint32_t b = 0; b < B; b++) {
    const tcv_ba_control& k = controls[b];
    Bundle g{};
    g.cam0 = cam;
    g.cam1 = cam + k.cameras;
    g.pt0 = pt;
    g.pt1 = pt + k.points;
    g.obs0 = obs;
    g.obs1 = obs + k.observations;
    if (ps[g.pt0] != g.obs0 || ps[g.pt1] != g.obs1 || cs[g.cam0] != g.obs0 || cs[g.cam1] != g.obs1) {
      set_error("bundle offsets do not match the bundle sizes");
      return TCV_BA_ERR_INDEX;
    }
    g.first_directional = g.pt1 - k.directional_points;
    g.anchor = cam + k.anchor;
    g.scale_view = cam + k.scale_view;
    g.normalize_scale = k.normalize_scale;
    g.focal_row = 6 * C + b;
    g.max_cg = k.max_cg;
    // tiles of whole points of this bundle
    g.tile0 = (uint32_t)plan.tiles.size() - 1;
    uint32_t first = g.pt0;
    for (uint32_t p = g.pt0 + 1; p < g.pt1; p++) {
      const bool too_many_obs = ps[p + 1] - ps[first] > (uint32_t)TILE;
      const bool too_many_points = p - first >= (uint32_t)TILE;
      if (too_many_obs || too_many_points) {
        plan.tiles.push_back(p);
        first = p;
      }
    }
    plan.tiles.push_back(g.pt1);
    g.tile1 = (uint32_t)plan.tiles.size() - 1;
    // each tile's stored words from a 64-observation block of their own
    if (plan.tile_slot.empty()) plan.tile_slot.push_back(0);
    for (uint32_t t = g.tile0; t < g.tile1; t++) {
      const uint32_t n = ps[plan.tiles[t + 1]] - ps[plan.tiles[t]];
      plan.tile_slot.push_back(plan.tile_slot.back() + (n + 63) / 64 * 64);
    }
    while (plan.tile_bundle.size() < g.tile1) plan.tile_bundle.push_back(b);
    // camera tasks
    g.task0 = (uint32_t)(plan.tasks.size() / 4);
    for (uint32_t c = g.cam0; c < g.cam1; c++) {
      plan.task_start[c] = (uint32_t)(plan.tasks.size() / 4);
      plan.camera_bundle[c] = b;
      uint32_t at = cs[c];
      const uint32_t e = cs[c + 1];
      do {
        const uint32_t end = e - at > kTaskObs ? at + kTaskObs : e;
        plan.tasks.push_back(c);
        plan.tasks.push_back(at);
        plan.tasks.push_back(end);
        plan.tasks.push_back(b);
        at = end;
      } while (at < e);
    }
    g.task1 = (uint32_t)(plan.tasks.size() / 4);
    // PCG-step groups of at most GROUP_CAMS cameras
    g.group0 = (uint32_t)(plan.groups.size() / GROUP_WORDS);
    g.group1 = g.group0 + (g.cam1 - g.cam0 + GROUP_CAMS - 1) / GROUP_CAMS;
    for (uint32_t c = g.cam0; c < g.cam1; c += GROUP_CAMS) {
      const uint32_t record[GROUP_WORDS] = {b, c, std::min<uint32_t>(c + GROUP_CAMS, g.cam1),
                                            g.group0, g.group1, 0, 0, 0};
      plan.groups.insert(plan.groups.end(), record, record + GROUP_WORDS);
    }
    for (uint32_t p = g.pt0; p < g.pt1; p++) plan.point_bundle[p] = b;
    // cost blocks
    g.cost0 = (uint32_t)(plan.cost_blocks.size() / 3);
    for (uint32_t at = g.obs0; at < g.obs1; at += COST_OBS) {
      plan.cost_blocks.push_back(b);
      plan.cost_blocks.push_back(at);
      plan.cost_blocks.push_back(std::min<uint32_t>(at + COST_OBS, g.obs1));
    }
    g.cost1 = (uint32_t)(plan.cost_blocks.size() / 3);
    plan.bundles.push_back(g);
    cam = g.cam1;
    pt = g.pt1;
    obs = g.obs1;
  }
  plan.task_start[C] = (uint32_t)(plan.tasks.size() / 4);
  if (cam != C || pt != P || obs != M) {
    set_error("bundle sizes do not add up to the batch");
    return TCV_BA_ERR_INDEX;
  }
  return TCV_BA_OK;
}

uint32_t blocks_for(size_t n, uint32_t threads) { return (uint32_t)((n + threads - 1) / threads); }
uint32_t tile_blocks(const Problem& q) { return blocks_for(q.n_tiles, TILE_WAVES); }
// The one-pass product's and its reduction's grids (the compacted ones when run_pcg set maps).
uint32_t op_grid(const Problem& q) { return q.op_map ? q.n_op_map : q.n_op; }
uint32_t reduce_grid(const Problem& q) {
  return blocks_for((size_t)(q.cam_map ? q.n_cam_map : q.C) * 64, 256);
}

// The launches of one form (W storage type and route).
template <class WT, bool SHARED>
struct Launch {
  static void point_pass(const Problem& q, hipStream_t s) {
    if (sizeof(WT) == 8 && q.factored64)
      hipLaunchKernelGGL((k_point_pass<WT, SHARED, false, true>), dim3(tile_blocks(q)),
                         dim3(TILE * TILE_WAVES), 0, s, q);
    else
      hipLaunchKernelGGL((k_point_pass<WT, SHARED, false>), dim3(tile_blocks(q)),
                         dim3(TILE * TILE_WAVES), 0, s, q);
  }
  static void camera_pass(const Problem& q, hipStream_t s) {
    hipLaunchKernelGGL((k_camera_pass<WT, SHARED, false>), dim3(q.n_tasks), dim3(TASK_THREADS), 0,
                       s, q);
    hipLaunchKernelGGL((k_camera_finish<SHARED>), dim3(blocks_for(q.C, 64)), dim3(64), 0, s, q);
  }
  // The exact-Hessian assembly (FP64 blocks) of the Newton polish and the adjoint.
  static void exact_passes(const Problem& q, hipStream_t s) {
    hipLaunchKernelGGL((k_point_pass<double, SHARED, true>), dim3(tile_blocks(q)),
                       dim3(TILE * TILE_WAVES), 0, s, q);
    hipLaunchKernelGGL((k_camera_pass<double, SHARED, true>), dim3(q.n_tasks), dim3(TASK_THREADS),
                       0, s, q);
    hipLaunchKernelGGL((k_camera_finish<SHARED>), dim3(blocks_for(q.C, 64)), dim3(64), 0, s, q);
  }
  static void solve_points(const Problem& q, hipStream_t s) {
    hipLaunchKernelGGL((k_point_sweep<WT, SHARED, SWEEP_SOLVE>), dim3(tile_blocks(q)),
                       dim3(TILE * TILE_WAVES), 0, s, q);
  }
  static void pixel_cotangent(const Problem& q, double* out, double* focal, hipStream_t s) {
    hipLaunchKernelGGL((k_pixel_cotangent<SHARED>), dim3(blocks_for(q.M, 256)), dim3(256), 0, s,
                       q, out, focal);
  }
  static void jacobian(const Problem& q, double* out, hipStream_t s) {
    hipLaunchKernelGGL((k_jacobian<SHARED>), dim3(blocks_for(q.M, 256)), dim3(256), 0, s, q, out);
  }
  static void init_cg(const Problem& q, hipStream_t s) {
    hipLaunchKernelGGL((k_init_cg<SHARED>), dim3(q.B), dim3(CG_THREADS), 0, s, q);
  }
  static void point_sweep(const Problem& q, hipStream_t s) {
    if (q.n_op) {
      if (op_grid(q))
        hipLaunchKernelGGL((k_schur_onepass<WT, SHARED>), dim3(op_grid(q)),
                           dim3(TILE * TILE_WAVES), 0, s, q);
    }
    else if constexpr (sizeof(WT) == 4)
      hipLaunchKernelGGL((k_point_sweep_factored<SHARED, SWEEP_SCHUR>), dim3(tile_blocks(q)),
                         dim3(TILE * TILE_WAVES), 0, s, q);
    else if (q.factored64)
      hipLaunchKernelGGL((k_point_sweep_factored<SHARED, SWEEP_SCHUR, false, double>),
                         dim3(tile_blocks(q)), dim3(TILE * TILE_WAVES), 0, s, q);
    else if (q.tprod)
      hipLaunchKernelGGL((k_point_sweep_fused<WT, SHARED>), dim3(tile_blocks(q)),
                         dim3(TILE * TILE_WAVES), 0, s, q);
    else
      hipLaunchKernelGGL((k_point_sweep<WT, SHARED, SWEEP_SCHUR>), dim3(tile_blocks(q)),
                         dim3(TILE * TILE_WAVES), 0, s, q);
  }
  static void camera_sweep(const Problem& q, hipStream_t s) {
    if (q.n_op) {
      if (reduce_grid(q))
        hipLaunchKernelGGL((k_onepass_reduce<SHARED>), dim3(reduce_grid(q)), dim3(256), 0, s, q);
    }
    else if constexpr (sizeof(WT) == 4)
      hipLaunchKernelGGL((k_camera_sweep_factored<SHARED>), dim3(q.n_tasks), dim3(TASK_THREADS),
                         q.camera_lds, s, q);
    else if (q.factored64)
      hipLaunchKernelGGL((k_camera_sweep_factored<SHARED, false, double>), dim3(q.n_tasks),
                         dim3(TASK_THREADS), q.camera_lds, s, q);
    else if (q.tprod)
      hipLaunchKernelGGL((k_camera_tsum<SHARED>), dim3(q.n_tasks), dim3(TASK_THREADS),
                         q.camera_lds, s, q);
    else
      hipLaunchKernelGGL((k_camera_sweep<WT, SHARED>), dim3(q.n_tasks), dim3(TASK_THREADS),
                         q.camera_lds, s, q);
  }
  // After the one-pass product: each camera's rows summed and S p finished (k_onepass_product),
  // then the PCG step without k_cg_product.
  static void onepass_product(const Problem& q, hipStream_t s) {
    if (reduce_grid(q))
      hipLaunchKernelGGL((k_onepass_product<SHARED>), dim3(reduce_grid(q)), dim3(256), 0, s, q);
  }
  static void advance_rows(const Problem& q, hipStream_t s) {
    if (q.n_single_map)
      hipLaunchKernelGGL((k_cg_single<SHARED, false>), dim3(q.n_single_map), dim3(ADV_THREADS), 0,
                         s, q);
    const uint32_t groups = q.group_map ? q.n_group_map : q.n_groups;
    if (!groups) return;
    hipLaunchKernelGGL((k_cg_step<SHARED>), dim3(groups), dim3(ADV_THREADS), 0, s, q);
    hipLaunchKernelGGL((k_cg_direction<SHARED>), dim3(groups), dim3(ADV_THREADS), 0, s, q);
  }
  static void advance(const Problem& q, hipStream_t s) {
    if (q.n_single_map)
      hipLaunchKernelGGL((k_cg_single<SHARED>), dim3(q.n_single_map), dim3(ADV_THREADS), 0, s, q);
    const uint32_t groups = q.group_map ? q.n_group_map : q.n_groups;
    if (!groups) return;
    hipLaunchKernelGGL((k_cg_product<SHARED>), dim3(groups), dim3(ADV_THREADS), 0, s, q);
    hipLaunchKernelGGL((k_cg_step<SHARED>), dim3(groups), dim3(ADV_THREADS), 0, s, q);
    hipLaunchKernelGGL((k_cg_direction<SHARED>), dim3(groups), dim3(ADV_THREADS), 0, s, q);
  }
  static void update_points(const Problem& q, hipStream_t s) {
    if constexpr (sizeof(WT) == 4)
      hipLaunchKernelGGL((k_point_sweep_factored<SHARED, SWEEP_UPDATE>), dim3(tile_blocks(q)),
                         dim3(TILE * TILE_WAVES), 0, s, q);
    else if (q.factored64)
      hipLaunchKernelGGL((k_point_sweep_factored<SHARED, SWEEP_UPDATE, false, double>),
                         dim3(tile_blocks(q)), dim3(TILE * TILE_WAVES), 0, s, q);
    else
      hipLaunchKernelGGL((k_point_sweep<WT, SHARED, SWEEP_UPDATE>), dim3(tile_blocks(q)),
                         dim3(TILE * TILE_WAVES), 0, s, q);
  }
  // The streamed route: the same sweeps with every cross block relinearised in registers.
  static void streamed_point_sweep(const Problem& q, hipStream_t s) {
    if (q.n_op) {
      if (op_grid(q))
        hipLaunchKernelGGL((k_schur_onepass<WT, SHARED, true>), dim3(op_grid(q)),
                           dim3(TILE * TILE_WAVES), 0, s, q);
    }
    else if constexpr (sizeof(WT) == 4)
      hipLaunchKernelGGL((k_point_sweep_factored<SHARED, SWEEP_SCHUR, true>), dim3(tile_blocks(q)),
                         dim3(TILE * TILE_WAVES), 0, s, q);
    else if (q.factored64)
      hipLaunchKernelGGL((k_point_sweep_factored<SHARED, SWEEP_SCHUR, true, double>),
                         dim3(tile_blocks(q)), dim3(TILE * TILE_WAVES), 0, s, q);
    else
      hipLaunchKernelGGL((k_point_sweep<WT, SHARED, SWEEP_SCHUR, true>), dim3(tile_blocks(q)),
                         dim3(TILE * TILE_WAVES), 0, s, q);
  }
  static void streamed_camera_sweep(const Problem& q, hipStream_t s) {
    if (q.n_op) {
      if (reduce_grid(q))
        hipLaunchKernelGGL((k_onepass_reduce<SHARED>), dim3(reduce_grid(q)), dim3(256), 0, s, q);
    }
    else if constexpr (sizeof(WT) == 4)
      hipLaunchKernelGGL((k_camera_sweep_factored<SHARED, true>), dim3(q.n_tasks),
                         dim3(TASK_THREADS), 0, s, q);
    else if (q.factored64)
      hipLaunchKernelGGL((k_camera_sweep_factored<SHARED, true, double>), dim3(q.n_tasks),
                         dim3(TASK_THREADS), 0, s, q);
    else
      hipLaunchKernelGGL((k_camera_sweep<WT, SHARED, false, true>), dim3(q.n_tasks),
                         dim3(TASK_THREADS), 0, s, q);
  }
  static void streamed_update(const Problem& q, hipStream_t s) {
    if constexpr (sizeof(WT) == 4)
      hipLaunchKernelGGL((k_point_sweep_factored<SHARED, SWEEP_UPDATE, true>),
                         dim3(tile_blocks(q)), dim3(TILE * TILE_WAVES), 0, s, q);
    else if (q.factored64)
      hipLaunchKernelGGL((k_point_sweep_factored<SHARED, SWEEP_UPDATE, true, double>),
                         dim3(tile_blocks(q)), dim3(TILE * TILE_WAVES), 0, s, q);
    else
      hipLaunchKernelGGL((k_point_sweep<WT, SHARED, SWEEP_UPDATE, true>), dim3(tile_blocks(q)),
                         dim3(TILE * TILE_WAVES), 0, s, q);
  }
  static void probe(const Problem& q, const double* v, double* sv, double* zv, double* rhs,
                    hipStream_t s) {
    hipLaunchKernelGGL((k_probe<SHARED>), dim3(1), dim3(CG_THREADS), 0, s, q, v, sv, zv, rhs);
  }
  // The implicit FP64 operator on p (PCG, guarded per bundle) and the refinement step on x.
  static void implicit_p_sweep(const Problem& q, hipStream_t s) {
    hipLaunchKernelGGL((k_point_sweep_implicit<SHARED>), dim3(tile_blocks(q)),
                       dim3(TILE * TILE_WAVES), 0, s, q, (const double*)(q.vec + 2 * (size_t)q.m),
                       (int)GUARD_PCG);
  }
  static void unstored_p_sweep(const Problem& q, hipStream_t s) {
    hipLaunchKernelGGL((k_point_sweep_implicit<SHARED>), dim3(tile_blocks(q)),
                       dim3(TILE * TILE_WAVES), 0, s, q, (const double*)(q.vec + 2 * (size_t)q.m),
                       (int)GUARD_PCG_UNSTORED);
  }
  static void unstored_camera_sweep(const Problem& q, hipStream_t s) {
    hipLaunchKernelGGL((k_camera_sweep_implicit<SHARED>), dim3(q.n_tasks), dim3(TASK_THREADS), 0,
                       s, q, (int)GUARD_PCG_UNSTORED);
  }
  static void implicit_update(const Problem& q, hipStream_t s) {
    hipLaunchKernelGGL((k_point_sweep_implicit<SHARED>), dim3(tile_blocks(q)),
                       dim3(TILE * TILE_WAVES), 0, s, q, (const double*)q.vec, (int)GUARD_UPDATE);
  }
  static void implicit_camera_sweep(const Problem& q, hipStream_t s) {
    hipLaunchKernelGGL((k_camera_sweep_implicit<SHARED>), dim3(q.n_tasks), dim3(TASK_THREADS), 0,
                       s, q, (int)GUARD_PCG);
  }
  // The direct route: assembly and factorisation, the certificate of x, a correction from the factor.
  static void dense_solve(const Problem& q, hipStream_t s) {
    hipLaunchKernelGGL((k_dense_rows<SHARED>), dim3(q.n_direct_cams), dim3(DENSE_THREADS), 0,
                       s, q);
    if (q.n_small)
      hipLaunchKernelGGL((k_dense_factor<SHARED>), dim3(q.n_small), dim3(DENSE_THREADS), 0, s, q);
  }
  // A large bundle (dense array n x n at word `off`): the factorisation, panel by panel.
  static void large_factor(const Problem& q, uint32_t b, uint32_t n, uint64_t off, hipStream_t s) {
    const uint32_t u = n - 1;
    const LargeDense d{q.dense + off, q.state + (size_t)b * S_WORDS, n};
    hipLaunchKernelGGL((k_lfact_begin<SHARED>), dim3(64), dim3(DENSE_THREADS), 0, s, q, b);
    // the first diagonal block; each later one is factored by the update launch before it
    hipLaunchKernelGGL(k_lfact_diag, dim3(1), dim3(DENSE_THREADS), 0, s, d, 0u);
    // panels in pairs: the first panel updates only the second's column, and the trailing matrix
    // takes both panels' products in one pass over its tiles
    for (uint32_t kb = 0;;) {
      const uint32_t t = (u - kb - DENSE_LNB) / DENSE_LNB;  // blocks below panel kb
      // the panel's launch also carries the forward substitution of block kb (the last one too)
      hipLaunchKernelGGL(k_lfact_panel, dim3(t ? t : 1), dim3(DENSE_THREADS), 0, s, d, kb);
      if (!t) break;
      if (t == 1) {
        hipLaunchKernelGGL(k_lfact_update, dim3(1), dim3(DENSE_THREADS), 0, s, d, kb, 1u, 0u);
        kb += DENSE_LNB;
        continue;
      }
      hipLaunchKernelGGL(k_lfact_update, dim3(t), dim3(DENSE_THREADS), 0, s, d, kb, 1u, 1u);
      hipLaunchKernelGGL(k_lfact_panel, dim3(t - 1), dim3(DENSE_THREADS), 0, s, d,
                         kb + DENSE_LNB);
      hipLaunchKernelGGL(k_lfact_update, dim3((t - 1) * t / 2), dim3(DENSE_THREADS), 0, s, d, kb,
                         2u, 0u);
      kb += 2 * DENSE_LNB;
    }
  }
  // The substitutions of a large bundle: mode 0 finishes the solve of the assembled right-hand
  // side (its forward substitution ran with the factorisation) into x, mode 1 adds the correction
  // S^-1 r.
  static void large_solve(const Problem& q, uint32_t b, uint32_t n, uint64_t off, int mode,
                          hipStream_t s) {
    const uint32_t u = n - 1;
    const LargeDense d{q.dense + off, q.state + (size_t)b * S_WORDS, n};
    if (mode == 1) {
      hipLaunchKernelGGL((k_lsolve_in<SHARED>), dim3(blocks_for(u, DENSE_THREADS)),
                         dim3(DENSE_THREADS), 0, s, q, b);
      for (uint32_t kb = 0; kb < u; kb += DENSE_LNB) {
        const uint32_t rows = u - kb - DENSE_LNB;
        hipLaunchKernelGGL(k_lsolve_fwd, dim3(rows ? blocks_for(rows, DENSE_THREADS) : 1),
                           dim3(DENSE_THREADS), 0, s, d, kb);
      }
    }
    for (uint32_t kb = u; kb > 0;) {
      kb -= DENSE_LNB;
      hipLaunchKernelGGL(k_lsolve_bwd, dim3(kb ? blocks_for(kb, DENSE_THREADS) : 1),
                         dim3(DENSE_THREADS), 0, s, d, kb, mode);
    }
    // mode 0 clears word 19, which every workgroup of the launch tests first: one workgroup
    hipLaunchKernelGGL((k_lsolve_out<SHARED>), dim3(mode == 0 ? 1 : blocks_for(u, DENSE_THREADS)),
                       dim3(DENSE_THREADS), 0, s, q, b, mode);
  }
  // The certificate's product S x: the implicit FP64 operator, or the stored FP64 cross blocks of an
  // exact-Hessian system.
  static void certify(const Problem& q, int final_pass, double tol, hipStream_t s) {
    if (q.exact) {
      hipLaunchKernelGGL((k_point_sweep<WT, SHARED, SWEEP_CERTIFY>), dim3(tile_blocks(q)),
                         dim3(TILE * TILE_WAVES), 0, s, q);
      hipLaunchKernelGGL((k_camera_sweep<WT, SHARED, true>), dim3(q.n_tasks), dim3(TASK_THREADS),
                         0, s, q);
    } else {
      hipLaunchKernelGGL((k_point_sweep_implicit<SHARED>), dim3(tile_blocks(q)),
                         dim3(TILE * TILE_WAVES), 0, s, q, (const double*)q.vec,
                         (int)GUARD_DIRECT);
      hipLaunchKernelGGL((k_camera_sweep_implicit<SHARED>), dim3(q.n_tasks), dim3(TASK_THREADS),
                         0, s, q, (int)GUARD_DIRECT);
    }
    hipLaunchKernelGGL((k_certify<SHARED>), dim3(q.n_direct), dim3(CG_THREADS), 0, s, q,
                       final_pass, tol);
  }
  static void dense_correct(const Problem& q, hipStream_t s) {
    if (q.n_small)
      hipLaunchKernelGGL((k_dense_correct<SHARED>), dim3(q.n_small), dim3(DENSE_THREADS), 0, s, q);
  }
  static void refine(const Problem& q, hipStream_t s) {
    hipLaunchKernelGGL((k_point_sweep_implicit<SHARED>), dim3(tile_blocks(q)),
                       dim3(TILE * TILE_WAVES), 0, s, q, (const double*)q.vec, (int)GUARD_NONE);
    hipLaunchKernelGGL((k_camera_sweep_implicit<SHARED>), dim3(q.n_tasks), dim3(TASK_THREADS), 0,
                       s, q, (int)GUARD_NONE);
    hipLaunchKernelGGL((k_refine<SHARED>), dim3(q.B), dim3(CG_THREADS), 0, s, q);
  }
};

struct Ops {
  void (*point_pass)(const Problem&, hipStream_t);
  void (*camera_pass)(const Problem&, hipStream_t);
  void (*init_cg)(const Problem&, hipStream_t);
  void (*point_sweep)(const Problem&, hipStream_t);
  void (*camera_sweep)(const Problem&, hipStream_t);
  void (*advance)(const Problem&, hipStream_t);
  void (*update_points)(const Problem&, hipStream_t);
  void (*probe)(const Problem&, const double*, double*, double*, double*, hipStream_t);
  void (*refine)(const Problem&, hipStream_t);
  void (*exact_passes)(const Problem&, hipStream_t);
  void (*solve_points)(const Problem&, hipStream_t);
  void (*pixel_cotangent)(const Problem&, double*, double*, hipStream_t);
  void (*jacobian)(const Problem&, double*, hipStream_t);
  void (*dense_solve)(const Problem&, hipStream_t);
  void (*certify)(const Problem&, int, double, hipStream_t);
  void (*dense_correct)(const Problem&, hipStream_t);
  // the bundles whose cross blocks are not stored: their PCG sweeps (null when point_sweep and
  // camera_sweep are the implicit operator, which covers every bundle) and their update
  void (*unstored_point_sweep)(const Problem&, hipStream_t);
  void (*unstored_camera_sweep)(const Problem&, hipStream_t);
  void (*unstored_update)(const Problem&, hipStream_t);
  void (*large_factor)(const Problem&, uint32_t, uint32_t, uint64_t, hipStream_t);
  void (*large_solve)(const Problem&, uint32_t, uint32_t, uint64_t, int, hipStream_t);
  // the PCG loop after the one-pass product (null for the implicit operator, whose sweeps never
  // run it): S p finished per camera with the rows' reduction, then the step and direction
  void (*onepass_product)(const Problem&, hipStream_t);
  void (*advance_rows)(const Problem&, hipStream_t);
};

// The PCG sweeps of one iteration over the stored and the unstored bundles.
void point_sweeps(const Problem& q, const Ops& ops, hipStream_t st) {
  if (!ops.unstored_point_sweep || q.n_unstored < q.B) ops.point_sweep(q, st);
  if (ops.unstored_point_sweep && q.n_unstored) ops.unstored_point_sweep(q, st);
}
void camera_sweeps(const Problem& q, const Ops& ops, hipStream_t st) {
  if (!ops.unstored_camera_sweep || q.n_unstored < q.B) ops.camera_sweep(q, st);
  if (ops.unstored_camera_sweep && q.n_unstored) ops.unstored_camera_sweep(q, st);
}
void update_points(const Problem& q, const Ops& ops, hipStream_t st) {
  if (q.n_unstored < q.B) ops.update_points(q, st);
  if (q.n_unstored) ops.unstored_update(q, st);
}

template <class WT, bool SHARED>
Ops ops_for(bool implicit, bool streamed) {
  using L = Launch<WT, SHARED>;
  Ops o{L::point_pass,   L::camera_pass,   L::init_cg,         L::point_sweep,
        L::camera_sweep, L::advance,       L::update_points,   L::probe,
        L::refine,       L::exact_passes,  L::solve_points,    L::pixel_cotangent,
        L::jacobian,     L::dense_solve,   L::certify,         L::dense_correct,
        L::unstored_p_sweep, L::unstored_camera_sweep, L::implicit_update,
        L::large_factor, L::large_solve, L::onepass_product, L::advance_rows};
  if (implicit) {
    o.point_sweep = L::implicit_p_sweep;
    o.camera_sweep = L::implicit_camera_sweep;
    o.unstored_point_sweep = nullptr;
    o.unstored_camera_sweep = nullptr;
    o.onepass_product = nullptr;
    o.advance_rows = nullptr;
  } else if (streamed) {
    o.point_sweep = L::streamed_point_sweep;
    o.camera_sweep = L::streamed_camera_sweep;
    o.update_points = L::streamed_update;
  }
  return o;
}

Ops select_ops(const tcv_ba_control* c, bool streamed = false) {
  const bool shared = c->shared_focal != 0;
  const bool implicit = (c->flags & TCV_BA_FLAG_IMPLICIT) != 0;
  if (c->precision == TCV_BA_MIXED)
    return shared ? ops_for<float, true>(implicit, streamed)
                  : ops_for<float, false>(implicit, streamed);
  return shared ? ops_for<double, true>(implicit, streamed)
                : ops_for<double, false>(implicit, streamed);
}

template <class T>
int upload(DevBuf& buf, const std::vector<T>& v, hipStream_t st) {
  TRY(ensure(buf, v.size() * sizeof(T)));
  if (!v.empty())
    HIP_TRY(hipMemcpyAsync(buf.ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
  return TCV_BA_OK;
}

int read_state(tcv_ba_context* ctx, const Problem& q, hipStream_t st, tcv_ba_report* rep) {
  if (ctx->fail_read_in && --ctx->fail_read_in == 0) {
    set_error("injected read failure of the state words");
    return TCV_BA_ERR_DEVICE;
  }
  HIP_TRY(hipMemcpyAsync(ctx->h_state, q.state, (size_t)q.B * S_WORDS * 8, hipMemcpyDeviceToHost,
                         st));
  HIP_TRY(hipStreamSynchronize(st));
  if (rep) rep->flag_reads++;
  return TCV_BA_OK;
}

double word(const tcv_ba_context* ctx, uint32_t b, int w) {
  return ctx->h_state[(size_t)b * S_WORDS + w];
}

bool any_word(const tcv_ba_context* ctx, const Problem& q, int w) {
  for (uint32_t b = 0; b < q.B; b++)
    if (word(ctx, b, w) != 0) return true;
  return false;
}

double max_word(const tcv_ba_context* ctx, const Problem& q, int w) {
  double v = 0;
  for (uint32_t b = 0; b < q.B; b++) v = std::max(v, word(ctx, b, w));
  return v;
}

// A device array of a solve and the bytes the solve needs in it.
struct Need {
  DevBuf* buf;
  size_t bytes;
};

// Every device array a solve of `plan` allocates (besides the caller's staging): the resident
// form stores each observation's Jacobian factors (FP64 or, mixed, FP32), or for the exact systems
// its FP64 cross block (and then the fused product's rows); the streamed form relinearises them in
// every sweep and stores none.
std::vector<Need> solve_needs(tcv_ba_context* ctx, const tcv_ba_control* c, uint32_t B,
                              uint32_t C, uint32_t P, uint32_t M, const Plan& plan, bool exact,
                              bool streamed) {
  const uint32_t shared = c->shared_focal ? 1u : 0u;
  const size_t ws = c->precision == TCV_BA_MIXED ? 4 : 8;
  const uint32_t K = shared ? 21 : 18;
  const size_t m = (size_t)C * 6 + shared * B;
  const uint32_t n_unstored = exact ? 0u : (uint32_t)plan.direct_bundles.size();
  const size_t n_tiles = plan.tiles.size() - 1, n_tasks = plan.tasks.size() / 4;
  const size_t n_cost = plan.cost_blocks.size() / 3, n_groups = plan.groups.size() / GROUP_WORDS;
  // the factors (KF words per observation) or, for the exact systems, the cross blocks (K)
  const uint32_t words = exact ? K : kf_words_of(ws, shared);
  const size_t w_words = (size_t)words * plan.tile_slot.back();
  const size_t wp_words = (size_t)words * M;  // tile_word: no padding
  std::vector<Need> n = {
      {&ctx->scene, ((size_t)C * 18 + (size_t)P * 3 + (size_t)M * 2) * 8},
      {&ctx->trial, ((size_t)C * 12 + (size_t)P * 3) * 8},
      {&ctx->cam_point, (size_t)M * 4},
      {&ctx->pblk, (size_t)P * 12 * 8},
      {&ctx->tile_fail, n_tiles * 4},
      {&ctx->tile_focal, n_tiles * 16},
      {&ctx->task_part, n_tasks * ASM_WORDS * 8},
      {&ctx->camA, (size_t)C * 36 * 8},
      {&ctx->camInv, (size_t)C * 36 * 8},
      {&ctx->camRhs, (size_t)C * 6 * 8},
      {&ctx->camHcf, (size_t)C * 6 * 8},
      {&ctx->camFocal, (size_t)C * 2 * 8},
      {&ctx->vec, (m * 5 + (size_t)P * 3) * 8},
      {&ctx->state, (size_t)B * S_WORDS * 8},
      {&ctx->cost_part, n_cost * 3 * 8},
      {&ctx->error, 8},
      {&ctx->group_part, n_groups * 4 * 8},
      {&ctx->bundles, plan.bundles.size() * sizeof(Bundle)},
      {&ctx->tiles, plan.tiles.size() * 4},
      {&ctx->tile_bundle, plan.tile_bundle.size() * 4},
      {&ctx->tile_slot, plan.tile_slot.size() * 4},
      {&ctx->tasks, plan.tasks.size() * 4},
      {&ctx->task_start, plan.task_start.size() * 4},
      {&ctx->cost_blocks, plan.cost_blocks.size() * 4},
      {&ctx->camera_bundle, plan.camera_bundle.size() * 4},
      {&ctx->point_bundle, plan.point_bundle.size() * 4},
      {&ctx->groups, plan.groups.size() * 4},
  };
  if (n_unstored < B && !streamed) {
    n.push_back({&ctx->Wp, wp_words * ws});
    n.push_back({&ctx->Wc, w_words * ws});
    // the fused FP64 product's rows over the cross blocks of the exact systems (the tuner times it
    // against the plain sweeps, which compute the same bits); over factors it is not offered (over
    // FP32 cross blocks it was slower: 70.8 + 27.4 against 42.6 + 40.0 us at 500 cameras, since
    // the rows stay FP64 while the blocks it saves are halved)
    if (ws == 8 && exact) n.push_back({&ctx->tprod, (size_t)M * 64});
  }
  if (!exact) {
    // the one-pass product's workgroups (at most OP_TILES tiles of one bundle each), windows, row
    // tables and partial rows (at most OP_SPAN per workgroup)
    const size_t n_op = (n_tiles + OP_TILES - 1) / OP_TILES + B;
    n.push_back({&ctx->op_tile, (n_op + 1) * 4});
    n.push_back({&ctx->op_lo, n_op * 4});
    n.push_back({&ctx->op_off, (n_op + 1) * 4});
    n.push_back({&ctx->op_cam, ((size_t)C + 1) * 4});
    n.push_back({&ctx->op_row, n_op * OP_SPAN * 4});
    n.push_back({&ctx->op_part, n_op * OP_SPAN * 8 * 8});
    // each stored slot's code word and each tile's record
    n.push_back({&ctx->op_code, (size_t)plan.tile_slot.back() * 4});
    n.push_back({&ctx->op_trec, n_tiles * OP_RECORD * 4});
    // the PCG loop's grids: one-pass workgroups, cameras and PCG-step groups
    n.push_back({&ctx->pcg_map, (n_op + C + n_groups) * 4});
    // each camera's p . S p and focal share after the one-pass product
    n.push_back({&ctx->cam_dot, (size_t)C * 2 * 8});
  }
  if (!plan.direct_bundles.empty()) {
    n.push_back({&ctx->camS, (size_t)C * 36 * 8});
    n.push_back({&ctx->camG, (size_t)M * 18 * 8});
    if (shared) n.push_back({&ctx->pfc, (size_t)P * 3 * 8});
    n.push_back({&ctx->dense, plan.dense_words * 8});
    n.push_back({&ctx->covis, (size_t)plan.covis_words * 4});
    n.push_back({&ctx->direct_bundles, plan.direct_bundles.size() * 4});
    n.push_back({&ctx->direct_cams, plan.direct_cams.size() * 4});
    n.push_back({&ctx->small_bundles, plan.small_bundles.size() * 4});
  }
  for (Need& x : n) x.bytes = x.bytes ? x.bytes : 8;
  return n;
}

size_t need_bytes(const std::vector<Need>& needs) {
  size_t total = 0;
  for (const Need& x : needs) total += x.bytes;
  return total;
}

// The bytes the context holds once every need is allocated (buffers grow and are never shrunk).
size_t held_after(const tcv_ba_context* ctx, const std::vector<Need>& needs) {
  size_t total = 0;
  for (const DevBuf* b : ctx->all) {
    size_t bytes = b->bytes;
    for (const Need& x : needs)
      if (x.buf == b) bytes = std::max(bytes, x.bytes);
    total += bytes;
  }
  return total;
}

int release_buffers(tcv_ba_context* ctx);

// The residency plan under the context's working budget (reference planBundleResidency,
// src/device/gpu/budget-plan.ts:114-188, and its refusal, src/sfm/bundle/gpu/bundle.ts:214-242):
// resident when every array of the solve, the caller's `staging` included, fits the budget;
// otherwise streamed when the arrays without the cross blocks fit; otherwise refused before any
// allocation. With a budget, a context whose held buffers would pass it releases them first.
// No budget (0): resident. `streamed` receives the decision.
int admit(tcv_ba_context* ctx, const tcv_ba_control* c, uint32_t B, uint32_t C, uint32_t P,
          uint32_t M, const Plan& plan, bool exact, const std::vector<Need>& staging,
          bool& streamed) {
  streamed = false;
  if (ctx->budget == 0) return TCV_BA_OK;
  std::vector<Need> resident = solve_needs(ctx, c, B, C, P, M, plan, exact, false);
  std::vector<Need> floor = solve_needs(ctx, c, B, C, P, M, plan, exact, true);
  resident.insert(resident.end(), staging.begin(), staging.end());
  floor.insert(floor.end(), staging.begin(), staging.end());
  const size_t need = need_bytes(floor);
  if (need > ctx->budget) {
    set_error("GPU BA needs " + std::to_string(need) +
              " bytes for the arrays that stay resident and a window of one track, more than the "
              "working budget of " + std::to_string(ctx->budget) + " bytes");
    return TCV_BA_ERR_MEMORY;
  }
  streamed = need_bytes(resident) > ctx->budget;
  if (held_after(ctx, streamed ? floor : resident) > ctx->budget) TRY(release_buffers(ctx));
  return TCV_BA_OK;
}

// The point and camera offsets of a packed index (host array) and the plan of its batch.
int host_plan(const tcv_ba_control* controls, uint32_t B, const uint32_t* index, uint32_t C,
              uint32_t P, uint32_t M, bool exact, Plan& plan) {
  std::vector<uint32_t> ps(index + 2 * (size_t)M, index + 2 * (size_t)M + P + 1);
  std::vector<uint32_t> cs(index + 2 * (size_t)M + P + 1, index + 2 * (size_t)M + P + C + 2);
  TRY(build_plan(ps, cs, controls, B, C, P, M, plan));
  plan_direct(plan, controls[0].flags, controls[0].shared_focal ? 1u : 0u, exact);
  return TCV_BA_OK;
}

// Residency of a setup call: decided by setup against the budget, or given by a caller that
// admitted its staging with the plan already.
enum Residency { DECIDE = -1, RESIDENT = 0, STREAMED = 1 };

int plan_onepass(tcv_ba_context* ctx, Problem& q, const Plan& plan, hipStream_t st,
                 tcv_ba_report* rep);

// Allocate, validate the index and build the plan of a batch of B bundles whose packed arrays
// `scene` / `index` (device pointers) hold the whole batch.
int setup(tcv_ba_context* ctx, const tcv_ba_control* controls, uint32_t B, const double* scene,
          const uint32_t* index, hipStream_t st, Problem& q, tcv_ba_report* rep,
          bool exact = false, int residency = DECIDE) {
  const double t0 = now_ms();
  if (B < 1) return TCV_BA_ERR_ARGUMENT;
  uint64_t C64 = 0, P64 = 0, M64 = 0;
  for (uint32_t b = 0; b < B; b++) {
    TRY(check_control(&controls[b]));
    if (!same_options(controls[0], controls[b])) {
      set_error("the bundles of a batch must share every option but their sizes and gauge");
      return TCV_BA_ERR_ARGUMENT;
    }
    C64 += controls[b].cameras;
    P64 += controls[b].points;
    M64 += controls[b].observations;
  }
  if (3 * M64 + P64 + 2 * C64 + 2 > 0xffffffffull || 18 * C64 + 3 * P64 + 2 * M64 > 0xffffffffull) {
    set_error("batch exceeds the u32 index bounds");
    return TCV_BA_ERR_ARGUMENT;
  }
  const tcv_ba_control* c = &controls[0];
  const uint32_t C = (uint32_t)C64, P = (uint32_t)P64, M = (uint32_t)M64;
  const uint32_t shared = c->shared_focal ? 1u : 0u;
  const size_t scene_size = (size_t)C * 18 + (size_t)P * 3 + (size_t)M * 2;
  const size_t trial_size = (size_t)C * 12 + (size_t)P * 3;
  const size_t ws = c->precision == TCV_BA_MIXED ? 4 : 8;
  const uint32_t K = shared ? 21 : 18;
  const uint32_t m = C * 6 + shared * B;
  std::vector<uint32_t> ps(P + 1), cs(C + 1);
  HIP_TRY(hipMemcpyAsync(ps.data(), index + 2 * (size_t)M, (P + 1) * 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(cs.data(), index + 2 * (size_t)M + P + 1, (C + 1) * 4,
                         hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  Plan plan;
  TRY(build_plan(ps, cs, controls, B, C, P, M, plan));
  plan_direct(plan, c->flags, shared, exact);
  const bool direct = !plan.direct_bundles.empty();
  const uint32_t n_unstored = exact ? 0u : (uint32_t)plan.direct_bundles.size();
  const uint32_t n_tiles = (uint32_t)plan.tiles.size() - 1;
  const uint32_t n_tasks = (uint32_t)plan.tasks.size() / 4;
  const uint32_t n_cost = (uint32_t)plan.cost_blocks.size() / 3;
  const uint32_t n_groups = (uint32_t)plan.groups.size() / GROUP_WORDS;
  // the stored words per observation: the factors, or the exact systems' cross blocks
  const uint32_t stored_words = exact ? K : kf_words_of(ws, shared);
  const size_t w_words = (size_t)stored_words * plan.tile_slot.back();
  const size_t vec_words = (size_t)m * 5 + (size_t)P * 3;
  bool streamed = residency == STREAMED;
  if (residency == DECIDE) TRY(admit(ctx, c, B, C, P, M, plan, exact, {}, streamed));
  for (const Need& x : solve_needs(ctx, c, B, C, P, M, plan, exact, streamed))
    TRY(ensure(*x.buf, x.bytes));
  const bool fused = !streamed && ws == 8 && exact && n_unstored < B;
  TRY(ensure_host_state(ctx, (size_t)B * S_WORDS));
  if (direct) {
    TRY(upload(ctx->direct_bundles, plan.direct_bundles, st));
    TRY(upload(ctx->direct_cams, plan.direct_cams, st));
    TRY(upload(ctx->small_bundles, plan.small_bundles, st));
  }
  TRY(upload(ctx->bundles, plan.bundles, st));
  TRY(upload(ctx->tiles, plan.tiles, st));
  TRY(upload(ctx->tile_bundle, plan.tile_bundle, st));
  TRY(upload(ctx->tile_slot, plan.tile_slot, st));
  TRY(upload(ctx->tasks, plan.tasks, st));
  TRY(upload(ctx->task_start, plan.task_start, st));
  TRY(upload(ctx->cost_blocks, plan.cost_blocks, st));
  TRY(upload(ctx->camera_bundle, plan.camera_bundle, st));
  TRY(upload(ctx->point_bundle, plan.point_bundle, st));
  TRY(upload(ctx->groups, plan.groups, st));
  HIP_TRY(hipMemcpyAsync(ctx->scene.ptr, scene, scene_size * 8, hipMemcpyDeviceToDevice, st));
  HIP_TRY(hipMemsetAsync(ctx->trial.ptr, 0, trial_size * 8, st));
  HIP_TRY(hipMemsetAsync(ctx->vec.ptr, 0, vec_words * 8, st));
  HIP_TRY(hipMemsetAsync(ctx->state.ptr, 0, (size_t)B * S_WORDS * 8, st));
  HIP_TRY(hipMemsetAsync(ctx->error.ptr, 0, 8, st));
  q = Problem{};
  q.C = C;
  q.P = P;
  q.M = M;
  q.B = B;
  q.m = m;
  q.K = K;
  q.shared = shared;
  q.n_tiles = n_tiles;
  q.n_tasks = n_tasks;
  q.n_cost_blocks = n_cost;
  q.n_groups = n_groups;
  q.huber = c->huber;
  q.tol = c->relative_tolerance;
  q.refine_tol = c->refinement_tolerance > 0 ? c->refinement_tolerance : c->relative_tolerance;
  q.gauge_tol = c->precision == TCV_BA_MIXED ? std::max(q.tol, q.refine_tol) : q.tol;
  q.cg_budget = 0;
  q.cg_max = 0;
  for (uint32_t b = 0; b < B; b++) q.cg_max = std::max(q.cg_max, controls[b].max_cg);
  q.stop_change = 1e-9;
  q.scene = (double*)ctx->scene.ptr;
  q.trial = (double*)ctx->trial.ptr;
  q.index = index;
  q.bundles = (const Bundle*)ctx->bundles.ptr;
  q.tiles = (const uint32_t*)ctx->tiles.ptr;
  q.tile_bundle = (const uint32_t*)ctx->tile_bundle.ptr;
  q.tile_slot = (const uint32_t*)ctx->tile_slot.ptr;
  q.tasks = (const uint32_t*)ctx->tasks.ptr;
  q.task_start = (const uint32_t*)ctx->task_start.ptr;
  q.cost_blocks = (const uint32_t*)ctx->cost_blocks.ptr;
  q.camera_bundle = (const uint32_t*)ctx->camera_bundle.ptr;
  q.point_bundle = (const uint32_t*)ctx->point_bundle.ptr;
  q.groups = (const uint32_t*)ctx->groups.ptr;
  q.group_part = (double*)ctx->group_part.ptr;
  q.cam_point = (uint32_t*)ctx->cam_point.ptr;
  q.Wp = n_unstored < B && !streamed ? ctx->Wp.ptr : nullptr;
  q.Wc = n_unstored < B && !streamed ? ctx->Wc.ptr : nullptr;
  q.streamed = streamed ? 1u : 0u;
  q.exact = exact ? 1u : 0u;
  q.n_unstored = n_unstored;
  q.pblk = (double*)ctx->pblk.ptr;
  q.tile_fail = (uint32_t*)ctx->tile_fail.ptr;
  q.tile_focal = (double*)ctx->tile_focal.ptr;
  q.task_part = (double*)ctx->task_part.ptr;
  q.camA = (double*)ctx->camA.ptr;
  q.camInv = (double*)ctx->camInv.ptr;
  q.camRhs = (double*)ctx->camRhs.ptr;
  q.camHcf = (double*)ctx->camHcf.ptr;
  q.camFocal = (double*)ctx->camFocal.ptr;
  q.vec = (double*)ctx->vec.ptr;
  q.state = (double*)ctx->state.ptr;
  q.cost_part = (double*)ctx->cost_part.ptr;
  q.error = (uint32_t*)ctx->error.ptr;
  q.tprod = fused ? (double*)ctx->tprod.ptr : nullptr;
  if (direct) {
    q.n_direct = (uint32_t)plan.direct_bundles.size();
    q.n_direct_cams = (uint32_t)plan.direct_cams.size();
    q.direct_bundles = (const uint32_t*)ctx->direct_bundles.ptr;
    q.direct_cams = (const uint32_t*)ctx->direct_cams.ptr;
    q.camS = (double*)ctx->camS.ptr;
    q.camG = (double*)ctx->camG.ptr;
    q.pfc = shared ? (double*)ctx->pfc.ptr : nullptr;
    q.dense = (double*)ctx->dense.ptr;
    q.covis = (uint32_t*)ctx->covis.ptr;
    q.n_small = (uint32_t)plan.small_bundles.size();
    q.small_bundles = (const uint32_t*)ctx->small_bundles.ptr;
  }
  hipLaunchKernelGGL(k_check_points, dim3(blocks_for(P, 256)), dim3(256), 0, st, q);
  hipLaunchKernelGGL(k_check_tasks, dim3(n_tasks), dim3(256), 0, st, q);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(ctx->h_error, ctx->error.ptr, 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if ((*ctx->h_error & 3u) != 0) {
    set_error(*ctx->h_error & 1 ? "observation point/camera ids do not match the point offsets"
                                : "camera lists do not match the observation cameras");
    return TCV_BA_ERR_INDEX;
  }
  if (direct && (*ctx->h_error & 4u) != 0) {
    // a camera observes one point twice: that bundle's camera lists are not sets of points, so it
    // keeps PCG (word 40 marks it); the others keep the direct route
    TRY(read_state(ctx, q, st, nullptr));
    std::vector<char> excluded(B);
    for (uint32_t b = 0; b < B; b++) excluded[b] = word(ctx, b, S_DUPLICATE) != 0;
    plan_direct(plan, c->flags, shared, exact, excluded);
    TRY(upload(ctx->bundles, plan.bundles, st));
    TRY(upload(ctx->direct_bundles, plan.direct_bundles, st));
    TRY(upload(ctx->direct_cams, plan.direct_cams, st));
    TRY(upload(ctx->small_bundles, plan.small_bundles, st));
    q.n_direct = (uint32_t)plan.direct_bundles.size();
    q.n_direct_cams = (uint32_t)plan.direct_cams.size();
    q.n_small = (uint32_t)plan.small_bundles.size();
    q.n_unstored = exact ? 0u : q.n_direct;
    // the excluded bundles need cross blocks: stored, or under a working budget relinearised in
    // every sweep (the same bits)
    if (q.n_unstored < B && ctx->budget) q.streamed = 1;
    if (q.n_unstored < B && !q.streamed) {
      TRY(ensure(ctx->Wp, (size_t)stored_words * M * ws));
      TRY(ensure(ctx->Wc, w_words * ws));
      q.Wp = ctx->Wp.ptr;
      q.Wc = ctx->Wc.ptr;
    }
  }
  ctx->bundle_cams.assign(B + 1, 0);
  ctx->bundle_groups.assign(B + 1, 0);
  for (uint32_t b = 0; b < B; b++) {
    ctx->bundle_cams[b + 1] = plan.bundles[b].cam1;
    ctx->bundle_groups[b + 1] = plan.bundles[b].group1;
  }
  ctx->bundle_ops.assign(B + 1, 0);
  ctx->map_on.clear();
  if (!exact) TRY(plan_onepass(ctx, q, plan, st, rep));
  // the FP64 form stores FP64 Jacobian factors (ba_factored.hip): 88 bytes per observation
  // instead of the cross block's 168 (the one-pass product 187 against 291 us at 2,000 cameras);
  // the exact systems keep the cross blocks
  q.factored64 = !exact && c->precision == TCV_BA_FP64 ? 1u : 0u;
  ctx->large_bundles = plan.large_bundles;
  ctx->large_n.clear();
  ctx->large_off.clear();
  for (uint32_t b : plan.large_bundles) {
    ctx->large_n.push_back(plan.bundles[b].dense_n);
    ctx->large_off.push_back(plan.bundles[b].dense_off);
  }
  if (q.n_direct) {
    // which cameras of a direct bundle share points: the assembly skips the pairs that do not
    HIP_TRY(hipMemsetAsync(q.covis, 0, (size_t)plan.covis_words * 4, st));
    hipLaunchKernelGGL(k_covisibility, dim3(blocks_for(P, 256)), dim3(256), 0, st, q);
    HIP_TRY(hipGetLastError());
  }
  // the shape that the camera sweep's schedule is tuned for
  ctx->shape = {C, P, M, B, n_tasks, ws, shared, q.n_unstored, q.streamed};
  if (rep) {
    rep->setup_ms = now_ms() - t0;
    rep->point_tiles = n_tiles;
    rep->camera_tasks = n_tasks;
  }
  return TCV_BA_OK;
}

// The one-pass Schur product (ba_onepass.hip): workgroups of at most OP_TILES tiles
// of one bundle, counted from the bundle's first tile (so a batch groups each bundle's tiles as its
// solve alone does), the camera window of each, and, when every window spans at most OP_SPAN
// cameras, each workgroup's first partial row, each camera's rows in workgroup order, each stored
// slot's code word and each tile's record. Otherwise q.n_op stays 0 and the two sweeps run.
int plan_onepass(tcv_ba_context* ctx, Problem& q, const Plan& plan, hipStream_t st,
                 tcv_ba_report* rep) {
  q.n_op = 0;
  std::vector<uint32_t> h_tile, h_bundle;
  for (uint32_t b = 0; b < plan.bundles.size(); b++) {
    const Bundle& g = plan.bundles[b];
    for (uint32_t t = g.tile0; t < g.tile1; t += OP_TILES) {
      h_tile.push_back(t);
      h_bundle.push_back(b);
    }
  }
  const uint32_t n_op = (uint32_t)h_tile.size();
  ctx->bundle_ops.assign(plan.bundles.size() + 1, 0);
  for (uint32_t w = 0; w < n_op; w++) ctx->bundle_ops[h_bundle[w] + 1]++;
  for (size_t b = 0; b < plan.bundles.size(); b++) ctx->bundle_ops[b + 1] += ctx->bundle_ops[b];
  if (n_op == 0) return TCV_BA_OK;
  h_tile.push_back(q.n_tiles);
  TRY(upload(ctx->op_tile, h_tile, st));
  q.op_tile = (const uint32_t*)ctx->op_tile.ptr;
  uint32_t* lo = (uint32_t*)ctx->op_lo.ptr;
  uint32_t* off = (uint32_t*)ctx->op_off.ptr;
  hipLaunchKernelGGL(k_onepass_windows, dim3(n_op), dim3(256), 0, st, q, lo, off);
  HIP_TRY(hipGetLastError());
  std::vector<uint32_t> h_lo(n_op), h_span(n_op);
  HIP_TRY(hipMemcpyAsync(h_lo.data(), lo, n_op * 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(h_span.data(), off, n_op * 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  // the camera of row k of workgroup w's window (Window in ba_onepass.hip)
  auto camera = [&](uint32_t w, uint32_t k) {
    const Bundle& g = plan.bundles[h_bundle[w]];
    const uint32_t v = (h_lo[w] & 0x7fffffffu) + k, n = g.cam1 - g.cam0;
    return (h_lo[w] >> 31) && v >= g.cam0 + n ? v - n : v;
  };
  std::vector<uint32_t> h_off(n_op + 1, 0), h_cam(q.C + 1, 0);
  for (uint32_t w = 0; w < n_op; w++) {
    if (h_span[w] > (uint32_t)OP_SPAN) return TCV_BA_OK;  // a wide window: the two sweeps
    h_off[w + 1] = h_off[w] + h_span[w];
    for (uint32_t k = 0; k < h_span[w]; k++) h_cam[camera(w, k) + 1]++;
  }
  for (uint32_t c = 0; c < q.C; c++) h_cam[c + 1] += h_cam[c];
  std::vector<uint32_t> h_row(h_cam[q.C]), fill(h_cam.begin(), h_cam.end() - 1);
  for (uint32_t w = 0; w < n_op; w++)
    for (uint32_t k = 0; k < h_span[w]; k++) h_row[fill[camera(w, k)]++] = h_off[w] + k;
  TRY(upload(ctx->op_off, h_off, st));
  TRY(upload(ctx->op_cam, h_cam, st));
  TRY(upload(ctx->op_row, h_row, st));
  q.op_lo = lo;
  // every window spans at most OP_SPAN rows: each slot's code and each tile's record
  hipLaunchKernelGGL(k_onepass_codes, dim3(n_op), dim3(TILE * TILE_WAVES), 0, st, q,
                     (uint32_t*)ctx->op_code.ptr, (uint32_t*)ctx->op_trec.ptr);
  HIP_TRY(hipGetLastError());
  q.n_op = n_op;
  q.op_off = (const uint32_t*)ctx->op_off.ptr;
  q.op_cam = (const uint32_t*)ctx->op_cam.ptr;
  q.op_row = (const uint32_t*)ctx->op_row.ptr;
  q.op_part = (double*)ctx->op_part.ptr;
  q.op_code = (const uint32_t*)ctx->op_code.ptr;
  q.op_trec = (const uint32_t*)ctx->op_trec.ptr;
  if (rep) rep->onepass_rows = h_off[n_op];
  // nothing reads the camera-major blocks or the fused product's rows then
  q.Wc = nullptr;
  q.tprod = nullptr;
  HIP_TRY(hipStreamSynchronize(st));  // the host tables are uploaded before they go out of scope
  return TCV_BA_OK;
}

int ensure_events(tcv_ba_context* ctx, size_t n) {
  while (ctx->events.size() < n) {
    hipEvent_t e;
    HIP_TRY(hipEventCreate(&e));
    ctx->events.push_back(e);
  }
  return TCV_BA_OK;
}

// The Schur product's schedule tuner. Its candidates differ in speed only, never in a bit: the
// two plain sweeps with the camera sweep's residency left to its registers or capped at three
// workgroups per CU (an LDS reservation per workgroup: 160 KiB per CU, at most 63 KiB, which
// needs no opt-in), and the fused product (k_point_sweep_fused + k_camera_tsum, whose 64-byte
// rows of t per observation pay off while the device's last-level cache holds them). Neither the
// task partition nor any sum changes. Measured at 500 cameras (FP64): plain 66.1 + 61.6 us (point
// + camera sweep), capped camera sweep 60.2 us, fused 95.0 + 27.6 us (-3.5 %); at 2,000 cameras
// the fused product is 13 % slower (its 621 MB of rows miss that cache). The first PCG chunk of
// a problem shape not yet tuned runs its iterations 1..6 with the candidates in turn (each
// iteration's one product, so nothing is added), times each product with HIP events and keeps
// the fastest candidate, if it beats the plain sweeps by 2 % (timing noise aside).
struct TuneCandidate {
  bool fused;
  uint32_t cap;
};
constexpr TuneCandidate kTune[3] = {{false, 0}, {false, 3}, {true, 0}};
constexpr uint32_t kTuneFirst = 1, kTuneRuns = 6;  // iterations [1, 7) of the first chunk

uint32_t camera_lds(uint32_t cap) {
  return cap ? std::min<uint32_t>(163840 / cap - 1024, 64512) : 0;
}

// The schedule `pick` applied to the problem q: its residency cap and product.
Problem scheduled(const Problem& q, const TuneCandidate& pick) {
  Problem s = q;
  s.camera_lds = camera_lds(pick.cap);
  if (!pick.fused) s.tprod = nullptr;
  return s;
}

int ensure_tune_events(tcv_ba_context* ctx) {
  while (ctx->tune_events.size() < 2 * kTuneRuns) {
    hipEvent_t e;
    HIP_TRY(hipEventCreate(&e));
    ctx->tune_events.push_back(e);
  }
  return TCV_BA_OK;
}

// After the tuned chunk has been read back: the candidates' summed times, the pick.
int tune_pick(tcv_ba_context* ctx, bool fused_available) {
  double total[3] = {0, 0, 0};
  for (uint32_t k = 0; k < kTuneRuns; k++) {
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->tune_events[2 * k], ctx->tune_events[2 * k + 1]));
    total[k % 3] += ms;
  }
  uint32_t best = 0;
  for (uint32_t j = 1; j < 3; j++)
    if ((fused_available || !kTune[j].fused) && total[j] < total[best]) best = j;
  if (total[best] > 0.98 * total[0]) best = 0;  // a schedule other than the plain one must win by 2 %
  ctx->tuned_fused = kTune[best].fused;
  ctx->camera_cap = kTune[best].cap;
  ctx->tuned_shape = ctx->shape;
  return TCV_BA_OK;
}

// The PCG loop's grids restricted to the bundles in `on`: their one-pass workgroups, the cameras
// of the one-pass reduction and their PCG-step groups, in order, the groups of bundles with one
// group apart (k_cg_single runs their three steps in one launch). The other bundles' workgroups
// would return at once, yet each costs its dispatch: a product launch of the 256-bundle batch
// while few of its bundles were still in PCG took 32 us at the median, 17.8 us at the 5th
// percentile. out = q with the maps set (without maps, when pcg_map is not held: every bundle's
// grid, three launches per step). Called only after a read of the state words (a stream
// synchronisation), so the previous upload from h_map has completed.
int compact_pcg(tcv_ba_context* ctx, const Problem& q, const std::vector<char>& on,
                hipStream_t st, Problem& out) {
  out = q;
  out.op_map = out.cam_map = out.group_map = out.single_map = nullptr;
  out.n_op_map = out.n_cam_map = out.n_group_map = out.n_single_map = 0;
  const size_t words = (size_t)ctx->bundle_ops.back() + q.C + q.n_groups;
  if (!ctx->pcg_map.ptr || ctx->pcg_map.bytes < words * 4 || ctx->bundle_ops.size() != q.B + 1 ||
      ctx->bundle_cams.size() != q.B + 1 || ctx->bundle_groups.size() != q.B + 1)
    return TCV_BA_OK;
  std::vector<uint32_t>& h = ctx->h_map;
  h.clear();
  const std::vector<uint32_t>& groups = ctx->bundle_groups;
  auto single = [&](uint32_t b) { return groups[b + 1] - groups[b] == 1; };
  // the lists: one-pass workgroups, cameras, groups of bundles with several, groups of those with one
  const std::vector<uint32_t>* ranges[4] = {&ctx->bundle_ops, &ctx->bundle_cams, &groups, &groups};
  uint32_t n[4];
  for (int k = 0; k < 4; k++) {
    const size_t first = h.size();
    for (uint32_t b = 0; b < q.B; b++)
      if (on[b] && (k < 2 || single(b) == (k == 3)))
        for (uint32_t i = (*ranges[k])[b]; i < (*ranges[k])[b + 1]; i++) h.push_back(i);
    n[k] = (uint32_t)(h.size() - first);
  }
  // a list that covers the whole grid in order stays a null map: no lookup in those launches; the
  // lists are uploaded when some map is not null and they differ from those pcg_map holds
  const bool identity = n[0] == ctx->bundle_ops.back() && n[1] == q.C && n[2] == q.n_groups;
  if (!identity && on != ctx->map_on) {
    HIP_TRY(hipMemcpyAsync(ctx->pcg_map.ptr, h.data(), h.size() * 4, hipMemcpyHostToDevice, st));
    ctx->map_on = on;
  }
  const uint32_t* base = (const uint32_t*)ctx->pcg_map.ptr;
  if (n[0] != ctx->bundle_ops.back()) out.op_map = base;
  if (n[1] != q.C) out.cam_map = base + n[0];
  if (n[2] != q.n_groups) out.group_map = base + n[0] + n[1];
  out.single_map = base + n[0] + n[1] + n[2];
  out.n_op_map = n[0];
  out.n_cam_map = n[1];
  out.n_group_map = n[2];
  out.n_single_map = n[3];
  return TCV_BA_OK;
}

// PCG in chunks with one read of the state words after each: the first chunk `first` iterations,
// later ones kChunk, at most `budget` in all. A chunk that outlives convergence costs only guarded
// no-op launches (state word 19), as the reference's chunks do. With TCV_BA_FLAG_TIMING the two
// sweeps of every kTimingStride-th launch in which some bundle's PCG was active are timed with HIP
// events (each event is a queue packet between kernels: timing every launch slows the solve by 7 %
// at 500 cameras).
int run_pcg(tcv_ba_context* ctx, const tcv_ba_control* c, const Problem& q, const Ops& ops,
            hipStream_t st, tcv_ba_report* rep, uint32_t first, uint32_t budget,
            bool known = false) {
  const bool timing = (c->flags & TCV_BA_FLAG_TIMING) != 0 && rep;
  uint32_t done = 0, chunk = first;
  double matvecs_before = max_word(ctx, q, S_MATVECS);
  // the product's schedule: tuned in the first chunk when this shape has not been (the stored-block
  // sweeps only; the tuning iterations are not multiples of kTimingStride); until then the plain
  // sweeps with no cap
  const bool tuned = ctx->tuned_shape == ctx->shape;
  bool tune = !tuned && !q.streamed && q.n_op == 0 && ops.unstored_camera_sweep != nullptr &&
              q.n_unstored < q.B && std::min(first, budget) >= kTuneFirst + kTuneRuns;
  static_assert(kTuneFirst > 0 && kTuneFirst + kTuneRuns <= kTimingStride, "tuning iterations");
  if (tune) TRY(ensure_tune_events(ctx));
  // after the one-pass product S p is finished per camera (k_onepass_product): one launch less
  const bool rows = q.n_op != 0 && ops.onepass_product != nullptr;
  Problem qr = q;
  qr.cam_dot = rows ? (double*)ctx->cam_dot.ptr : nullptr;
  Problem qt = scheduled(qr, tuned ? TuneCandidate{ctx->tuned_fused, ctx->camera_cap} : kTune[0]);
  // the launches cover the bundles whose PCG may run: those whose PCG the last read of the state
  // words showed active (`known`: it followed this pass's set-up), else (a batch) those whose LM
  // loop it showed active, or the one bundle; after each chunk those still in PCG
  std::vector<char> on(q.B);
  for (uint32_t b = 0; b < q.B; b++)
    on[b] = known ? word(ctx, b, S_PCG_ACTIVE) != 0 : q.B == 1 || word(ctx, b, S_ACTIVE) != 0;
  Problem qa;
  TRY(compact_pcg(ctx, qt, on, st, qa));
  bool tune_done = false;  // qt changed: qa follows it
  while (done < budget) {
    const uint32_t n = chunk < budget - done ? chunk : budget - done;
    for (uint32_t k = 0; k < n; k++) {
      const bool timed = timing && k % kTimingStride == 0;
      const bool trial = tune && k >= kTuneFirst && k < kTuneFirst + kTuneRuns;
      if (trial) {
        const uint32_t r = k - kTuneFirst;
        const Problem qc = scheduled(q, kTune[r % 3]);
        HIP_TRY(hipEventRecord(ctx->tune_events[2 * r], st));
        point_sweeps(qc, ops, st);
        camera_sweeps(qc, ops, st);
        HIP_TRY(hipEventRecord(ctx->tune_events[2 * r + 1], st));
      } else {
        if (timed) HIP_TRY(hipEventRecord(ctx->events[3 * k], st));
        point_sweeps(qa, ops, st);
        if (timed) HIP_TRY(hipEventRecord(ctx->events[3 * k + 1], st));
        if (!rows) {
          camera_sweeps(qa, ops, st);
        } else {
          if (ops.unstored_camera_sweep && q.n_unstored) ops.unstored_camera_sweep(qa, st);
          ops.onepass_product(qa, st);
        }
        if (timed) HIP_TRY(hipEventRecord(ctx->events[3 * k + 2], st));
      }
      (rows ? ops.advance_rows : ops.advance)(qa, st);
    }
    HIP_TRY(hipGetLastError());
    if (rep) rep->matvec_launches += n;
    done += n;
    TRY(read_state(ctx, q, st, rep));
    const double now = max_word(ctx, q, S_MATVECS);
    const uint32_t active = (uint32_t)(now - matvecs_before);
    matvecs_before = now;
    if (rep) rep->matvecs += active;
    if (timing) {
      for (uint32_t k = 0; k < active && k < n; k += kTimingStride) {
        float a = 0, b = 0;
        HIP_TRY(hipEventElapsedTime(&a, ctx->events[3 * k], ctx->events[3 * k + 1]));
        HIP_TRY(hipEventElapsedTime(&b, ctx->events[3 * k + 1], ctx->events[3 * k + 2]));
        rep->schur_point_ms += a;
        rep->schur_camera_ms += b;
        rep->schur_ms += a + b;
        rep->timed_matvecs++;
      }
    }
    if (tune) {
      TRY(tune_pick(ctx, q.tprod != nullptr));
      qt = scheduled(qr, TuneCandidate{ctx->tuned_fused, ctx->camera_cap});
      tune = false;
      tune_done = true;
    }
    if (!any_word(ctx, q, S_PCG_ACTIVE)) break;
    // (the lists are rebuilt, between the chunks on the host, only when they change)
    bool changed = tune_done;
    for (uint32_t b = 0; b < q.B; b++) {
      const char v = word(ctx, b, S_PCG_ACTIVE) != 0;
      changed |= v != on[b];
      on[b] = v;
    }
    if (changed) TRY(compact_pcg(ctx, qt, on, st, qa));
    tune_done = false;
    chunk = kChunk;
  }
  if (rep) {
    rep->camera_sweep_cap = qt.camera_lds ? ctx->camera_cap : 0;
    rep->fused_product = qt.tprod != nullptr;
  }
  return TCV_BA_OK;
}

// The linear solve of a step, after init_cg. The direct route first: factor and solve, then
// certify x with the matrix-free residual at `tol` (the form's final residual), correcting from the
// factor in between; a solution that still misses it hands over to PCG, which also runs the
// bundles that are not direct and those whose factorisation failed.
int solve_linear(tcv_ba_context* ctx, const tcv_ba_control* c, const Problem& q, const Ops& ops,
                 hipStream_t st, tcv_ba_report* rep, uint32_t first, uint32_t budget, double tol) {
  if (!q.n_direct) return run_pcg(ctx, c, q, ops, st, rep, first, budget);
  ops.dense_solve(q, st);
  for (size_t k = 0; k < ctx->large_bundles.size(); k++) {
    ops.large_factor(q, ctx->large_bundles[k], ctx->large_n[k], ctx->large_off[k], st);
    ops.large_solve(q, ctx->large_bundles[k], ctx->large_n[k], ctx->large_off[k], 0, st);
  }
  for (uint32_t k = 0; k < kDirectPasses; k++) {
    ops.certify(q, k + 1 == kDirectPasses, tol, st);
    HIP_TRY(hipGetLastError());
    TRY(read_state(ctx, q, st, rep));
    uint32_t pending = 0;
    for (uint32_t b = 0; b < q.B; b++) pending += word(ctx, b, S_DIRECT) == 1 ? 1u : 0u;
    if (!pending) break;
    if (rep && k + 1 < kDirectPasses) rep->direct_corrections += pending;
    ops.dense_correct(q, st);
    for (size_t j = 0; j < ctx->large_bundles.size(); j++)
      if (word(ctx, ctx->large_bundles[j], S_DIRECT) == 1)
        ops.large_solve(q, ctx->large_bundles[j], ctx->large_n[j], ctx->large_off[j], 1, st);
  }
  if (!any_word(ctx, q, S_PCG_ACTIVE)) return TCV_BA_OK;
  return run_pcg(ctx, c, q, ops, st, rep, first, budget, true);
}

// Ends the PCG of every bundle whose budget ran out (the reference stops at maxCgIterations);
// its candidate is then rejected by the gauge's convergence check, as in the reference.
__global__ void k_stop_pcg(Problem q) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < q.B) q.state[(size_t)b * S_WORDS + S_PCG_ACTIVE] = 0;
}

// The LM loop on a prepared batch. `finish_ms` (optional, B entries) receives each bundle's time
// from the start of the loop until its LM loop ended.
int run_lm(tcv_ba_context* ctx, const tcv_ba_control* c, const Problem& q, hipStream_t st,
           tcv_ba_report* rep, double* finish_ms) {
  const Ops ops = select_ops(c, q.streamed != 0);
  const bool timing = (c->flags & TCV_BA_FLAG_TIMING) != 0 && rep;
  const double t0 = now_ms();
  hipLaunchKernelGGL(k_init_bundle, dim3(q.B), dim3(64), 0, st, q);
  hipLaunchKernelGGL(k_cost<false>, dim3(q.n_cost_blocks), dim3(COST_OBS), 0, st, q);
  hipLaunchKernelGGL(k_finish_cost<false>, dim3(q.B), dim3(COST_OBS), 0, st, q);
  HIP_TRY(hipGetLastError());
  TRY(read_state(ctx, q, st, rep));
  const bool monitor = c->stall_tolerance > 1e-9;
  std::vector<double> monitor_cost(q.B), finish(q.B, -1.0);
  std::vector<char> active(q.B);
  for (uint32_t b = 0; b < q.B; b++) {
    monitor_cost[b] = word(ctx, b, S_COST);
    active[b] = word(ctx, b, S_ACTIVE) != 0;
    if (!active[b]) finish[b] = now_ms() - t0;
  }
  uint32_t hint = 16;
  std::vector<char> unstored;  // the mixed form's refinement skips these (their PCG is FP64)
  const uint32_t update_blocks = blocks_for(q.C > q.P ? q.C : q.P, 256);
  const uint32_t trial_blocks = blocks_for((size_t)q.C * 12 + (size_t)q.P * 3, 256);
  if (timing) TRY(ensure_events(ctx, 3 * 256 + 3));
  auto any_active = [&]() {
    for (uint32_t b = 0; b < q.B; b++)
      if (active[b]) return true;
    return false;
  };
  for (uint32_t iteration = 0; iteration < c->iterations && any_active(); iteration++) {
    ops.point_pass(q, st);
    ops.camera_pass(q, st);
    ops.init_cg(q, st);
    HIP_TRY(hipGetLastError());
    TRY(solve_linear(ctx, c, q, ops, st, rep, hint, q.cg_max,
                     c->precision == TCV_BA_MIXED ? q.refine_tol : q.tol));
    // The mixed form: replace the FP32-operator residual by the FP64 one and restart PCG from x
    // until the FP64 residual meets the refinement tolerance (each restart has its own budget).
    if (c->precision == TCV_BA_MIXED) {
      if (unstored.empty()) {
        unstored.resize(q.B);
        std::vector<Bundle> bundles(q.B);
        HIP_TRY(hipMemcpy(bundles.data(), q.bundles, q.B * sizeof(Bundle), hipMemcpyDeviceToHost));
        for (uint32_t b = 0; b < q.B; b++) unstored[b] = bundles[b].unstored != 0;
      }
      if (any_word(ctx, q, S_PCG_ACTIVE))
        hipLaunchKernelGGL(k_stop_pcg, dim3(blocks_for(q.B, 64)), dim3(64), 0, st, q);
      for (uint32_t k = 0; k < kMaxRefinements; k++) {
        bool pending = false;
        // every pass that converged or broke down (k_refine leaves the failures of the PCG
        // set-up alone)
        for (uint32_t b = 0; b < q.B; b++)
          pending |= (word(ctx, b, S_CONVERGED) != 0 || word(ctx, b, S_PCG_FAILED) != 0) &&
                     word(ctx, b, S_ACTIVE) != 0 && word(ctx, b, S_DIRECT) != 2 && !unstored[b];
        if (!pending) break;
        ops.refine(q, st);
        HIP_TRY(hipGetLastError());
        TRY(read_state(ctx, q, st, rep));
        if (!any_word(ctx, q, S_PCG_ACTIVE)) break;
        TRY(run_pcg(ctx, c, q, ops, st, rep, kChunk, q.cg_max, true));
        if (any_word(ctx, q, S_PCG_ACTIVE))
          hipLaunchKernelGGL(k_stop_pcg, dim3(blocks_for(q.B, 64)), dim3(64), 0, st, q);
      }
      if (rep) {
        double total = 0;
        for (uint32_t b = 0; b < q.B; b++) total += word(ctx, b, S_REFINEMENTS);
        rep->refinement_steps = total;
      }
    }
    const uint32_t steps = (uint32_t)max_word(ctx, q, S_STEPS);
    // the next LM iteration's first chunk: half of this one's PCG iterations (three quarters left
    // 25 no-op iterations per LM iteration past convergence at 500 cameras, half leaves 8; a
    // quarter saves no more)
    hint = steps / 2 < kChunk ? kChunk : steps / 2;
    hipLaunchKernelGGL(k_update_cameras, dim3(blocks_for(q.C, 64)), dim3(64), 0, st, q);
    update_points(q, ops, st);
    hipLaunchKernelGGL(k_gauge, dim3(q.B), dim3(256), 0, st, q);
    hipLaunchKernelGGL(k_normalize, dim3(update_blocks), dim3(256), 0, st, q);
    hipLaunchKernelGGL(k_cost<true>, dim3(q.n_cost_blocks), dim3(COST_OBS), 0, st, q);
    hipLaunchKernelGGL(k_finish_cost<true>, dim3(q.B), dim3(COST_OBS), 0, st, q);
    hipLaunchKernelGGL(k_commit, dim3(trial_blocks), dim3(256), 0, st, q);
    HIP_TRY(hipGetLastError());
    TRY(read_state(ctx, q, st, rep));
    for (uint32_t b = 0; b < q.B; b++) {
      if (!active[b]) continue;
      bool on = word(ctx, b, S_ACTIVE) != 0;
      if (monitor && on) {
        // the host stall rule of bundleStallMonitor (bundle-shared.ts:83-123)
        const double next = word(ctx, b, S_COST), change = monitor_cost[b] - next;
        const bool accepted = word(ctx, b, S_COMMIT) == 1;
        monitor_cost[b] = next;
        if (accepted && change < c->stall_tolerance * (next > 1 ? next : 1)) {
          const double zero = 0;
          HIP_TRY(hipMemcpyAsync(q.state + (size_t)b * S_WORDS + S_ACTIVE, &zero, 8,
                                 hipMemcpyHostToDevice, st));
          on = false;
        }
      }
      if (!on) {
        active[b] = 0;
        finish[b] = now_ms() - t0;
      }
    }
  }
  HIP_TRY(hipStreamSynchronize(st));
  const double end = now_ms() - t0;
  if (finish_ms)
    for (uint32_t b = 0; b < q.B; b++) finish_ms[b] = finish[b] < 0 ? end : finish[b];
  if (rep) {
    double cg = 0, lm = 0, direct = 0;
    for (uint32_t b = 0; b < q.B; b++) {
      cg += word(ctx, b, S_CG_STEPS);
      lm += word(ctx, b, S_ACCEPTED) + word(ctx, b, S_REJECTED);
      direct += word(ctx, b, S_DIRECT_SOLVES);
    }
    rep->cg_iterations = (uint64_t)cg;
    rep->lm_iterations = (uint64_t)lm;
    rep->direct_solves = (uint64_t)direct;
  }
  return TCV_BA_OK;
}

// Free every device buffer of the context (the reference's `drop`; also what a failed call does,
// as the reference drops a failed job's buffers, device.ts:592-598).
int release_buffers(tcv_ba_context* ctx) {
  HIP_TRY(hipSetDevice(ctx->device));
  if (ctx->stream) HIP_TRY(hipStreamSynchronize(ctx->stream));
  for (DevBuf* b : ctx->all) {
    if (b->ptr) HIP_TRY(hipFree(b->ptr));
    b->ptr = nullptr;
    b->bytes = 0;
  }
  return TCV_BA_OK;
}

uint64_t held_bytes(const tcv_ba_context* ctx) {
  uint64_t total = 0;
  for (const DevBuf* b : ctx->all) total += b->bytes;
  return total;
}

// Solve a batch whose packed arrays are on the device; `out` receives the batch's coordinates
// then (26 + shared) state words per bundle.
int solve_device(tcv_ba_context* ctx, const tcv_ba_control* controls, uint32_t B,
                 const double* scene, const uint32_t* index, double* out, tcv_ba_report* rep,
                 double* finish_ms, hipStream_t st, int residency = DECIDE) {
  const double t0 = now_ms();
  HIP_TRY(hipSetDevice(ctx->device));
  Problem q;
  TRY(setup(ctx, controls, B, scene, index, st, q, rep, false, residency));
  if (rep) rep->streamed = q.streamed;
  TRY(run_lm(ctx, &controls[0], q, st, rep, finish_ms));
  const size_t total = (size_t)q.C * 12 + (size_t)q.P * 3 + (size_t)B * (26 + q.shared);
  hipLaunchKernelGGL(k_pack, dim3(blocks_for(total, 256)), dim3(256), 0, st, q, out);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(st));
  if (rep) {
    rep->solve_ms = now_ms() - t0;
    rep->device_bytes = held_bytes(ctx);
  }
  return TCV_BA_OK;
}

// ------------------------------------------------------------------ Newton polish and adjoint

// Upload a host scene and index and set up a one-bundle FP64 problem at log focal scale
// `log_focal`, with the PCG tolerance `tolerance`; the scene's cost words are filled. Rejects the
// inputs the forward rejects.
int prepare_exact(tcv_ba_context* ctx, const tcv_ba_control* c, const double* scene,
                  const uint32_t* index, double log_focal, double tolerance, hipStream_t st,
                  tcv_ba_control& fc, Problem& q, Ops& ops, tcv_ba_report* rep) {
  TRY(check_control(c));
  if (!(tolerance >= 1e-14 && tolerance <= 1e-2) || !(std::fabs(log_focal) < 700)) {
    set_error("tolerance outside [1e-14, 1e-2] or log focal scale outside (-700, 700)");
    return TCV_BA_ERR_ARGUMENT;
  }
  HIP_TRY(hipSetDevice(ctx->device));
  fc = *c;
  fc.precision = TCV_BA_FP64;
  fc.flags &= ~TCV_BA_FLAG_IMPLICIT;
  const size_t ss = tcv_ba_scene_size(c), is = tcv_ba_index_size(c);
  TRY(ensure(ctx->probe, ss * 8));
  TRY(ensure(ctx->index, is * 4));
  HIP_TRY(hipMemcpyAsync(ctx->probe.ptr, scene, ss * 8, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(ctx->index.ptr, index, is * 4, hipMemcpyHostToDevice, st));
  TRY(setup(ctx, &fc, 1, (const double*)ctx->probe.ptr, (const uint32_t*)ctx->index.ptr, st, q,
            rep, true));
  q.tol = tolerance;
  q.exact_gauge = 1;
  q.exact = 1;
  ops = select_ops(&fc);
  hipLaunchKernelGGL(k_init_bundle, dim3(1), dim3(64), 0, st, q);
  const double lf = c->shared_focal ? log_focal : 0.0;
  HIP_TRY(hipMemcpyAsync(q.state + S_LOG_FOCAL, &lf, 8, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_cost<false>, dim3(q.n_cost_blocks), dim3(COST_OBS), 0, st, q);
  hipLaunchKernelGGL(k_finish_cost<false>, dim3(1), dim3(COST_OBS), 0, st, q);
  HIP_TRY(hipGetLastError());
  TRY(read_state(ctx, q, st, rep));
  const double input_error = word(ctx, 0, S_INPUT_ERROR);
  if (input_error != 0) {
    set_error(input_error == 1 ? "GPU BA gauge requires a nonzero baseline"
                               : "GPU BA observations must initially project in front of their cameras");
    return TCV_BA_ERR_ARGUMENT;
  }
  return TCV_BA_OK;
}

// One undamped solve of the exact-Hessian system (the right-hand side is -gradient, or q.rhs_in):
// assembly, then the direct route (TCV_BA_FLAG_SCHUR_CACHE) and/or PCG to q.tol within `max_cg`
// iterations. True when the solve converged (or the right-hand side was zero).
int exact_solve(tcv_ba_context* ctx, const tcv_ba_control* fc, const Problem& q, const Ops& ops,
                hipStream_t st, uint32_t max_cg, tcv_ba_report* rep, bool& converged) {
  hipLaunchKernelGGL(k_newton_begin, dim3(1), dim3(64), 0, st, q);
  ops.exact_passes(q, st);
  ops.init_cg(q, st);
  HIP_TRY(hipGetLastError());
  TRY(read_state(ctx, q, st, rep));
  if (word(ctx, 0, S_PCG_FAILED) != 0) {
    converged = false;
    return TCV_BA_OK;
  }
  if (word(ctx, 0, S_PCG_ACTIVE) != 0) {
    TRY(solve_linear(ctx, fc, q, ops, st, rep, 32, max_cg, q.tol));
    if (any_word(ctx, q, S_PCG_ACTIVE))
      hipLaunchKernelGGL(k_stop_pcg, dim3(1), dim3(64), 0, st, q);
    TRY(read_state(ctx, q, st, rep));
    converged = word(ctx, 0, S_CONVERGED) != 0 && word(ctx, 0, S_PCG_FAILED) == 0;
  } else {
    converged = true;  // zero right-hand side: x = 0
  }
  if (rep) rep->cg_iterations += (uint64_t)word(ctx, 0, S_STEPS);
  return TCV_BA_OK;
}

// The right-hand side of the adjoint solve (layout of Problem::rhs_in) from a cotangent on the
// solve's outputs (R row-major and t per camera, X per point, then the log focal scale):
//  * each variable camera's chart cotangent for R' = exp(omega) R, t' = exp(omega) t + v:
//    omega: sum_j R_j x dR_j + t x dt (R_j, dR_j the columns), v: dt; fixed cameras do not move;
//  * then the transpose of the derivative of the gauge map the forward applies to every candidate:
//    directional points are renormalised (v_X - X (X . v_X), X unit), and with one fixed camera the
//    similarity about the anchor centre o rescales the baseline b to the scale view, which adds
//    -grad(b) (n . v) / b with n the similarity generator (X - o per finite point, -R (C - o) on each
//    variable camera's v) and grad(b) = -R_s (C_s - o) / b on the scale view's v.
std::vector<double> adjoint_rhs(const tcv_ba_control* c, const double* scene,
                                const uint32_t* index, const double* cot) {
  const size_t C = c->cameras, P = c->points, M = c->observations;
  const size_t sh = c->shared_focal ? 1 : 0, m = 6 * C + sh;
  const uint32_t* active = index + 3 * M + P + C + 2;
  const double* X = scene + 12 * C;
  std::vector<double> rhs(m + 3 * P, 0.0);
  auto cross = [](const double* a, const double* b, double* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
  };
  auto center = [&](size_t k, double* o) {
    const double* p = scene + 12 * k;
    for (int i = 0; i < 3; i++) o[i] = -(p[i] * p[9] + p[3 + i] * p[10] + p[6 + i] * p[11]);
  };
  for (size_t k = 0; k < C; k++) {
    if (!active[k]) continue;
    const double* p = scene + 12 * k;
    const double* d = cot + 12 * k;
    double w[3] = {0, 0, 0}, x[3];
    for (int j = 0; j < 3; j++) {
      const double Rj[3] = {p[j], p[3 + j], p[6 + j]}, dRj[3] = {d[j], d[3 + j], d[6 + j]};
      cross(Rj, dRj, x);
      for (int i = 0; i < 3; i++) w[i] += x[i];
    }
    cross(p + 9, d + 9, x);
    for (int i = 0; i < 3; i++) {
      rhs[6 * k + i] = w[i] + x[i];
      rhs[6 * k + 3 + i] = d[9 + i];
    }
  }
  if (sh) rhs[6 * C] = cot[12 * C + 3 * P];
  const size_t first_dir = P - c->directional_points;
  for (size_t p = 0; p < P; p++) {
    double* r = &rhs[m + 3 * p];
    for (int i = 0; i < 3; i++) r[i] = cot[12 * C + 3 * p + i];
    if (p >= first_dir) {
      const double* xp = X + 3 * p;
      const double n2 = xp[0] * xp[0] + xp[1] * xp[1] + xp[2] * xp[2];
      if (n2 > 0) {
        const double dot = (xp[0] * r[0] + xp[1] * r[1] + xp[2] * r[2]) / n2;
        for (int i = 0; i < 3; i++) r[i] -= xp[i] * dot;
      }
    }
  }
  if (c->normalize_scale) {
    double o[3], cs[3];
    center(c->anchor, o);
    center(c->scale_view, cs);
    const double d0 = cs[0] - o[0], d1 = cs[1] - o[1], d2 = cs[2] - o[2];
    const double b = std::sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    double nv = 0;
    for (size_t p = 0; p < first_dir; p++)
      for (int i = 0; i < 3; i++) nv += (X[3 * p + i] - o[i]) * rhs[m + 3 * p + i];
    for (size_t k = 0; k < C; k++) {
      if (!active[k]) continue;
      const double* R = scene + 12 * k;
      double ck[3];
      center(k, ck);
      for (int i = 0; i < 3; i++) {
        const double n_i = -(R[i * 3] * (ck[0] - o[0]) + R[i * 3 + 1] * (ck[1] - o[1]) +
                             R[i * 3 + 2] * (ck[2] - o[2]));
        nv += n_i * rhs[6 * k + 3 + i];
      }
    }
    const double* Rs = scene + 12 * (size_t)c->scale_view;
    const double dd[3] = {d0, d1, d2};
    for (int i = 0; i < 3; i++) {
      const double grad = -(Rs[i * 3] * dd[0] + Rs[i * 3 + 1] * dd[1] + Rs[i * 3 + 2] * dd[2]) / b;
      rhs[6 * (size_t)c->scale_view + 3 + i] -= grad * nv / b;
    }
  }
  return rhs;
}

// fn(begin, end) over [0, n) in contiguous chunks on the host's threads (at most 16); the chunks
// write disjoint outputs, so the result does not depend on the split.
template <class Fn>
void parallel_for(size_t n, Fn fn) {
  const size_t threads = std::max<size_t>(1, std::min<size_t>(16, std::thread::hardware_concurrency()));
  const size_t parts = std::min(threads, std::max<size_t>(1, n / 65536));
  if (parts <= 1) return fn(size_t(0), n);
  std::vector<std::thread> pool;
  for (size_t k = 0; k < parts; k++)
    pool.emplace_back(fn, n * k / parts, n * (k + 1) / parts);
  for (std::thread& t : pool) t.join();
}

// Tiles of whole points of one bundle, as build_plan forms them (at most TILE observations and
// TILE points; a longer point alone): the first point of each, then P.
std::vector<uint32_t> point_tiles(const uint32_t* ps, uint32_t P) {
  std::vector<uint32_t> tiles{0};
  uint32_t first = 0;
  for (uint32_t p = 1; p < P; p++)
    if (ps[p + 1] - ps[first] > (uint32_t)TILE || p - first >= (uint32_t)TILE) {
      tiles.push_back(p);
      first = p;
    }
  tiles.push_back(P);
  return tiles;
}

// Per point the extremes of its track's cameras, plain and wrapped round the camera loop (a camera
// k < C/2 counted as k + C), as k_onepass_windows measures a window; lo > hi for a point without
// observations.
struct TrackSpan {
  uint32_t lo = UINT32_MAX, hi = 0, wlo = UINT32_MAX, whi = 0;
};

// The one-pass rows of the points taken in `order` (given point of each), whose track offsets in
// that order are `ps`: the sum of the window spans of the one-pass workgroups (OP_TILES consecutive
// tiles each), as set-up will count them; `fits` says whether every window spans at most OP_SPAN
// cameras, the condition for the one-pass product.
uint64_t window_rows(const std::vector<TrackSpan>& ts, const uint32_t* ps,
                     const std::vector<uint32_t>& order, bool& fits) {
  const uint32_t P = (uint32_t)ts.size();
  const std::vector<uint32_t> tiles = point_tiles(ps, P);
  const uint32_t n_tiles = (uint32_t)tiles.size() - 1;
  const uint32_t n_op = (n_tiles + OP_TILES - 1) / OP_TILES;
  std::vector<uint32_t> span(n_op, 0);
  parallel_for(n_op, [&](size_t w0, size_t w1) {
    for (size_t w = w0; w < w1; w++) {
      const uint32_t t0 = (uint32_t)w * OP_TILES, t1 = std::min<uint32_t>(t0 + OP_TILES, n_tiles);
      TrackSpan x;
      for (uint32_t j = tiles[t0]; j < tiles[t1]; j++) {
        const TrackSpan& y = ts[order[j]];
        x.lo = std::min(x.lo, y.lo);
        x.hi = std::max(x.hi, y.hi);
        x.wlo = std::min(x.wlo, y.wlo);
        x.whi = std::max(x.whi, y.whi);
      }
      if (x.lo <= x.hi) span[w] = std::min(x.hi - x.lo, x.whi - x.wlo) + 1;
    }
  });
  uint64_t rows = 0;
  fits = true;
  for (uint32_t v : span) {
    rows += v;
    fits = fits && v <= (uint32_t)OP_SPAN;
  }
  return rows;
}

// The given point of each new one: ascending lowest camera of the track, then ascending highest
// camera (C and C for a track without observations), by two stable counting sorts within each
// group, the directional points (from first_dir on) after the others.
std::vector<uint32_t> sort_points(const std::vector<TrackSpan>& ts, uint32_t C,
                                  uint32_t first_dir) {
  const uint32_t P = (uint32_t)ts.size();
  auto lo = [&](uint32_t p) { return std::min(ts[p].lo, C); };
  auto hi = [&](uint32_t p) { return ts[p].lo > ts[p].hi ? C : ts[p].hi; };
  std::vector<uint32_t> order(P), by_hi(P), count(C + 2);
  for (const auto& [g0, g1] : {std::pair<uint32_t, uint32_t>{0, first_dir}, {first_dir, P}}) {
    std::fill(count.begin(), count.end(), 0u);
    for (uint32_t p = g0; p < g1; p++) count[hi(p) + 1]++;
    for (uint32_t k = 0; k <= C; k++) count[k + 1] += count[k];
    for (uint32_t p = g0; p < g1; p++) by_hi[g0 + count[hi(p)]++] = p;
    std::fill(count.begin(), count.end(), 0u);
    for (uint32_t p = g0; p < g1; p++) count[lo(p) + 1]++;
    for (uint32_t k = 0; k <= C; k++) count[k + 1] += count[k];
    for (uint32_t j = g0; j < g1; j++) order[g0 + count[lo(by_hi[j])]++] = by_hi[j];
  }
  return order;
}

// Whether the solve entries consider another order for the points of a one-bundle problem: more
// cameras than one window holds, and a solve that runs the one-pass product (a bundle the direct
// route takes runs it only for PCG hand-overs, so its points stay as given).
bool may_order(const tcv_ba_control* c) {
  if (c->cameras <= (uint32_t)OP_SPAN || c->points == 0 || c->observations == 0) return false;
  return (c->flags & TCV_BA_FLAG_DIRECT) == 0 || c->cameras > DENSE_LARGE_CAMERAS;
}

// The order the solve entries give the points of one problem (tracks `ts`, given offsets `ps`):
// by the lowest, then the highest camera of their tracks (sort_points), taken when its one-pass
// windows fit and the given order's do not, or take at least an eighth more rows (partial rows
// written per product and summed per camera). True with `order` (the given point of each new one) and `nps` (the new
// track offsets) filled; false to keep the given order.
bool choose_order(const std::vector<TrackSpan>& ts, const uint32_t* ps, uint32_t C,
                  uint32_t first_dir, std::vector<uint32_t>& order, std::vector<uint32_t>& nps) {
  const uint32_t P = (uint32_t)ts.size();
  order = sort_points(ts, C, first_dir);
  nps.assign(P + 1, 0);
  for (uint32_t j = 0; j < P; j++) nps[j + 1] = nps[j] + (ps[order[j] + 1] - ps[order[j]]);
  std::vector<uint32_t> given(P);
  for (uint32_t p = 0; p < P; p++) given[p] = p;
  bool fits = false, nfits = false;
  const uint64_t rows = window_rows(ts, ps, given, fits);
  const uint64_t nrows = window_rows(ts, nps.data(), order, nfits);
  return nfits && (!fits || 8 * nrows <= 7 * rows);
}

// Host arrays of one problem whose points are not in an order the one-pass product can take
// (points listed as their tracks were created, say, rather than by the cameras that see them) would
// run the two sweeps: on the 2,000-camera trajectory scene with its points shuffled, 495 us per
// FP64 product against 189 us; points in an order that fits but spreads each workgroup's tracks
// over more cameras write more partial rows. So the points are ordered by the lowest, then the
// highest camera of their tracks (sort_points; the directional points stay after the others, and
// a point without observations goes last in its group) when choose_order takes that order, their
// observations follow them in their given order and each camera's list is rebuilt ascending:
// `oscene` / `oindex` receive the reordered arrays and `perm` the given point of each new one, and
// the result is true. Nothing changes (false, `perm` empty) when may_order or choose_order keeps
// the given order or when the index is not one the engine accepts (its own checks then report it
// as before). The
// reordered problem is the same problem: only sums over points change their order.
bool order_points(const tcv_ba_control* c, const double* scene, const uint32_t* index,
                  std::vector<double>& oscene, std::vector<uint32_t>& oindex,
                  std::vector<uint32_t>& perm) {
  perm.clear();
  const uint32_t C = c->cameras, P = c->points, M = c->observations;
  if (!may_order(c)) return false;
  const uint32_t *ocam = index, *opt = index + M, *ps = index + 2 * (size_t)M,
                 *cs = ps + P + 1, *co = cs + C + 1, *act = co + M;
  // the offsets' checks of build_plan and those of k_check_points / k_check_tasks: observations
  // of each point name it and a camera, each camera's list names its own observations ascending
  if (ps[0] != 0 || ps[P] != M || cs[0] != 0 || cs[C] != M) return false;
  for (uint32_t p = 0; p < P; p++)
    if (ps[p + 1] < ps[p]) return false;
  for (uint32_t k = 0; k < C; k++)
    if (cs[k + 1] < cs[k]) return false;
  std::vector<char> bad(2, 0);
  parallel_for(P, [&](size_t p0, size_t p1) {
    for (size_t p = p0; p < p1; p++)
      for (uint32_t i = ps[p]; i < ps[p + 1]; i++)
        if (opt[i] != p || ocam[i] >= C) bad[0] = 1;
  });
  parallel_for(C, [&](size_t k0, size_t k1) {
    for (size_t k = k0; k < k1; k++)
      for (uint32_t n = cs[k]; n < cs[k + 1]; n++)
        if (co[n] >= M || ocam[co[n]] != k || (n > cs[k] && co[n - 1] >= co[n])) bad[1] = 1;
  });
  if (bad[0] || bad[1]) return false;
  std::vector<TrackSpan> ts(P);
  const uint32_t half = C / 2;
  parallel_for(P, [&](size_t p0, size_t p1) {
    for (size_t p = p0; p < p1; p++)
      for (uint32_t i = ps[p]; i < ps[p + 1]; i++) {
        const uint32_t cam = ocam[i], v = cam < half ? cam + C : cam;
        ts[p].lo = std::min(ts[p].lo, cam);
        ts[p].hi = std::max(ts[p].hi, cam);
        ts[p].wlo = std::min(ts[p].wlo, v);
        ts[p].whi = std::max(ts[p].whi, v);
      }
  });
  std::vector<uint32_t> order, hnps;
  if (!choose_order(ts, ps, C, P - c->directional_points, order, hnps)) return false;
  // the reordered index: observations in the new point order, then each camera's list, its new
  // observations ascending
  oindex.assign(tcv_ba_index_size(c), 0);
  uint32_t *ncam = oindex.data(), *npt = ncam + M, *nps = npt + M, *ncs = nps + P + 1,
           *nco = ncs + C + 1, *nact = nco + M;
  std::memcpy(nps, hnps.data(), ((size_t)P + 1) * 4);
  parallel_for(P, [&](size_t j0, size_t j1) {
    for (size_t j = j0; j < j1; j++)
      for (uint32_t i = ps[order[j]], n = nps[j]; i < ps[order[j] + 1]; i++, n++) {
        ncam[n] = ocam[i];
        npt[n] = (uint32_t)j;
      }
  });
  std::memcpy(ncs, cs, ((size_t)C + 1) * 4);
  std::vector<uint32_t> fill(cs, cs + C);
  for (uint32_t n = 0; n < M; n++) nco[fill[ncam[n]]++] = n;
  std::memcpy(nact, act, (size_t)C * 4);
  // the reordered scene: poses and models as given, points and pixels moved
  oscene.resize(tcv_ba_scene_size(c));
  const size_t o_pt = (size_t)12 * C, o_model = o_pt + 3 * (size_t)P, o_pix = o_model + 6 * (size_t)C;
  std::memcpy(oscene.data(), scene, o_pt * 8);
  std::memcpy(oscene.data() + o_model, scene + o_model, (size_t)C * 6 * 8);
  parallel_for(P, [&](size_t j0, size_t j1) {
    for (size_t j = j0; j < j1; j++) {
      const size_t g = order[j];
      for (int k = 0; k < 3; k++) oscene[o_pt + 3 * j + k] = scene[o_pt + 3 * g + k];
      for (uint32_t i = ps[g], n = nps[j]; i < ps[g + 1]; i++, n++) {
        oscene[o_pix + 2 * (size_t)n] = scene[o_pix + 2 * (size_t)i];
        oscene[o_pix + 2 * (size_t)n + 1] = scene[o_pix + 2 * (size_t)i + 1];
      }
    }
  });
  perm = std::move(order);
  return true;
}

// The bits of a camera index below C (the radix sort's key width).
int camera_bits(uint32_t C) {
  int bits = 1;
  while (bits < 32 && (C >> bits) != 0) bits++;
  return bits;
}

// order_points for device arrays: the same choice of order (choose_order), from the tracks' camera
// extremes measured on the device; when it takes the new order, the reordered scene goes into
// `nscene` and the index into `nindex`, ctx->ord_perm keeps the given point of each new one, and
// `moved` is set (unpermute_points puts the solved points back). Nothing changes when the given
// order is kept (may_order, choose_order) or when the index is not one the engine accepts (its
// checks then report it as before).
int order_device(tcv_ba_context* ctx, const tcv_ba_control* c, const double* scene,
                 const uint32_t* index, hipStream_t st, DevBuf& nscene_buf, DevBuf& nindex_buf,
                 bool& moved) {
  moved = false;
  const uint32_t C = c->cameras, P = c->points, M = c->observations;
  if (!may_order(c)) return TCV_BA_OK;
  const uint32_t *ocam = index, *opt = index + M, *ps = index + 2 * (size_t)M,
                 *cs = ps + P + 1, *co = cs + C + 1, *act = co + M;
  std::vector<uint32_t> hps(P + 1), hcs(C + 1);
  HIP_TRY(hipMemcpyAsync(hps.data(), ps, (P + 1) * 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(hcs.data(), cs, (C + 1) * 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (hps[0] != 0 || hps[P] != M || hcs[0] != 0 || hcs[C] != M) return TCV_BA_OK;
  for (uint32_t p = 0; p < P; p++)
    if (hps[p + 1] < hps[p]) return TCV_BA_OK;
  for (uint32_t k = 0; k < C; k++)
    if (hcs[k + 1] < hcs[k]) return TCV_BA_OK;
  // the tracks' camera extremes and the checks of the words the reordering reads
  static_assert(sizeof(TrackSpan) == 16, "four words per point");
  TRY(ensure(ctx->ord_key, (size_t)P * 16 + 16));
  uint32_t* span = (uint32_t*)ctx->ord_key.ptr;
  uint32_t* flag = span + 4 * (size_t)P;
  HIP_TRY(hipMemsetAsync(flag, 0, 4, st));
  hipLaunchKernelGGL(k_order_spans, dim3(blocks_for(P, ORDER_THREADS)), dim3(ORDER_THREADS), 0, st,
                     ocam, opt, ps, P, C, span, flag);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_order_lists, dim3(C), dim3(ORDER_THREADS), 0, st, ocam, cs, co, M, flag);
  HIP_TRY(hipGetLastError());
  std::vector<TrackSpan> ts(P);
  uint32_t bad = 0;
  HIP_TRY(hipMemcpyAsync(ts.data(), span, (size_t)P * 16, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(&bad, flag, 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (bad) return TCV_BA_OK;
  std::vector<uint32_t> order, nps;
  if (!choose_order(ts, hps.data(), C, P - c->directional_points, order, nps)) return TCV_BA_OK;
  // the reordered index and scene
  const size_t ss = tcv_ba_scene_size(c), is = tcv_ba_index_size(c);
  TRY(upload(ctx->ord_perm, order, st));
  TRY(ensure(nscene_buf, ss * 8));
  TRY(ensure(nindex_buf, is * 4));
  double* nscene = (double*)nscene_buf.ptr;
  uint32_t* nidx = (uint32_t*)nindex_buf.ptr;
  uint32_t *ncam = nidx, *npt = ncam + M, *nps_d = npt + M, *ncs = nps_d + P + 1,
           *nco = ncs + C + 1, *nact = nco + M;
  HIP_TRY(hipMemcpyAsync(nps_d, nps.data(), ((size_t)P + 1) * 4, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(ncs, cs, ((size_t)C + 1) * 4, hipMemcpyDeviceToDevice, st));
  HIP_TRY(hipMemcpyAsync(nact, act, (size_t)C * 4, hipMemcpyDeviceToDevice, st));
  const size_t o_pt = (size_t)12 * C, o_model = o_pt + 3 * (size_t)P;
  const size_t o_pix = o_model + 6 * (size_t)C;
  HIP_TRY(hipMemcpyAsync(nscene, scene, o_pt * 8, hipMemcpyDeviceToDevice, st));
  HIP_TRY(hipMemcpyAsync(nscene + o_model, scene + o_model, (size_t)C * 6 * 8,
                         hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(k_order_gather, dim3(blocks_for(P, ORDER_THREADS)), dim3(ORDER_THREADS), 0,
                     st, (const uint32_t*)ctx->ord_perm.ptr, ps, nps_d, ocam, scene + o_pt,
                     scene + o_pix, P, ncam, npt, nscene + o_pt, nscene + o_pix);
  HIP_TRY(hipGetLastError());
  // each camera's list: its new observations ascending, a stable sort of the observations by
  // camera
  const int bits = camera_bits(C);
  size_t temp = 0;
  HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, temp, (const uint32_t*)nullptr,
                                             (uint32_t*)nullptr, (const uint32_t*)nullptr,
                                             (uint32_t*)nullptr, (int)M, 0, bits, st));
  // the sort's temporary storage 256-byte aligned after the ids and sorted cameras
  const size_t words = ((2 * (size_t)M + 63) / 64) * 64;
  TRY(ensure(ctx->ord_sort, words * 4 + temp));
  uint32_t* ids = (uint32_t*)ctx->ord_sort.ptr;
  uint32_t* sorted_cam = ids + M;
  hipLaunchKernelGGL(k_order_iota, dim3(blocks_for(M, ORDER_THREADS)), dim3(ORDER_THREADS), 0, st,
                     ids, M);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipcub::DeviceRadixSort::SortPairs((void*)(ids + words), temp, (const uint32_t*)ncam,
                                             sorted_cam, (const uint32_t*)ids, nco, (int)M, 0,
                                             bits, st));
  moved = true;
  return TCV_BA_OK;
}

// The points of a solve of reordered arrays back in the given order: `out` (device) holds the
// solve's output, `staging` (device, 3P doubles) takes its points while k_order_scatter writes
// each back to its given index (ctx->ord_perm).
int unpermute_points(tcv_ba_context* ctx, const tcv_ba_control* c, double* out, double* staging,
                     hipStream_t st) {
  const uint32_t P = c->points;
  double* pts = out + (size_t)12 * c->cameras;
  HIP_TRY(hipMemcpyAsync(staging, pts, (size_t)P * 24, hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(k_order_scatter, dim3(blocks_for(P, ORDER_THREADS)), dim3(ORDER_THREADS), 0,
                     st, (const uint32_t*)ctx->ord_perm.ptr, (const double*)staging, P, pts);
  HIP_TRY(hipGetLastError());
  return TCV_BA_OK;
}

// The device bytes order_device may hold beyond the solve's for one problem: the reordered scene
// and index, the tracks' extremes, the permutation and the radix sort's arrays.
int order_bytes(const tcv_ba_control* c, size_t& scene, size_t& index, size_t& key, size_t& perm,
                size_t& sort) {
  const uint32_t M = c->observations;
  size_t temp = 0;
  HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, temp, (const uint32_t*)nullptr,
                                             (uint32_t*)nullptr, (const uint32_t*)nullptr,
                                             (uint32_t*)nullptr, (int)M, 0, camera_bits(c->cameras),
                                             (hipStream_t)0));
  scene = tcv_ba_scene_size(c) * 8;
  index = tcv_ba_index_size(c) * 4;
  key = (size_t)c->points * 16 + 16;
  perm = (size_t)c->points * 4;
  sort = ((2 * (size_t)M + 63) / 64) * 64 * 4 + temp;
  return TCV_BA_OK;
}

}  // namespace

// ------------------------------------------------------------------ C ABI

extern "C" {

size_t tcv_ba_output_size(const tcv_ba_control* c) {
  return (size_t)c->cameras * 12 + (size_t)c->points * 3 + TCV_BA_STATE_WORDS +
         (c->shared_focal ? 1 : 0);
}
size_t tcv_ba_scene_size(const tcv_ba_control* c) {
  return (size_t)c->cameras * 18 + (size_t)c->points * 3 + (size_t)c->observations * 2;
}
size_t tcv_ba_index_size(const tcv_ba_control* c) {
  return 3 * (size_t)c->observations + c->points + 2 * (size_t)c->cameras + 2;
}

const char* tcv_ba_status_string(int status) {
  switch (status) {
    case TCV_BA_OK: return "ok";
    case TCV_BA_ERR_ARGUMENT: return "Invalid GPU BA scene/options";
    case TCV_BA_ERR_INDEX: return "GPU BA index arrays are not a packed scene of these sizes";
    case TCV_BA_ERR_DEVICE: return "GPU BA device call failed";
    case TCV_BA_ERR_MEMORY: return "GPU BA needs more device memory than is free";
    case TCV_BA_ERR_NO_DEVICE: return "no gfx950 HIP device is visible";
    case TCV_BA_ERR_SOLVE: return "GPU BA adjoint system was not solved";
    default: return "unknown status";
  }
}

const char* tcv_ba_last_error(void) { return g_last_error.c_str(); }

int tcv_ba_create(int device, tcv_ba_context** out) {
  if (!out) return TCV_BA_ERR_ARGUMENT;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) {
    set_error("no HIP device " + std::to_string(device));
    return TCV_BA_ERR_NO_DEVICE;
  }
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    set_error(std::string("device is ") + prop.gcnArchName + ", the engine is built for gfx950");
    return TCV_BA_ERR_NO_DEVICE;
  }
  HIP_TRY(hipSetDevice(device));
  tcv_ba_context* ctx = new tcv_ba_context();
  ctx->device = device;
  hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipHostMalloc((void**)&ctx->h_error, 8, hipHostMallocDefault);
  if (e != hipSuccess) {
    set_error(std::string("context set-up: ") + hipGetErrorName(e));
    tcv_ba_destroy(ctx);
    return TCV_BA_ERR_DEVICE;
  }
  if (ensure_host_state(ctx, S_WORDS) != TCV_BA_OK) {
    tcv_ba_destroy(ctx);
    return TCV_BA_ERR_DEVICE;
  }
  *out = ctx;
  return TCV_BA_OK;
}

int tcv_ba_release(tcv_ba_context* ctx) {
  if (!ctx) return TCV_BA_ERR_ARGUMENT;
  std::lock_guard<std::mutex> guard(ctx->lock);
  return release_buffers(ctx);
}

void tcv_ba_destroy(tcv_ba_context* ctx) {
  if (!ctx) return;
  {
    // wait for a call still running on another thread
    std::lock_guard<std::mutex> guard(ctx->lock);
    (void)hipSetDevice(ctx->device);
    release_buffers(ctx);
    for (hipEvent_t e : ctx->events) (void)hipEventDestroy(e);
    for (hipEvent_t e : ctx->tune_events) (void)hipEventDestroy(e);
    if (ctx->h_state) (void)hipHostFree(ctx->h_state);
    if (ctx->h_error) (void)hipHostFree(ctx->h_error);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  }
  delete ctx;
}

uint64_t tcv_ba_device_bytes(const tcv_ba_context* ctx) {
  if (!ctx) return 0;
  std::lock_guard<std::mutex> guard(ctx->lock);
  return held_bytes(ctx);
}


static int body_ba_solve_device(tcv_ba_context* ctx, const tcv_ba_control* c, const double* scene,
                                const uint32_t* index, double* out, tcv_ba_report* rep,
                                void* stream) {
  if (!ctx || !c || !scene || !index || !out) return TCV_BA_ERR_ARGUMENT;
  if (rep) std::memset(rep, 0, sizeof *rep);
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  TRY(check_control(c));
  HIP_TRY(hipSetDevice(ctx->device));
  // points in the order the one-pass product takes best (order_device); the output's points go back
  // to the given order. Under a working budget the given order is kept: set-up admits the solve's
  // own arrays only.
  bool moved = false;
  if (!ctx->budget) TRY(order_device(ctx, c, scene, index, st, ctx->probe, ctx->index, moved));
  if (!moved) return solve_device(ctx, c, 1, scene, index, out, rep, nullptr, st);
  TRY(solve_device(ctx, c, 1, (const double*)ctx->probe.ptr, (const uint32_t*)ctx->index.ptr, out,
                   rep, nullptr, st));
  // the solved points staged in the reordered scene's point words
  TRY(unpermute_points(ctx, c, out, (double*)ctx->probe.ptr + (size_t)12 * c->cameras, st));
  HIP_TRY(hipStreamSynchronize(st));
  return TCV_BA_OK;
}

static int body_ba_solve(tcv_ba_context* ctx, const tcv_ba_control* c, const double* scene,
                         const uint32_t* index, double* out, tcv_ba_report* rep) {
  if (!ctx || !c || !scene || !index || !out) return TCV_BA_ERR_ARGUMENT;
  if (rep) std::memset(rep, 0, sizeof *rep);
  TRY(check_control(c));
  HIP_TRY(hipSetDevice(ctx->device));
  const size_t ss = tcv_ba_scene_size(c), is = tcv_ba_index_size(c), os = tcv_ba_output_size(c);
  // the residency plan under the working budget, with the staging of the host arrays (and, when
  // the points may be reordered, of the reordered arrays), before anything is allocated
  int residency = RESIDENT;
  if (ctx->budget) {
    Plan plan;
    bool streamed = false;
    std::vector<Need> staging = {{&ctx->probe, ss * 8}, {&ctx->index, is * 4}, {&ctx->out, os * 8}};
    if (may_order(c)) {
      size_t b_scene, b_index, b_key, b_perm, b_sort;
      TRY(order_bytes(c, b_scene, b_index, b_key, b_perm, b_sort));
      staging.insert(staging.end(), {{&ctx->ord_scene, b_scene}, {&ctx->ord_index, b_index},
                                     {&ctx->ord_key, b_key}, {&ctx->ord_perm, b_perm},
                                     {&ctx->ord_sort, b_sort}});
    }
    TRY(host_plan(c, 1, index, c->cameras, c->points, c->observations, false, plan));
    TRY(admit(ctx, c, 1, c->cameras, c->points, c->observations, plan, false, staging, streamed));
    residency = streamed ? STREAMED : RESIDENT;
  }
  TRY(ensure(ctx->probe, ss * 8));  // staging of the host scene
  TRY(ensure(ctx->index, is * 4));
  TRY(ensure(ctx->out, os * 8));
  hipStream_t st = ctx->stream;
  HIP_TRY(hipMemcpyAsync(ctx->probe.ptr, scene, ss * 8, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(ctx->index.ptr, index, is * 4, hipMemcpyHostToDevice, st));
  // points in the order the one-pass product takes best (order_device, into the context's
  // reordered arrays); the output's points go back to the given order
  bool moved = false;
  TRY(order_device(ctx, c, (const double*)ctx->probe.ptr, (const uint32_t*)ctx->index.ptr, st,
                   ctx->ord_scene, ctx->ord_index, moved));
  const DevBuf& s_scene = moved ? ctx->ord_scene : ctx->probe;
  const DevBuf& s_index = moved ? ctx->ord_index : ctx->index;
  TRY(solve_device(ctx, c, 1, (const double*)s_scene.ptr, (const uint32_t*)s_index.ptr,
                   (double*)ctx->out.ptr, rep, nullptr, st, residency));
  if (moved)
    TRY(unpermute_points(ctx, c, (double*)ctx->out.ptr,
                         (double*)ctx->ord_scene.ptr + (size_t)12 * c->cameras, st));
  HIP_TRY(hipMemcpyAsync(out, ctx->out.ptr, os * 8, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (rep) rep->device_bytes = held_bytes(ctx);
  return TCV_BA_OK;
}

static int body_ba_solve_batch(tcv_ba_context* ctx, uint32_t count, const tcv_ba_control* controls,
                               const double* const* scenes, const uint32_t* const* indices,
                               double* const* outs, tcv_ba_report* rep, double* finish_ms) {
  if (!ctx || count < 1 || !controls || !scenes || !indices || !outs) return TCV_BA_ERR_ARGUMENT;
  if (rep) std::memset(rep, 0, sizeof *rep);
  for (uint32_t b = 0; b < count; b++) {
    TRY(check_control(&controls[b]));
    if (!scenes[b] || !indices[b] || !outs[b]) return TCV_BA_ERR_ARGUMENT;
  }
  HIP_TRY(hipSetDevice(ctx->device));
  const double t0 = now_ms();
  // The batch as one block-diagonal packed problem: every section concatenated in bundle order,
  // ids shifted by the bundle's first camera, point and observation.
  uint64_t C = 0, P = 0, M = 0;
  for (uint32_t b = 0; b < count; b++) {
    C += controls[b].cameras;
    P += controls[b].points;
    M += controls[b].observations;
  }
  if (3 * M + P + 2 * C + 2 > 0xffffffffull || 18 * C + 3 * P + 2 * M > 0xffffffffull) {
    set_error("batch exceeds the u32 index bounds");
    return TCV_BA_ERR_ARGUMENT;
  }
  std::vector<double> scene(18 * C + 3 * P + 2 * M);
  std::vector<uint32_t> index(3 * M + P + 2 * C + 2);
  const size_t o_pt = 12 * C, o_model = 12 * C + 3 * P, o_pix = 18 * C + 3 * P;
  // each bundle's points in the order its solve alone takes (order_points), so that the batch
  // equals its bundles alone; perms[b] is empty where the given order stays
  std::vector<std::vector<uint32_t>> perms(count);
  const size_t i_op = M, i_ps = 2 * M, i_cs = 2 * M + P + 1, i_co = 2 * M + P + C + 2,
               i_ca = 3 * M + P + C + 2;
  uint64_t c0 = 0, p0 = 0, m0 = 0;
  for (uint32_t b = 0; b < count; b++) {
    const tcv_ba_control& k = controls[b];
    const double* s = scenes[b];
    const uint32_t* x = indices[b];
    std::vector<double> oscene;
    std::vector<uint32_t> oindex;
    if (order_points(&k, s, x, oscene, oindex, perms[b])) {
      s = oscene.data();
      x = oindex.data();
    } else {
      perms[b].clear();
    }
    const size_t Cb = k.cameras, Pb = k.points, Mb = k.observations;
    std::memcpy(&scene[12 * c0], s, 12 * Cb * 8);
    std::memcpy(&scene[o_pt + 3 * p0], s + 12 * Cb, 3 * Pb * 8);
    std::memcpy(&scene[o_model + 6 * c0], s + 12 * Cb + 3 * Pb, 6 * Cb * 8);
    std::memcpy(&scene[o_pix + 2 * m0], s + 18 * Cb + 3 * Pb, 2 * Mb * 8);
    for (size_t i = 0; i < Mb; i++) {
      if (x[i] >= Cb || x[Mb + i] >= Pb || x[2 * Mb + Pb + Cb + 2 + i] >= Mb) {
        set_error("bundle " + std::to_string(b) + ": index entry out of range");
        return TCV_BA_ERR_INDEX;
      }
      index[m0 + i] = (uint32_t)(x[i] + c0);
      index[i_op + m0 + i] = (uint32_t)(x[Mb + i] + p0);
      index[i_co + m0 + i] = (uint32_t)(x[2 * Mb + Pb + Cb + 2 + i] + m0);
    }
    for (size_t p = 0; p <= Pb; p++) index[i_ps + p0 + p] = (uint32_t)(x[2 * Mb + p] + m0);
    for (size_t c = 0; c <= Cb; c++) index[i_cs + c0 + c] = (uint32_t)(x[2 * Mb + Pb + 1 + c] + m0);
    for (size_t c = 0; c < Cb; c++) index[i_ca + c0 + c] = x[3 * Mb + Pb + Cb + 2 + c];
    c0 += Cb;
    p0 += Pb;
    m0 += Mb;
  }
  const size_t words = 26 + (controls[0].shared_focal ? 1 : 0);
  const size_t out_size = 12 * C + 3 * P + count * words;
  int residency = RESIDENT;
  if (ctx->budget) {
    Plan plan;
    bool streamed = false;
    TRY(host_plan(controls, count, index.data(), (uint32_t)C, (uint32_t)P, (uint32_t)M, false,
                  plan));
    TRY(admit(ctx, controls, count, (uint32_t)C, (uint32_t)P, (uint32_t)M, plan, false,
              {{&ctx->probe, scene.size() * 8}, {&ctx->index, index.size() * 4},
               {&ctx->out, out_size * 8}},
              streamed));
    residency = streamed ? STREAMED : RESIDENT;
  }
  TRY(ensure(ctx->probe, scene.size() * 8));
  TRY(ensure(ctx->index, index.size() * 4));
  TRY(ensure(ctx->out, out_size * 8));
  hipStream_t st = ctx->stream;
  HIP_TRY(hipMemcpyAsync(ctx->probe.ptr, scene.data(), scene.size() * 8, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(ctx->index.ptr, index.data(), index.size() * 4, hipMemcpyHostToDevice, st));
  TRY(solve_device(ctx, controls, count, (const double*)ctx->probe.ptr,
                   (const uint32_t*)ctx->index.ptr, (double*)ctx->out.ptr, rep, finish_ms, st,
                   residency));
  std::vector<double> all(out_size);
  HIP_TRY(hipMemcpyAsync(all.data(), ctx->out.ptr, out_size * 8, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  c0 = p0 = 0;
  for (uint32_t b = 0; b < count; b++) {
    const size_t Cb = controls[b].cameras, Pb = controls[b].points;
    double* o = outs[b];
    std::memcpy(o, &all[12 * c0], 12 * Cb * 8);
    if (perms[b].empty()) {
      std::memcpy(o + 12 * Cb, &all[12 * C + 3 * p0], 3 * Pb * 8);
    } else {
      for (size_t j = 0; j < Pb; j++)
        for (int d = 0; d < 3; d++) o[12 * Cb + 3 * (size_t)perms[b][j] + d] = all[12 * C + 3 * (p0 + j) + d];
    }
    std::memcpy(o + 12 * Cb + 3 * Pb, &all[12 * C + 3 * P + b * words], words * 8);
    c0 += Cb;
    p0 += Pb;
  }
  if (rep) {
    rep->solve_ms = now_ms() - t0;
    rep->device_bytes = held_bytes(ctx);
  }
  return TCV_BA_OK;
}

static int body_ba_schur_probe(tcv_ba_context* ctx, const tcv_ba_control* c, const double* scene,
                               const uint32_t* index, double lambda, double log_focal,
                               const double* p, double* rhs, double* sp, double* z, int32_t* status,
                               uint32_t repeat, tcv_ba_report* rep) {
  if (!ctx || !c || !scene || !index || !p || !rhs || !sp || !z || !status)
    return TCV_BA_ERR_ARGUMENT;
  if (rep) std::memset(rep, 0, sizeof *rep);
  TRY(check_control(c));
  HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const size_t ss = tcv_ba_scene_size(c), is = tcv_ba_index_size(c);
  const uint32_t m = c->cameras * 6 + (c->shared_focal ? 1 : 0);
  TRY(ensure(ctx->probe, ss * 8));
  TRY(ensure(ctx->index, is * 4));
  TRY(ensure(ctx->out, (size_t)m * 4 * 8));
  HIP_TRY(hipMemcpyAsync(ctx->probe.ptr, scene, ss * 8, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(ctx->index.ptr, index, is * 4, hipMemcpyHostToDevice, st));
  Problem q;
  TRY(setup(ctx, c, 1, (const double*)ctx->probe.ptr, (const uint32_t*)ctx->index.ptr, st, q, rep));
  const Ops ops = select_ops(c);
  // state: lambda, LM active, the log focal scale, PCG active (so that the sweeps run)
  std::vector<double> init(S_WORDS, 0.0);
  init[S_LAMBDA] = lambda;
  init[S_ACTIVE] = 1;
  init[S_LOG_FOCAL] = log_focal;
  init[S_PCG_ACTIVE] = 1;
  HIP_TRY(hipMemcpyAsync(q.state, init.data(), S_WORDS * 8, hipMemcpyHostToDevice, st));
  double* dv = (double*)ctx->out.ptr;
  double* dsp = dv + m;
  double* dz = dv + 2 * (size_t)m;
  double* drhs = dv + 3 * (size_t)m;
  HIP_TRY(hipMemcpyAsync(dv, p, (size_t)m * 8, hipMemcpyHostToDevice, st));
  ops.point_pass(q, st);
  ops.camera_pass(q, st);
  ops.init_cg(q, st);  // the focal scalars of the assembled system
  // the probe vector is the PCG direction p; PCG active so that the sweeps run
  HIP_TRY(hipMemcpyAsync(q.vec + 2 * (size_t)m, p, (size_t)m * 8, hipMemcpyHostToDevice, st));
  const double one = 1;
  HIP_TRY(hipMemcpyAsync(q.state + S_PCG_ACTIVE, &one, 8, hipMemcpyHostToDevice, st));
  HIP_TRY(hipStreamSynchronize(st));
  const bool timing = (c->flags & TCV_BA_FLAG_TIMING) != 0 && rep;
  if (repeat < 1) repeat = 1;
  TRY(ensure_events(ctx, 3));
  for (uint32_t r = 0; r < repeat; r++) {
    if (timing) HIP_TRY(hipEventRecord(ctx->events[0], st));
    point_sweeps(q, ops, st);
    if (timing) HIP_TRY(hipEventRecord(ctx->events[1], st));
    camera_sweeps(q, ops, st);
    if (timing) HIP_TRY(hipEventRecord(ctx->events[2], st));
    if (timing) {
      HIP_TRY(hipEventSynchronize(ctx->events[2]));
      float a = 0, b = 0;
      HIP_TRY(hipEventElapsedTime(&a, ctx->events[0], ctx->events[1]));
      HIP_TRY(hipEventElapsedTime(&b, ctx->events[1], ctx->events[2]));
      rep->schur_point_ms += a;
      rep->schur_camera_ms += b;
      rep->schur_ms += a + b;
    }
  }
  if (rep) rep->matvecs = repeat;
  if (rep && timing) rep->timed_matvecs = repeat;
  ops.probe(q, dv, dsp, dz, drhs, st);
  HIP_TRY(hipGetLastError());
  // validity: every point tile and camera block
  std::vector<uint32_t> fail(q.n_tiles);
  std::vector<double> inv((size_t)q.C * 36);
  HIP_TRY(hipMemcpyAsync(fail.data(), q.tile_fail, (size_t)q.n_tiles * 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(inv.data(), q.camInv, (size_t)q.C * 36 * 8, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(sp, dsp, (size_t)m * 8, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(z, dz, (size_t)m * 8, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(rhs, drhs, (size_t)m * 8, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(ctx->h_state, q.state, S_WORDS * 8, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  int32_t bad = 0;
  for (uint32_t t = 0; t < q.n_tiles; t++) bad |= (fail[t] & 1u) != 0;
  for (uint32_t k = 0; k < q.C; k++) bad |= !(inv[(size_t)k * 36] > 0);
  if (c->shared_focal && !(ctx->h_state[S_FOCAL_INV] > 0)) bad = 1;
  *status = bad;
  if (rep) rep->device_bytes = held_bytes(ctx);
  return TCV_BA_OK;
}

static int body_ba_polish(tcv_ba_context* ctx, const tcv_ba_control* c, const double* scene,
                          const uint32_t* index, double log_focal, uint32_t lm_iterations,
                          uint32_t newton_steps, uint32_t max_cg, double tolerance, double* out,
                          tcv_ba_report* rep) {
  if (!ctx || !c || !scene || !index || !out || lm_iterations > 1000 || newton_steps > 100 ||
      max_cg < 1 || max_cg > 100000)
    return TCV_BA_ERR_ARGUMENT;
  if (rep) std::memset(rep, 0, sizeof *rep);
  const double t0 = now_ms();
  hipStream_t st = ctx->stream;
  tcv_ba_control fc;
  Problem q;
  Ops ops;
  TRY(prepare_exact(ctx, c, scene, index, log_focal, tolerance, st, fc, q, ops, rep));
  q.cg_budget = max_cg;
  q.cg_max = max_cg;
  const double cost0 = word(ctx, 0, S_COST0), sse0 = word(ctx, 0, S_SSE0);
  const uint32_t update_blocks = blocks_for(q.C > q.P ? q.C : q.P, 256);
  const uint32_t trial_blocks = blocks_for((size_t)q.C * 12 + (size_t)q.P * 3, 256);
  auto candidate = [&]() {
    hipLaunchKernelGGL(k_update_cameras, dim3(blocks_for(q.C, 64)), dim3(64), 0, st, q);
    update_points(q, ops, st);
    hipLaunchKernelGGL(k_gauge, dim3(1), dim3(256), 0, st, q);
    hipLaunchKernelGGL(k_normalize, dim3(update_blocks), dim3(256), 0, st, q);
  };
  // 1. Levenberg-Marquardt on the exact Hessian with the forward's damping and accept/reject rule
  //    (acceptBundleCandidate, bundle.shade.ts:1071-1096) and PCG tolerance, going on while the
  //    cost still decreases (until the damping passes 1e12 or the iterations end)
  Problem q1 = q;
  q1.stop_change = 1e-15;
  q1.tol = c->relative_tolerance;
  for (uint32_t it = 0; it < lm_iterations && word(ctx, 0, S_ACTIVE) != 0; it++) {
    ops.exact_passes(q1, st);
    ops.init_cg(q1, st);
    HIP_TRY(hipGetLastError());
    TRY(read_state(ctx, q1, st, rep));
    if (word(ctx, 0, S_PCG_ACTIVE) != 0) {
      TRY(solve_linear(ctx, &fc, q1, ops, st, rep, 32, max_cg, q1.tol));
      if (any_word(ctx, q1, S_PCG_ACTIVE))
        hipLaunchKernelGGL(k_stop_pcg, dim3(1), dim3(64), 0, st, q1);
    }
    hipLaunchKernelGGL(k_update_cameras, dim3(blocks_for(q.C, 64)), dim3(64), 0, st, q1);
    update_points(q1, ops, st);
    hipLaunchKernelGGL(k_gauge, dim3(1), dim3(256), 0, st, q1);
    hipLaunchKernelGGL(k_normalize, dim3(update_blocks), dim3(256), 0, st, q1);
    hipLaunchKernelGGL(k_cost<true>, dim3(q.n_cost_blocks), dim3(COST_OBS), 0, st, q1);
    hipLaunchKernelGGL(k_finish_cost<true>, dim3(1), dim3(COST_OBS), 0, st, q1);
    hipLaunchKernelGGL(k_commit, dim3(trial_blocks), dim3(256), 0, st, q1);
    HIP_TRY(hipGetLastError());
    TRY(read_state(ctx, q, st, rep));
    if (rep) {
      rep->lm_iterations++;
      rep->cg_iterations += (uint64_t)word(ctx, 0, S_STEPS);
    }
  }
  // 2. undamped Newton steps without a cost test (near the solution the cost change falls below
  //    the rounding of the cost itself)
  double taken_newton = 0;
  for (uint32_t it = 0; it < newton_steps; it++) {
    bool converged = false;
    TRY(exact_solve(ctx, &fc, q, ops, st, max_cg, rep, converged));
    if (!converged) break;
    candidate();
    hipLaunchKernelGGL(k_newton_accept, dim3(1), dim3(64), 0, st, q);
    hipLaunchKernelGGL(k_commit, dim3(trial_blocks), dim3(256), 0, st, q);
    HIP_TRY(hipGetLastError());
    TRY(read_state(ctx, q, st, rep));
    if (word(ctx, 0, S_COMMIT) == 0) break;
    taken_newton += 1;
  }
  if (rep) {
    rep->refinement_steps = taken_newton;
    rep->direct_solves = (uint64_t)word(ctx, 0, S_DIRECT_SOLVES);
    rep->held_points = (uint64_t)word(ctx, 0, S_HELD);
  }
  // the final cost; the initial words keep the polish input's
  hipLaunchKernelGGL(k_cost<false>, dim3(q.n_cost_blocks), dim3(COST_OBS), 0, st, q);
  hipLaunchKernelGGL(k_finish_cost<false>, dim3(1), dim3(COST_OBS), 0, st, q);
  const double initial[2] = {cost0, sse0};
  HIP_TRY(hipMemcpyAsync(q.state + S_COST0, initial, 16, hipMemcpyHostToDevice, st));
  const size_t os = tcv_ba_output_size(c);
  TRY(ensure(ctx->out, os * 8));
  hipLaunchKernelGGL(k_pack, dim3(blocks_for(os, 256)), dim3(256), 0, st, q, (double*)ctx->out.ptr);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out, ctx->out.ptr, os * 8, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (rep) {
    rep->solve_ms = now_ms() - t0;
    rep->device_bytes = held_bytes(ctx);
  }
  return TCV_BA_OK;
}

static int body_ba_backward(tcv_ba_context* ctx, const tcv_ba_control* c, const double* scene,
                            const uint32_t* index, double log_focal, const double* cotangent,
                            uint32_t max_cg, double tolerance, uint32_t mode,
                            double* pixel_cotangent, double* focal_cotangent, double* adjoint,
                            tcv_ba_report* rep) {
  if (!ctx || !c || !scene || !index || !cotangent || !pixel_cotangent || max_cg < 1 ||
      max_cg > 100000)
    return TCV_BA_ERR_ARGUMENT;
  if (mode != TCV_BA_ADJOINT_EXACT && mode != TCV_BA_ADJOINT_GAUSS_NEWTON) {
    set_error("unknown adjoint mode");
    return TCV_BA_ERR_ARGUMENT;
  }
  if (rep) std::memset(rep, 0, sizeof *rep);
  const double t0 = now_ms();
  hipStream_t st = ctx->stream;
  tcv_ba_control fc;
  Problem q;
  Ops ops;
  TRY(prepare_exact(ctx, c, scene, index, log_focal, tolerance, st, fc, q, ops, rep));
  q.cg_budget = max_cg;
  q.cg_max = max_cg;
  q.gauss_newton = mode == TCV_BA_ADJOINT_GAUSS_NEWTON;
  const size_t C = q.C, P = q.P, M = q.M, m = q.m;
  for (size_t k = 0; k < 12 * C + 3 * P + q.shared; k++)
    if (!std::isfinite(cotangent[k])) {
      set_error("cotangent has a non-finite entry");
      return TCV_BA_ERR_ARGUMENT;
    }
  const std::vector<double> rhs = adjoint_rhs(c, scene, index, cotangent);
  TRY(ensure(ctx->rhs, rhs.size() * 8));
  // the pixel cotangent (2M), each observation's focal share (M), the per-camera sums (C)
  TRY(ensure(ctx->grad, (3 * M + C) * 8));
  HIP_TRY(hipMemcpyAsync(ctx->rhs.ptr, rhs.data(), rhs.size() * 8, hipMemcpyHostToDevice, st));
  q.rhs_in = (const double*)ctx->rhs.ptr;
  const double ts = now_ms();
  bool converged = false;
  TRY(exact_solve(ctx, &fc, q, ops, st, max_cg, rep, converged));
  if (!converged) {
    char text[256];
    if (word(ctx, 0, S_PCG_FAILED) != 0)
      std::snprintf(text, sizeof text,
                    "the adjoint system is not positive definite at this scene (not a minimum?): "
                    "%.0f point tiles, %.0f camera blocks, focal %.0f failed; %.0f PCG steps",
                    word(ctx, 0, S_FAIL_TILES), word(ctx, 0, S_FAIL_CAMERAS),
                    word(ctx, 0, S_FAIL_FOCAL), word(ctx, 0, S_STEPS));
    else
      std::snprintf(text, sizeof text,
                    "the adjoint PCG did not reach its tolerance within %u iterations", max_cg);
    set_error(text);
    return TCV_BA_ERR_SOLVE;
  }
  ops.solve_points(q, st);
  double* grad = (double*)ctx->grad.ptr;
  ops.pixel_cotangent(q, grad, focal_cotangent ? grad + 2 * M : nullptr, st);
  if (focal_cotangent && C)
    hipLaunchKernelGGL(k_focal_cotangent, dim3(C), dim3(256), 0, st, q, grad + 2 * M,
                       grad + 3 * M);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(pixel_cotangent, grad, 2 * M * 8, hipMemcpyDeviceToHost, st));
  if (focal_cotangent)
    HIP_TRY(hipMemcpyAsync(focal_cotangent, grad + 3 * M, C * 8, hipMemcpyDeviceToHost, st));
  if (adjoint) {
    HIP_TRY(hipMemcpyAsync(adjoint, q.vec, 6 * C * 8, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(adjoint + 6 * C, q.vec + 5 * m, 3 * P * 8, hipMemcpyDeviceToHost, st));
    if (q.shared)
      HIP_TRY(hipMemcpyAsync(adjoint + 6 * C + 3 * P, q.vec + 6 * C, 8, hipMemcpyDeviceToHost, st));
  }
  HIP_TRY(hipStreamSynchronize(st));
  if (rep) {
    rep->solve_ms = now_ms() - t0;
    rep->schur_ms = now_ms() - ts;
    rep->device_bytes = held_bytes(ctx);
    rep->direct_solves = (uint64_t)word(ctx, 0, S_DIRECT_SOLVES);
    rep->held_points = (uint64_t)word(ctx, 0, S_HELD);
  }
  return TCV_BA_OK;
}

static int body_obs_evaluate(tcv_ba_context* ctx, const tcv_obs_control* c, const double* cameras,
                             const double* points, const double* pixels, const uint32_t* indices,
                             const uint32_t* offsets, uint32_t* flags, double* summary) {
  if (!ctx || !c || !flags || !summary) return TCV_BA_ERR_ARGUMENT;
  const uint32_t M = c->observations, T = c->tracks;
  const bool guarded = c->max_relative_depth_std_dev > 0;
  // the reference's request checks (observations.ts:141-150, multiview.ts:26-36)
  if (!(c->max_error_px >= 1e-12 && c->max_error_px <= 1e6) ||
      !(c->pixel_noise_std_dev_px > 0 && c->pixel_noise_std_dev_px < INFINITY) ||
      !(c->max_relative_depth_std_dev >= 0 && c->max_relative_depth_std_dev < INFINITY)) {
    set_error("Invalid GPU observation threshold");
    return TCV_BA_ERR_ARGUMENT;
  }
  if ((M && (!cameras || !points || !pixels || !indices || !c->views || !c->points)) ||
      (guarded && T && !offsets) || (!guarded && T) || (guarded && T != c->points)) {
    set_error("observation arrays do not match the control");
    return TCV_BA_ERR_ARGUMENT;
  }
  // every index in range, offsets ascending and covering the observations: checked on the host,
  // as the arrays are host arrays here
  for (uint32_t i = 0; i < M; i++)
    if (indices[2 * (size_t)i] >= c->views || indices[2 * (size_t)i + 1] >= c->points) {
      set_error("observation index outside the camera or point records");
      return TCV_BA_ERR_INDEX;
    }
  if (guarded && T) {
    if (offsets[0] != 0 || offsets[T] != M) {
      set_error("track offsets do not cover the observations");
      return TCV_BA_ERR_INDEX;
    }
    for (uint32_t h = 0; h < T; h++)
      if (offsets[h + 1] < offsets[h]) {
        set_error("track offsets decrease");
        return TCV_BA_ERR_INDEX;
      }
  }
  HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const uint32_t blocks = (M ? M : 1) / 64 + ((M ? M : 1) % 64 != 0);
  const size_t nflags = (size_t)M + (guarded ? T : 0);
  TRY(ensure(ctx->obs_cameras, (size_t)c->views * 18 * 8 + 8));
  TRY(ensure(ctx->obs_points, (size_t)c->points * 3 * 8 + 8));
  TRY(ensure(ctx->obs_pixels, (size_t)M * 2 * 8 + 8));
  TRY(ensure(ctx->obs_indices, (size_t)M * 2 * 4 + 8));
  TRY(ensure(ctx->obs_offsets, (size_t)(T + 1) * 4 + 8));
  TRY(ensure(ctx->obs_flags, nflags * 4 + 8));
  TRY(ensure(ctx->obs_errors, (size_t)M * 8 + 8));
  TRY(ensure(ctx->obs_partials, (size_t)blocks * 2 * 8));
  TRY(ensure(ctx->obs_summary, 2 * 8));
  if (M) {
    HIP_TRY(hipMemcpyAsync(ctx->obs_cameras.ptr, cameras, (size_t)c->views * 18 * 8,
                           hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(ctx->obs_points.ptr, points, (size_t)c->points * 3 * 8,
                           hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(ctx->obs_pixels.ptr, pixels, (size_t)M * 2 * 8, hipMemcpyHostToDevice,
                           st));
    HIP_TRY(hipMemcpyAsync(ctx->obs_indices.ptr, indices, (size_t)M * 2 * 4,
                           hipMemcpyHostToDevice, st));
  }
  if (guarded && T)
    HIP_TRY(hipMemcpyAsync(ctx->obs_offsets.ptr, offsets, (size_t)(T + 1) * 4,
                           hipMemcpyHostToDevice, st));
  ObsProblem o;
  o.cameras = (const double*)ctx->obs_cameras.ptr;
  o.points = (const double*)ctx->obs_points.ptr;
  o.pixels = (const double*)ctx->obs_pixels.ptr;
  o.indices = (const uint32_t*)ctx->obs_indices.ptr;
  o.offsets = (const uint32_t*)ctx->obs_offsets.ptr;
  o.flags = (uint32_t*)ctx->obs_flags.ptr;
  o.errors = (double*)ctx->obs_errors.ptr;
  o.partials = (double*)ctx->obs_partials.ptr;
  o.summary = (double*)ctx->obs_summary.ptr;
  o.count = M;
  o.tracks = guarded ? T : 0;
  o.blocks = blocks;
  o.max_error = c->max_error_px;
  o.max_relative_depth_std = c->max_relative_depth_std_dev;
  o.pixel_noise_std = c->pixel_noise_std_dev_px;
  if (M) hipLaunchKernelGGL(k_obs_project, dim3(blocks_for(M, 256)), dim3(256), 0, st, o);
  if (o.tracks) hipLaunchKernelGGL(k_obs_depth, dim3(blocks_for(o.tracks, 256)), dim3(256), 0, st, o);
  if (M) hipLaunchKernelGGL(k_obs_partials, dim3(blocks_for(blocks, 256)), dim3(256), 0, st, o);
  else HIP_TRY(hipMemsetAsync(o.partials, 0, (size_t)blocks * 2 * 8, st));
  hipLaunchKernelGGL(k_obs_summary, dim3(1), dim3(SUMMARY_THREADS), 0, st, o);
  HIP_TRY(hipGetLastError());
  if (nflags)
    HIP_TRY(hipMemcpyAsync(flags, o.flags, nflags * 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(summary, o.summary, 2 * 8, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  return TCV_BA_OK;
}

// The scratch of tcv_obs_evaluate_noalloc: the per-observation errors, the per-64-observation
// partials and the index-check word of at most kNoallocObservations observations, in one device
// array of the library (one copy per GPU, held until the process ends; 16.5 MiB). Host code uses
// it only through the address hipGetSymbolAddress returns for the current device.
constexpr uint32_t kNoallocObservations = TCV_OBS_NOALLOC_MAX_OBSERVATIONS;
constexpr uint32_t kNoallocBlocks = kNoallocObservations / 64;
constexpr size_t kNoallocWords = (size_t)kNoallocObservations + 2 * (size_t)kNoallocBlocks + 32;
static __device__ __align__(256) double scratch[kNoallocWords];

namespace {
std::mutex g_scratch_lock;
double* g_scratch_ptr[64] = {};  // the address of `scratch` on each device, once looked up

// The address of `scratch` on the current device (looked up on the entry's first call there).
int scratch_address(double** out) {
  int device = 0;
  HIP_TRY(hipGetDevice(&device));
  if (device < 0 || device >= 64) {
    set_error("device index outside the scratch table");
    return TCV_BA_ERR_NO_DEVICE;
  }
  std::lock_guard<std::mutex> guard(g_scratch_lock);
  if (!g_scratch_ptr[device]) {
    void* ptr = nullptr;
    HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = (double*)ptr;
  }
  *out = g_scratch_ptr[device];
  return TCV_BA_OK;
}
}  // namespace

static int body_obs_evaluate_noalloc(const tcv_obs_control* c, const double* cameras,
                                     const double* points, const double* pixels,
                                     const uint32_t* indices, const uint32_t* offsets,
                                     uint32_t* flags, double* summary, void* stream) {
  if (!c || !flags || !summary) return TCV_BA_ERR_ARGUMENT;
  const uint32_t M = c->observations, T = c->tracks;
  const bool guarded = c->max_relative_depth_std_dev > 0;
  if (!(c->max_error_px >= 1e-12 && c->max_error_px <= 1e6) ||
      !(c->pixel_noise_std_dev_px > 0 && c->pixel_noise_std_dev_px < INFINITY) ||
      !(c->max_relative_depth_std_dev >= 0 && c->max_relative_depth_std_dev < INFINITY)) {
    set_error("Invalid GPU observation threshold");
    return TCV_BA_ERR_ARGUMENT;
  }
  if ((M && (!cameras || !points || !pixels || !indices || !c->views || !c->points)) ||
      (guarded && T && !offsets) || (!guarded && T) || (guarded && T != c->points)) {
    set_error("observation arrays do not match the control");
    return TCV_BA_ERR_ARGUMENT;
  }
  if (M > kNoallocObservations) {
    set_error("tcv_obs_evaluate_noalloc takes at most " + std::to_string(kNoallocObservations) +
              " observations (its fixed scratch); " + std::to_string(M) + " given");
    return TCV_BA_ERR_ARGUMENT;
  }
  double* base = nullptr;
  TRY(scratch_address(&base));
  hipStream_t st = (hipStream_t)stream;
  const uint32_t blocks = (M ? M : 1) / 64 + ((M ? M : 1) % 64 != 0);
  ObsProblem o;
  o.cameras = cameras;
  o.points = points;
  o.pixels = pixels;
  o.indices = indices;
  o.offsets = offsets;
  o.flags = flags;
  o.errors = base;
  o.partials = base + kNoallocObservations;
  o.summary = summary;
  o.count = M;
  o.tracks = guarded ? T : 0;
  o.blocks = blocks;
  o.max_error = c->max_error_px;
  o.max_relative_depth_std = c->max_relative_depth_std_dev;
  o.pixel_noise_std = c->pixel_noise_std_dev_px;
  // the device arrays' indices are checked on the device before anything reads through them
  uint32_t* bad = (uint32_t*)(base + kNoallocObservations + 2 * (size_t)kNoallocBlocks);
  hipLaunchKernelGGL(k_obs_zero, dim3(1), dim3(64), 0, st, bad, 1u);
  const uint32_t checked = M > o.tracks ? M : o.tracks;
  if (checked)
    hipLaunchKernelGGL(k_obs_check, dim3(blocks_for(checked, 256)), dim3(256), 0, st, o, c->views,
                       c->points, bad);
  HIP_TRY(hipGetLastError());
  uint32_t word = 0;
  HIP_TRY(hipMemcpyAsync(&word, bad, 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (word) {
    set_error(word & 1 ? "observation index outside the camera or point records"
                       : "track offsets do not ascend from 0 to the observation count");
    return TCV_BA_ERR_INDEX;
  }
  if (M) hipLaunchKernelGGL(k_obs_project, dim3(blocks_for(M, 256)), dim3(256), 0, st, o);
  if (o.tracks) hipLaunchKernelGGL(k_obs_depth, dim3(blocks_for(o.tracks, 256)), dim3(256), 0, st, o);
  if (M) hipLaunchKernelGGL(k_obs_partials, dim3(blocks_for(blocks, 256)), dim3(256), 0, st, o);
  else hipLaunchKernelGGL(k_obs_zero_f64, dim3(1), dim3(64), 0, st, o.partials, 2u * blocks);
  hipLaunchKernelGGL(k_obs_summary, dim3(1), dim3(SUMMARY_THREADS), 0, st, o);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(st));
  return TCV_BA_OK;
}

static int body_ba_jacobian(tcv_ba_context* ctx, const tcv_ba_control* c, const double* scene,
                            const uint32_t* index, double log_focal, double* out,
                            tcv_ba_report* rep) {
  if (!ctx || !c || !scene || !index || !out) return TCV_BA_ERR_ARGUMENT;
  if (rep) std::memset(rep, 0, sizeof *rep);
  hipStream_t st = ctx->stream;
  tcv_ba_control fc;
  Problem q;
  Ops ops;
  TRY(prepare_exact(ctx, c, scene, index, log_focal, c->relative_tolerance, st, fc, q, ops, rep));
  TRY(ensure(ctx->grad, (size_t)q.M * 23 * 8));
  ops.jacobian(q, (double*)ctx->grad.ptr, st);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out, ctx->grad.ptr, (size_t)q.M * 23 * 8, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  return TCV_BA_OK;
}

}  // extern "C"

namespace {

// An entry point's call with the context's lock held for its whole duration. A call that fails
// past its argument checks drops every device buffer of the context, as the reference drops a
// failed job's buffers (the pending work is waited for first); the error text is the call's.
template <class F>
int locked(tcv_ba_context* ctx, F&& body) {
  if (!ctx) return TCV_BA_ERR_ARGUMENT;
  std::lock_guard<std::mutex> guard(ctx->lock);
  const int status = body();
  if (status != TCV_BA_OK && status != TCV_BA_ERR_ARGUMENT) {
    const std::string text = g_last_error;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    (void)release_buffers(ctx);
    g_last_error = text;
  }
  return status;
}

}  // namespace

extern "C" {

int tcv_ba_solve_device(tcv_ba_context* ctx, const tcv_ba_control* c, const double* scene,
                        const uint32_t* index, double* out, tcv_ba_report* rep, void* stream) {
  return locked(ctx, [&] { return body_ba_solve_device(ctx, c, scene, index, out, rep, stream); });
}

int tcv_ba_solve(tcv_ba_context* ctx, const tcv_ba_control* c, const double* scene,
                 const uint32_t* index, double* out, tcv_ba_report* rep) {
  return locked(ctx, [&] { return body_ba_solve(ctx, c, scene, index, out, rep); });
}

int tcv_ba_solve_batch(tcv_ba_context* ctx, uint32_t count, const tcv_ba_control* controls,
                       const double* const* scenes, const uint32_t* const* indices,
                       double* const* outs, tcv_ba_report* report, double* finish_ms) {
  return locked(ctx, [&] {
    return body_ba_solve_batch(ctx, count, controls, scenes, indices, outs, report, finish_ms);
  });
}

int tcv_ba_schur_probe(tcv_ba_context* ctx, const tcv_ba_control* c, const double* scene,
                       const uint32_t* index, double lambda, double log_focal, const double* p,
                       double* rhs, double* sp, double* z, int32_t* status, uint32_t repeat,
                       tcv_ba_report* rep) {
  return locked(ctx, [&] {
    return body_ba_schur_probe(ctx, c, scene, index, lambda, log_focal, p, rhs, sp, z, status,
                               repeat, rep);
  });
}

int tcv_ba_polish(tcv_ba_context* ctx, const tcv_ba_control* c, const double* scene,
                  const uint32_t* index, double log_focal, uint32_t lm_iterations,
                  uint32_t newton_steps, uint32_t max_cg, double tolerance, double* out,
                  tcv_ba_report* rep) {
  return locked(ctx, [&] {
    return body_ba_polish(ctx, c, scene, index, log_focal, lm_iterations, newton_steps, max_cg,
                          tolerance, out, rep);
  });
}

int tcv_ba_backward(tcv_ba_context* ctx, const tcv_ba_control* c, const double* scene,
                    const uint32_t* index, double log_focal, const double* cotangent,
                    uint32_t max_cg, double tolerance, uint32_t mode, double* pixel_cotangent,
                    double* focal_cotangent, double* adjoint, tcv_ba_report* rep) {
  return locked(ctx, [&] {
    return body_ba_backward(ctx, c, scene, index, log_focal, cotangent, max_cg, tolerance, mode,
                            pixel_cotangent, focal_cotangent, adjoint, rep);
  });
}

int tcv_obs_evaluate(tcv_ba_context* ctx, const tcv_obs_control* c, const double* cameras,
                     const double* points, const double* pixels, const uint32_t* indices,
                     const uint32_t* offsets, uint32_t* flags, double* summary) {
  return locked(ctx, [&] {
    return body_obs_evaluate(ctx, c, cameras, points, pixels, indices, offsets, flags, summary);
  });
}

int tcv_ba_jacobian(tcv_ba_context* ctx, const tcv_ba_control* c, const double* scene,
                    const uint32_t* index, double log_focal, double* out, tcv_ba_report* rep) {
  return locked(ctx, [&] { return body_ba_jacobian(ctx, c, scene, index, log_focal, out, rep); });
}

int tcv_obs_evaluate_noalloc(const tcv_obs_control* c, const double* cameras,
                             const double* points, const double* pixels, const uint32_t* indices,
                             const uint32_t* offsets, uint32_t* flags, double* summary,
                             void* stream) {
  return body_obs_evaluate_noalloc(c, cameras, points, pixels, indices, offsets, flags, summary,
                                   stream);
}

int tcv_ba_set_working_bytes(tcv_ba_context* ctx, uint64_t bytes) {
  if (!ctx) return TCV_BA_ERR_ARGUMENT;
  std::lock_guard<std::mutex> guard(ctx->lock);
  ctx->budget = bytes;
  return TCV_BA_OK;
}

int tcv_ba_fail_reads(tcv_ba_context* ctx, uint32_t n) {
  if (!ctx) return TCV_BA_ERR_ARGUMENT;
  std::lock_guard<std::mutex> guard(ctx->lock);
  ctx->fail_read_in = n;
  return TCV_BA_OK;
}

}  // extern "C"
