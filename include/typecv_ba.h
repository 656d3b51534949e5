/*
 * typecv_ba.h — C ABI of the MI355X bundle-adjustment engine behind `bundleAdjustGpu`.
 *
 * The engine runs the whole Levenberg–Marquardt loop of the reference operator on one gfx950 GPU:
 * linearisation, point elimination, matrix-free Schur products, block-Jacobi PCG, back-substitution,
 * gauge handling and the LM accept/reject rule. The host (a Node-API addon, the Python mirror in
 * `typecv_amd/bundle.py`, or any FFI) keeps the reference's validation, packing and unpacking and
 * hands over the two packed arrays of `packBundleScene`:
 *
 *   scene  f64 [poses 12C][points 3P][models 6C: fx fy cx cy k1 k2][pixels 2M]
 *   index  u32 [obsCamera M][obsPoint M][pointStart P+1][cameraStart C+1]
 *              [cameraObservations M][cameraActive C]                       (3M + P + 2C + 2 words)
 *
 * and receives the array of `packBundleResult`:
 *
 *   out    f64 [poses 12C][points 3P][state 26] (+ [log focal scale] on the shared-focal route)
 *
 * State words (reference `BUNDLE_STATE`, src/sfm/bundle/gpu/bundle-shared.ts:38-54):
 *   0 lambda, 1 LM active, 2 baseline, 3..5 gauge origin, 6 cost, 7 squared error,
 *   8 initial cost, 9 initial squared error, 10 accepted, 11 rejected, 12 commit, 13 gauge scale,
 *   14 input error (1 baseline, 2 cheirality), 15 trial valid, 16 CG steps, 17 failed solves,
 *   18 rho, 19 PCG active, 20 PCG failed, 21 |r0|^2, 22 |r|^2, 23 steps, 24 initial RMSE,
 *   25 final RMSE (24/25 hold the RMSEs in `out`).
 *
 * A failed linear solve is not an error: it shows in words 11 and 17 with the accepted geometry
 * unchanged. Input errors detected on the device (words 14) are reported in `out` as the reference
 * reports them; the host turns them into its errors (reference `checkBundleOutput`,
 * src/sfm/bundle/gpu/bundle-host.ts:278-283).
 *
 * Thread safety: every entry point that takes a context holds that context's lock for the whole
 * call, so calls on one context from several threads run one after another (the reference's
 * per-device FIFO of whole jobs, src/device/gpu/device.ts:592-598); distinct contexts run
 * concurrently.
 */
#ifndef TYPECV_BA_H
#define TYPECV_BA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TCV_BA_STATE_WORDS 26

/* Arithmetic forms. */
enum tcv_ba_precision {
  TCV_BA_FP64 = 0,  /* plain IEEE binary64 everywhere */
  TCV_BA_MIXED = 1  /* FP32 cross blocks and Schur products with FP64 accumulation, FP64 V^-1,
                       preconditioner, dots and LM decisions, outer iterative refinement */
};

/* Status codes of every entry point. */
enum tcv_ba_status {
  TCV_BA_OK = 0,
  TCV_BA_ERR_ARGUMENT = 1, /* sizes/options outside the reference's admitted ranges */
  TCV_BA_ERR_INDEX = 2,    /* the index array is not a packBundleScene layout of these sizes */
  TCV_BA_ERR_DEVICE = 3,   /* a HIP runtime call failed */
  TCV_BA_ERR_MEMORY = 4,   /* device allocation failed (the reference's budget refusal) */
  TCV_BA_ERR_NO_DEVICE = 5, /* no gfx950 device is visible */
  TCV_BA_ERR_SOLVE = 6      /* the adjoint system was not solved (not a minimum, or PCG budget) */
};

/*
 * The problem and options. Mirrors `BundleControl` (reference src/sfm/bundle/gpu/bundle.shade.ts:17-48
 * and joint-focal.shade.ts:14-42) for a resident run, plus the host-side loop options of
 * `bundleAdjustGpuFixed` (bundle.ts:89-132) and `bundleAdjustJointFocalGpu` (joint-focal.ts:78-130).
 */
typedef struct tcv_ba_control {
  uint32_t cameras;            /* C >= 2 */
  uint32_t points;             /* P: finite points, then directional points */
  uint32_t directional_points; /* trailing points that hold unit world bearings */
  uint32_t observations;       /* M */
  uint32_t anchor;             /* camera index of anchorViewId */
  uint32_t scale_view;         /* camera index of scaleViewId */
  uint32_t normalize_scale;    /* 1 when exactly one camera is fixed */
  uint32_t shared_focal;       /* 0: fixed intrinsics (bundle.shade.ts); 1: shared log focal */
  uint32_t iterations;         /* LM iterations, 0..100 */
  uint32_t max_cg;             /* PCG iterations per LM iteration, 1..256 (per bundle in a batch;
                                  each refinement restart has the same budget again) */
  uint32_t precision;          /* enum tcv_ba_precision */
  uint32_t flags;              /* 0; TCV_BA_FLAG_TIMING times the Schur sweeps with HIP events */
  double huber;                /* Huber threshold in pixels, [1e-12, 1e6] */
  double relative_tolerance;   /* PCG relative residual tolerance, [1e-12, 1e-2] */
  double stall_tolerance;      /* host LM stall rule; <= 1e-9 leaves only the device rule */
  double refinement_tolerance; /* mixed form: relative FP64 normal-equation residual that the
                                  outer refinement reaches (0: relative_tolerance). A step is
                                  accepted when its final residual meets the larger of the two
                                  tolerances (the one its solution was certified at) */
} tcv_ba_control;

#define TCV_BA_FLAG_TIMING 1u
/* PCG (and the probe) use the implicit FP64 Schur operator: every cross block is relinearised from
 * the scene in each sweep instead of streamed from storage. Same results to rounding. */
#define TCV_BA_FLAG_IMPLICIT 2u
/* The reference's opt-in `schurCache` (src/sfm/bundle/gpu/joint-focal.ts:102-108). It is honoured
 * only where the reference honours it: on the shared-focal route, for a bundle of at most 16
 * cameras with M (6C + 1) <= 262,144. Such a bundle takes the direct route below (the reference runs
 * PCG on the cached dense matrix instead; both are certified by the same matrix-free residual).
 * Everywhere else the flag is validated and has no effect, as on the reference's fixed route
 * (bundle.ts:825-826): the solve returns the matrix-free route's bits. */
#define TCV_BA_FLAG_SCHUR_CACHE 4u
/* The direct route, this engine's option `directSolve` (no counterpart in the reference): each
 * bundle of at most 2,048 cameras assembles its reduced camera system densely, factors it by
 * Cholesky (one workgroup per bundle up to 128 cameras, a sequence of launches above) and solves it
 * in place of the PCG loop; the matrix-free residual certifies the solution (PCG restarts from it
 * when it misses the tolerance), as joint-focal.shade.ts:1310-1383 certifies a cached solve.
 * Larger bundles, and bundles in which a camera observes one point twice, keep the PCG route. */
#define TCV_BA_FLAG_DIRECT 8u

/* What one solve did. Times are milliseconds of host wall clock unless named otherwise. */
typedef struct tcv_ba_report {
  double solve_ms;          /* device-resident solve: setup + LM loop, inputs already in HBM */
  double setup_ms;          /* index checks, tile tables, camera-major permutation */
  double schur_ms;          /* HIP-event time of the timed Schur sweeps (TCV_BA_FLAG_TIMING) */
  double schur_point_ms;    /* of which the point sweep */
  double schur_camera_ms;   /* of which the camera sweep */
  double refinement_steps;  /* outer refinement corrections taken (mixed form) */
  uint64_t lm_iterations;   /* LM iterations executed (accepted + rejected) */
  uint64_t cg_iterations;   /* PCG iterations that completed */
  uint64_t matvecs;         /* Schur products computed while PCG was active */
  uint64_t matvec_launches; /* Schur products launched (chunks end past convergence) */
  uint64_t flag_reads;      /* host reads of the activity words */
  uint64_t device_bytes;    /* device memory held by the context after the solve */
  uint64_t point_tiles;     /* wave tiles of whole points */
  uint64_t camera_tasks;    /* camera-list tasks */
  uint64_t direct_solves;   /* certified direct solves (TCV_BA_FLAG_DIRECT / SCHUR_CACHE) */
  uint64_t timed_matvecs;   /* Schur products timed into schur_ms (every 8th, TCV_BA_FLAG_TIMING) */
  uint64_t direct_corrections; /* refinement corrections of direct solutions from their factor
                                  (one per bundle whose certificate missed, at most two per solve) */
  uint64_t camera_sweep_cap;   /* camera-sweep workgroups per CU that the PCG launches used, as
                                  picked by the context's schedule tuner for this problem shape
                                  (0: as many as the sweep's registers allow) */
  uint64_t fused_product;      /* 1: the PCG launches used the fused FP64 Schur product (one read
                                  of the cross blocks, a row of products per observation), as
                                  picked by the same tuner; 0: the two plain sweeps */
  uint64_t streamed;           /* 1: the solve ran streamed under the context's working budget
                                  (no stored cross blocks: every sweep relinearises them) */
  uint64_t held_points;        /* tcv_ba_polish, tcv_ba_backward: points held fixed by the last
                                  exact-Hessian system, their blocks singular to rounding (tracks
                                  run off towards infinity); zero adjoint rows */
  uint64_t onepass_rows;       /* the one-pass Schur product's rows of camera partials (the sum of
                                  its workgroups' camera windows); 0: the two sweeps ran */
} tcv_ba_report;

typedef struct tcv_ba_context tcv_ba_context;

/* Size of the `out` array in doubles: 12C + 3P + 26 (+1 with shared focal). */
size_t tcv_ba_output_size(const tcv_ba_control* control);
/* Size of the scene array (12C + 3P + 6C + 2M) and of the index array (3M + P + 2C + 2). */
size_t tcv_ba_scene_size(const tcv_ba_control* control);
size_t tcv_ba_index_size(const tcv_ba_control* control);

/* Human-readable text of a status code. */
const char* tcv_ba_status_string(int status);
/* The last error text of this thread (HIP error names included), or "". */
const char* tcv_ba_last_error(void);

/*
 * Create / destroy a context on HIP device `device`. The context owns the device buffers, streams
 * and pinned flag words and keeps them across solves (grown as needed). Replaces the buffer pool of
 * the reference's `ShadeDevice` (src/device/gpu/device.ts:608-710).
 */
int tcv_ba_create(int device, tcv_ba_context** out);
void tcv_ba_destroy(tcv_ba_context* context);
/* Device bytes currently held by the context (the reference's `statistics.liveBytes`). */
uint64_t tcv_ba_device_bytes(const tcv_ba_context* context);
/* Release every device buffer of the context (the reference's `drop`). A call that fails past its
 * argument checks releases them too, as the reference drops a failed job's buffers. */
int tcv_ba_release(tcv_ba_context* context);

/*
 * The working budget of the solve entry points (tcv_ba_solve, tcv_ba_solve_device,
 * tcv_ba_solve_batch) in bytes of device memory, staging of host arrays included; 0 (the default)
 * sets none. The reference's `workingBytes` device option and residency plan
 * (src/device/gpu/device.ts:360, src/device/gpu/budget-plan.ts:114-188, bundle.ts:195-242): a solve
 * whose arrays all fit runs resident; otherwise it runs streamed when the arrays without the
 * stored cross blocks fit (every Schur sweep then relinearises each block in registers, with the
 * same results bit for bit); otherwise it is refused with TCV_BA_ERR_MEMORY and the reference's
 * "GPU BA needs N bytes ... more than the working budget of B bytes" text before anything is
 * allocated. A context whose held buffers would pass the budget releases them before allocating.
 */
int tcv_ba_set_working_bytes(tcv_ba_context* context, uint64_t bytes);

/*
 * Failure injection for tests of the error path: the n-th next read of the LM state words on this
 * context fails with TCV_BA_ERR_DEVICE ("injected read failure"), as the reference's tests replace
 * the device's read (tests/typeshade/cpu.test.mjs:508-531); 0 turns it off. No counterpart in the
 * reference's operator API.
 */
int tcv_ba_fail_reads(tcv_ba_context* context, uint32_t n);

/*
 * Solve from host arrays: upload, run the LM loop, read back `out`. The call that the Node-API
 * addon runs as async work for `bundleAdjustGpu` (reference src/sfm/bundle/gpu/bundle.ts:803-839:
 * everything inside `device.enqueue`, bundle.ts:178-799 and joint-focal.ts equivalents). Points
 * may be solved in another order than the given one: the order of the lowest camera of their
 * tracks, then their highest, taken when every 32 consecutive tiles of points then fit a window of 64 cameras (the
 * one-pass Schur product's condition) and the given order's do not, or write at least an eighth
 * more partial rows; not for a problem the direct route takes (TCV_BA_FLAG_DIRECT, at most 2,048
 * cameras). `out` holds the points back in the given order; the results differ from the given
 * order's only by the order of sums over points. tcv_ba_solve_batch and tcv_ba_solve_device do
 * the same per problem (the device entry on the device, not under a working budget).
 */
int tcv_ba_solve(tcv_ba_context* context, const tcv_ba_control* control, const double* scene,
                 const uint32_t* index, double* out, tcv_ba_report* report);

/*
 * Solve from device arrays already resident in HBM (`scene`, `index`, `out` are device pointers of
 * the context's device). `scene` and `index` are read, never written; a reordering of the points
 * (tcv_ba_solve) works on copies in the context's buffers. Work runs on `stream` (a hipStream_t,
 * or NULL for the context's own stream) and the call returns when `out` is complete.
 */
int tcv_ba_solve_device(tcv_ba_context* context, const tcv_ba_control* control,
                        const double* scene, const uint32_t* index, double* out,
                        tcv_ba_report* report, void* stream);

/*
 * Solve `count` independent problems together (host arrays, one set per problem). The problems
 * become one block-diagonal problem whose every kernel launch covers all of them, while each keeps
 * its own LM and PCG state (damping, accept/reject, convergence): the same results as `count`
 * separate tcv_ba_solve calls, to rounding. All controls must share every option except the
 * sizes, gauge (anchor, scale_view, normalize_scale) and directional_points. `outs[b]` receives
 * problem b's packBundleResult array. `finish_ms` (optional, `count` entries) receives each
 * problem's time from the start of the LM loop until its loop ended. The reference solves its
 * local/windowed bundles one operator call at a time (src/sfm/mapper/mapper.ts:873-908 via
 * extractLocalBundle, src/sfm/bundle/local-bundle.ts:130-190); this entry has no single
 * counterpart there and replaces a loop of `bundleAdjustGpu` calls.
 */
int tcv_ba_solve_batch(tcv_ba_context* context, uint32_t count, const tcv_ba_control* controls,
                       const double* const* scenes, const uint32_t* const* indices,
                       double* const* outs, tcv_ba_report* report, double* finish_ms);

/*
 * One linear system of the LM loop, for checks and benchmarks: linearise the scene at damping
 * `lambda` (and log focal scale `log_focal` on the shared-focal route), assemble the point blocks,
 * camera blocks and block-Jacobi preconditioner, and return
 *   rhs[m]    the reduced right-hand side (reference `cameraRhs` + focal RHS),
 *   sp[m]     the matrix-free Schur product S·p for the host vector p[m] (m = 6C + shared_focal),
 *   z[m]      the preconditioned vector M^-1·p,
 *   status[1] 0 when every point block and preconditioner block is valid, 1 otherwise.
 * `repeat` > 1 repeats the Schur product (timing); the result is that of the last product.
 * Host pointers. Reference kernels: linearizeBundle, bundlePointBlocks, bundleCameraBlocks,
 * bundleSchurPoint/bundleSchurCamera (bundle.shade.ts:185-225, 251-489; joint-focal.shade.ts).
 */
int tcv_ba_schur_probe(tcv_ba_context* context, const tcv_ba_control* control,
                       const double* scene, const uint32_t* index, double lambda,
                       double log_focal, const double* p, double* rhs, double* sp, double* z,
                       int32_t* status, uint32_t repeat, tcv_ba_report* report);

/*
 * Fixed-scene observation evaluation: evaluateObservationsGpu / evaluateContinuedObservationsGpu
 * (src/geometry/gpu/observations.ts:34-127; kernels observation-evaluation.shade.ts:44-137 and
 * observationSummary, triangulation.shade.ts:134-143). The host packs the request as the
 * reference's observationRequest does (observations.ts:132-197):
 *   cameras  18 doubles per referenced view: the pose record (R row-major, t), then fx fy cx cy k1 k2
 *   points   3 doubles per landmark
 *   pixels   2 doubles per observation, observations in landmark order
 *   indices  2 words per observation: the view's record, the landmark
 *   offsets  (guarded runs: max_relative_depth_std_dev > 0) tracks + 1 words, each track's first
 *            observation; NULL otherwise
 * flags (out) receives `observations` words, 1 for an inlier (projected in front of the camera
 * within max_error_px), then on guarded runs `tracks` words, 1 for a track whose depth was not
 * stable (its observations' flags are then 0); summary (out) receives the inliers' error sum and
 * their count. All arrays are host arrays.
 */
typedef struct tcv_obs_control {
  uint32_t views;         /* camera records */
  uint32_t points;        /* point records */
  uint32_t observations;  /* observations (may be 0) */
  uint32_t tracks;        /* tracks of a guarded run (offsets holds tracks + 1), else 0 */
  double max_error_px;    /* inlier threshold, [1e-12, 1e6] */
  double max_relative_depth_std_dev;  /* > 0: guarded run; 0: no depth filter */
  double pixel_noise_std_dev_px;      /* > 0 */
} tcv_obs_control;

int tcv_obs_evaluate(tcv_ba_context* context, const tcv_obs_control* control,
                     const double* cameras, const double* points, const double* pixels,
                     const uint32_t* indices, const uint32_t* offsets, uint32_t* flags,
                     double* summary);

/*
 * The fixed-scene observation evaluation of tcv_obs_evaluate from DEVICE arrays of the current HIP
 * device, allocating nothing: no hipMalloc, hipFree or allocating library call. No counterpart in
 * the reference; for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). The arguments mean what they mean for
 * tcv_obs_evaluate, but every array is a device array, `flags` and `summary` included, and work
 * runs on `stream` (a hipStream_t, NULL for the default stream); the call returns when flags and
 * summary are complete, with the same bits as tcv_obs_evaluate. The kernels' scratch (per-
 * observation errors, per-64-observation partials, an index-check word) is one fixed device array
 * of the library, 16.5 MiB on each GPU it runs on, held until the process ends: at most
 * TCV_OBS_NOALLOC_MAX_OBSERVATIONS observations (more: TCV_BA_ERR_ARGUMENT). Calls that overlap on
 * one GPU share that array and race: serialise them. Indices and track offsets are checked on the
 * device before any kernel reads through them (TCV_BA_ERR_INDEX).
 */
#define TCV_OBS_NOALLOC_MAX_OBSERVATIONS 2097152u
int tcv_obs_evaluate_noalloc(const tcv_obs_control* control, const double* cameras,
                             const double* points, const double* pixels, const uint32_t* indices,
                             const uint32_t* offsets, uint32_t* flags, double* summary,
                             void* stream);

/*
 * Differentiation of the solve. The reference operator has no derivative entries: these extend
 * `bundleAdjustGpu` (src/sfm/bundle/gpu/bundle.ts:803-839) with the pieces of a backward pass. All
 * take a host scene/index in the packBundleScene layouts (reference bundle-host.ts:121-218) holding
 * the poses and points to differentiate at, and `log_focal`, the shared log focal scale (ignored on
 * the fixed-focal route). Derivatives are of the half robust cost E = 1/2 sum_i robustCost(|r_i|),
 * r_i = pixel_i - u_i (numerics.shade.ts:141-187), in the left-perturbation chart of movePose
 * (numerics.shade.ts:254-281): R' = exp([omega]x) R, t' = exp([omega]x) t + v, X' = X + dX,
 * log f' = log f + df.
 */

/*
 * The Jacobian kernels' output: 23 doubles per observation (packed observation order), r (2),
 * d(u,v)/d(omega, v) (12: row u, then row v; zero for a fixed camera), d(u,v)/dX (6), d(u,v)/d log f
 * (2) and the Huber weight w; all zero with w = -1 where the point does not project. The forward's
 * linearizeBundle (bundle.shade.ts:251-297, joint-focal.shade.ts:287-336).
 */
int tcv_ba_jacobian(tcv_ba_context* context, const tcv_ba_control* control, const double* scene,
                    const uint32_t* index, double log_focal, double* out, tcv_ba_report* report);

/*
 * Polish of a solve to a stationary point of E, on the exact Hessian (Gauss-Newton with the Huber
 * curvature plus the residual-weighted second derivatives), each system solved by the forward's
 * Schur/PCG to `tolerance` within `max_cg` iterations and every step mapped to the forward's gauge
 * (normalizeBundleCandidate, bundle.shade.ts:994-1023):
 *   1. up to `lm_iterations` damped steps with the forward's damping and accept/reject rule
 *      (acceptBundleCandidate, bundle.shade.ts:1071-1096), until that rule stops;
 *   2. up to `newton_steps` undamped Newton steps without a cost test (near the solution the cost
 *      change falls below its own rounding), stopping at the first that cannot be solved.
 * `out` has the tcv_ba_output_size layout of tcv_ba_solve (word 10 counts the accepted damped
 * steps and Newton steps; words 8/9 hold the input's cost); report->refinement_steps counts the
 * Newton steps taken. Makes the stationarity that implicit differentiation assumes hold to rounding.
 */
int tcv_ba_polish(tcv_ba_context* context, const tcv_ba_control* control, const double* scene,
                  const uint32_t* index, double log_focal, uint32_t lm_iterations,
                  uint32_t newton_steps, uint32_t max_cg, double tolerance, double* out,
                  tcv_ba_report* report);

/* Systems of the adjoint (tcv_ba_backward's `mode`). */
enum tcv_ba_adjoint {
  /* the exact Hessian of E (Gauss-Newton with the Huber curvature, plus the residual-weighted
     second derivatives of the projection and the chart): the derivative at a stationary scene */
  TCV_BA_ADJOINT_EXACT = 0,
  /* the Gauss-Newton Hessian sum_i J_i^T D_i J_i (the Huber curvature kept, the second derivatives
     dropped): positive semi-definite, and definite after the gauge regularisation, at any scene; the
     usual linearised sensitivity of a least-squares solution, defined where the exact Hessian at the
     forward's output is not positive definite. Its error against the exact derivative is of the
     order of the residual-weighted curvature it drops */
  TCV_BA_ADJOINT_GAUSS_NEWTON = 1
};

/*
 * The adjoint (backward) of the solve by implicit differentiation: for a cotangent v on the outputs
 * (12 per camera: R row-major, t; 3 per point; then, on the shared route, one for the log focal
 * scale) returns pixel_cotangent[2M] = d<v, theta*>/d(observation pixels) and, when
 * `focal_cotangent` is given, focal_cotangent[C] = d<v, theta*>/d(log of camera c's input focal,
 * fx and fy scaled together). Solves H lambda = (dN)^T v~ with the `mode` system H (enum
 * tcv_ba_adjoint) by the forward's Schur/PCG (to `tolerance`, at most `max_cg` iterations), where
 * v~ is v in the chart and N the gauge map (directional renormalisation, the similarity of the
 * one-fixed-camera gauge), then pixel_cotangent_i = D_i J_i lambda with D_i = d(w r)/dr and, since
 * every column of J_i scales with the focal (d J_i/d log f = J_i),
 *   focal_cotangent_c = -sum over camera c's observations of (D_i J_i lambda) . Jf_i
 *                       + (J_i lambda) . (w_i r_i)   [the second term in the exact mode only]
 * with Jf_i = d pixel_i/d log f. Cross blocks are relinearised from the scene, not kept from the
 * forward. `adjoint` (optional, 6C + 3P + shared doubles: chart rows of each camera, point rows,
 * focal) receives lambda. TCV_BA_ERR_SOLVE when the system is not positive definite or PCG does not
 * converge (the exact mode needs a stationary scene, tcv_ba_polish).
 */
int tcv_ba_backward(tcv_ba_context* context, const tcv_ba_control* control, const double* scene,
                    const uint32_t* index, double log_focal, const double* cotangent,
                    uint32_t max_cg, double tolerance, uint32_t mode, double* pixel_cotangent,
                    double* focal_cotangent, double* adjoint, tcv_ba_report* report);

#ifdef __cplusplus
}
#endif

#endif /* TYPECV_BA_H */
