/*
 * typecv_ba_napi.c — the Node-API addon over the engine's C ABI (include/typecv_ba.h).
 *
 * The reference operator `bundleAdjustGpu` (src/sfm/bundle/gpu/bundle.ts:803-839) is a Promise
 * that enqueues one whole job per device (src/device/gpu/device.ts:592-598). This addon gives its
 * TypeScript host the same shape: every solve runs on a libuv worker thread (napi_async_work) and
 * resolves a Promise with the packBundleResult array; errors reject it with the engine's status
 * string and detail. The typed arrays passed in stay referenced until the job completes.
 *
 * Exports:
 *   open(device)                                   -> context (External; destroyed when collected)
 *   outputSize(control)                            -> number of doubles of solve's result
 *   deviceBytes(context), release(context)         -> tcv_ba_device_bytes, tcv_ba_release
 *   solve(context, control, scene, index)          -> Promise<Float64Array>
 *   setWorkingBytes(context, bytes)                -> tcv_ba_set_working_bytes
 *   backward(context, control, scene, index, logFocal, cotangent, maxCg, tolerance[, mode])
 *                                                  -> Promise<Float64Array(2M + C)> (tcv_ba_backward:
 *                                                     the pixel cotangent, then the input-focal
 *                                                     cotangent of each camera; mode 0 exact,
 *                                                     1 Gauss-Newton)
 *   polish(context, control, scene, index, logFocal, lmIterations, newtonSteps, maxCg, tolerance)
 *                                                  -> Promise<Float64Array> (tcv_ba_polish: the
 *                                                     solve's output layout)
 *   jacobian(context, control, scene, index, logFocal)
 *                                                  -> Promise<Float64Array(23M)> (tcv_ba_jacobian)
 *   solveBatch(context, controls[], scenes[], indices[])
 *                                                  -> Promise<Float64Array[]> (tcv_ba_solve_batch)
 *   evaluate(context, control, cameras, points, pixels, indices, offsets|null)
 *                                                  -> Promise<{flags: Uint32Array,
 *                                                     summary: Float64Array(2)}> (tcv_obs_evaluate)
 *     control: Float64Array(7) = views, points, observations, tracks, maxErrorPx,
 *     maxRelativeDepthStdDev, pixelNoiseStdDevPx (the tcv_obs_control fields)
 * `control` is a Float64Array(16): the tcv_ba_control fields in declaration order (12 integers,
 * then huber, relative_tolerance, stall_tolerance, refinement_tolerance); `scene` a Float64Array
 * and `index` a Uint32Array in the packBundleScene layouts (bundle-host.ts:121-218).
 */
#include <node_api.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "typecv_ba.h"

#define CHECK(env, call)                                             \
  do {                                                               \
    if ((call) != napi_ok) {                                         \
      napi_throw_error((env), NULL, "typecv_ba addon: " #call);      \
      return NULL;                                                   \
    }                                                                \
  } while (0)

static void finalize_context(napi_env env, void* data, void* hint) {
  (void)env;
  (void)hint;
  tcv_ba_destroy((tcv_ba_context*)data);
}

static tcv_ba_context* context_of(napi_env env, napi_value v) {
  void* p = NULL;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, NULL, "expected a context from open()");
    return NULL;
  }
  return (tcv_ba_context*)p;
}

/* A typed array's data and length, checked against the expected element type. */
static void* typed_data(napi_env env, napi_value v, napi_typedarray_type want, size_t* length) {
  bool is = false;
  napi_is_typedarray(env, v, &is);
  if (!is) {
    napi_throw_type_error(env, NULL, "expected a typed array");
    return NULL;
  }
  napi_typedarray_type type;
  void* data = NULL;
  napi_value buffer;
  size_t offset = 0;
  napi_get_typedarray_info(env, v, &type, length, &data, &buffer, &offset);
  if (type != want) {
    napi_throw_type_error(env, NULL,
                          want == napi_float64_array ? "expected a Float64Array" : "expected a Uint32Array");
    return NULL;
  }
  return data;
}

static int control_of(napi_env env, napi_value v, tcv_ba_control* c) {
  size_t n = 0;
  const double* f = typed_data(env, v, napi_float64_array, &n);
  if (!f) return 0;
  if (n != 16) {
    napi_throw_range_error(env, NULL, "control must be a Float64Array(16)");
    return 0;
  }
  uint32_t* u = &c->cameras;
  for (int k = 0; k < 12; k++) {
    if (!(f[k] >= 0 && f[k] <= 4294967295.0) || f[k] != (double)(uint32_t)f[k]) {
      napi_throw_range_error(env, NULL, "control integer field out of range");
      return 0;
    }
    u[k] = (uint32_t)f[k];
  }
  c->huber = f[12];
  c->relative_tolerance = f[13];
  c->stall_tolerance = f[14];
  c->refinement_tolerance = f[15];
  return 1;
}

static napi_value open_device(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  int32_t device = 0;
  if (argc > 0) CHECK(env, napi_get_value_int32(env, argv[0], &device));
  tcv_ba_context* ctx = NULL;
  const int status = tcv_ba_create(device, &ctx);
  if (status != TCV_BA_OK) {
    char text[512];
    snprintf(text, sizeof text, "%s (%s)", tcv_ba_status_string(status), tcv_ba_last_error());
    napi_throw_error(env, NULL, text);
    return NULL;
  }
  napi_value out;
  CHECK(env, napi_create_external(env, ctx, finalize_context, NULL, &out));
  return out;
}

static napi_value output_size(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  tcv_ba_control c;
  if (argc < 1 || !control_of(env, argv[0], &c)) return NULL;
  napi_value out;
  CHECK(env, napi_create_double(env, (double)tcv_ba_output_size(&c), &out));
  return out;
}

static napi_value device_bytes(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  tcv_ba_context* ctx = argc > 0 ? context_of(env, argv[0]) : NULL;
  if (!ctx) return NULL;
  napi_value out;
  CHECK(env, napi_create_double(env, (double)tcv_ba_device_bytes(ctx), &out));
  return out;
}

static napi_value set_working_bytes(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  tcv_ba_context* ctx = argc > 1 ? context_of(env, argv[0]) : NULL;
  if (!ctx) return NULL;
  double bytes = 0;
  CHECK(env, napi_get_value_double(env, argv[1], &bytes));
  if (!(bytes >= 0 && bytes <= 9007199254740991.0) || bytes != (double)(uint64_t)bytes) {
    napi_throw_range_error(env, NULL, "workingBytes must be a non-negative integer");
    return NULL;
  }
  tcv_ba_set_working_bytes(ctx, (uint64_t)bytes);
  return NULL;
}

static napi_value release(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  tcv_ba_context* ctx = argc > 0 ? context_of(env, argv[0]) : NULL;
  if (!ctx) return NULL;
  tcv_ba_release(ctx);
  return NULL;
}

/* ---------------------------------------------------------------- observation evaluation */

typedef struct {
  napi_async_work work;
  napi_deferred deferred;
  tcv_ba_context* ctx;
  tcv_obs_control control;
  napi_ref refs[7];
  int nrefs;
  const double *cameras, *points, *pixels;
  const uint32_t *indices, *offsets;
  uint32_t* flags;
  size_t nflags;
  double summary[2];
  int status;
  char detail[256];
} ObsJob;

static void obs_execute(napi_env env, void* data) {
  (void)env;
  ObsJob* j = data;
  j->status = tcv_obs_evaluate(j->ctx, &j->control, j->cameras, j->points, j->pixels, j->indices,
                               j->offsets, j->flags, j->summary);
  if (j->status != TCV_BA_OK) snprintf(j->detail, sizeof j->detail, "%s", tcv_ba_last_error());
}

static void obs_complete(napi_env env, napi_status s, void* data) {
  ObsJob* j = data;
  if (s == napi_ok && j->status == TCV_BA_OK) {
    napi_value result, fbuf, flags, sbuf, summary;
    void* bytes = NULL;
    napi_create_object(env, &result);
    napi_create_arraybuffer(env, j->nflags * sizeof(uint32_t), &bytes, &fbuf);
    if (j->nflags) memcpy(bytes, j->flags, j->nflags * sizeof(uint32_t));
    napi_create_typedarray(env, napi_uint32_array, j->nflags, fbuf, 0, &flags);
    napi_create_arraybuffer(env, 2 * sizeof(double), &bytes, &sbuf);
    memcpy(bytes, j->summary, 2 * sizeof(double));
    napi_create_typedarray(env, napi_float64_array, 2, sbuf, 0, &summary);
    napi_set_named_property(env, result, "flags", flags);
    napi_set_named_property(env, result, "summary", summary);
    napi_resolve_deferred(env, j->deferred, result);
  } else {
    char text[512];
    snprintf(text, sizeof text, "%s (%s)",
             s == napi_ok ? tcv_ba_status_string(j->status) : "job cancelled", j->detail);
    napi_value msg, err;
    napi_create_string_utf8(env, text, NAPI_AUTO_LENGTH, &msg);
    if (j->status == TCV_BA_ERR_ARGUMENT)
      napi_create_range_error(env, NULL, msg, &err);
    else
      napi_create_error(env, NULL, msg, &err);
    napi_reject_deferred(env, j->deferred, err);
  }
  for (int k = 0; k < j->nrefs; k++) napi_delete_reference(env, j->refs[k]);
  napi_delete_async_work(env, j->work);
  free(j->flags);
  free(j);
}

static napi_value evaluate(napi_env env, napi_callback_info info) {
  size_t argc = 7;
  napi_value argv[7];
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 6) {
    napi_throw_type_error(env, NULL, "missing arguments");
    return NULL;
  }
  ObsJob* j = calloc(1, sizeof(ObsJob));
  if (!j) {
    napi_throw_error(env, NULL, "out of host memory");
    return NULL;
  }
  size_t nc = 0, ncam = 0, npt = 0, npx = 0, nix = 0, noff = 0;
  const double* c = NULL;
  j->ctx = context_of(env, argv[0]);
  if (!j->ctx || !(c = typed_data(env, argv[1], napi_float64_array, &nc)) ||
      !(j->cameras = typed_data(env, argv[2], napi_float64_array, &ncam)) ||
      !(j->points = typed_data(env, argv[3], napi_float64_array, &npt)) ||
      !(j->pixels = typed_data(env, argv[4], napi_float64_array, &npx)) ||
      !(j->indices = typed_data(env, argv[5], napi_uint32_array, &nix))) {
    free(j);
    return NULL;
  }
  napi_valuetype t = napi_null;
  if (argc > 6) napi_typeof(env, argv[6], &t);
  if (t != napi_null && t != napi_undefined &&
      !(j->offsets = typed_data(env, argv[6], napi_uint32_array, &noff))) {
    free(j);
    return NULL;
  }
  if (nc != 7) {
    napi_throw_range_error(env, NULL, "control must be a Float64Array(7)");
    free(j);
    return NULL;
  }
  for (int k = 0; k < 4; k++)
    if (!(c[k] >= 0 && c[k] <= 4294967295.0) || c[k] != (double)(uint32_t)c[k]) {
      napi_throw_range_error(env, NULL, "control integer field out of range");
      free(j);
      return NULL;
    }
  j->control.views = (uint32_t)c[0];
  j->control.points = (uint32_t)c[1];
  j->control.observations = (uint32_t)c[2];
  j->control.tracks = (uint32_t)c[3];
  j->control.max_error_px = c[4];
  j->control.max_relative_depth_std_dev = c[5];
  j->control.pixel_noise_std_dev_px = c[6];
  const size_t M = j->control.observations, T = j->control.tracks;
  if (ncam < 18 * (size_t)j->control.views || npt < 3 * (size_t)j->control.points ||
      npx < 2 * M || nix < 2 * M || (j->offsets && noff < T + 1) || (T && !j->offsets)) {
    napi_throw_range_error(env, NULL, "observation arrays are shorter than the control implies");
    free(j);
    return NULL;
  }
  j->nflags = M + T;
  j->flags = malloc((j->nflags ? j->nflags : 1) * sizeof(uint32_t));
  if (!j->flags) {
    free(j);
    napi_throw_error(env, NULL, "out of host memory");
    return NULL;
  }
  for (size_t k = 0; k < (j->offsets ? 7u : 6u); k++)
    if (k != 1) napi_create_reference(env, argv[k], 1, &j->refs[j->nrefs++]);
  napi_value promise, resource;
  if (napi_create_promise(env, &j->deferred, &promise) != napi_ok ||
      napi_create_string_utf8(env, "tcv_obs_evaluate", NAPI_AUTO_LENGTH, &resource) != napi_ok ||
      napi_create_async_work(env, NULL, resource, obs_execute, obs_complete, j, &j->work) !=
          napi_ok ||
      napi_queue_async_work(env, j->work) != napi_ok) {
    for (int k = 0; k < j->nrefs; k++) napi_delete_reference(env, j->refs[k]);
    free(j->flags);
    free(j);
    napi_throw_error(env, NULL, "typecv_ba addon: could not queue the job");
    return NULL;
  }
  return promise;
}

/* ---------------------------------------------------------------- asynchronous jobs */

enum { JOB_SOLVE = 0, JOB_BACKWARD = 1, JOB_POLISH = 2, JOB_JACOBIAN = 3 };

typedef struct {
  int kind;
  napi_async_work work;
  napi_deferred deferred;
  tcv_ba_context* ctx;
  tcv_ba_control control;
  napi_ref refs[4]; /* context, scene, index, cotangent */
  int nrefs;
  const double* scene;
  const uint32_t* index;
  const double* cotangent;
  double log_focal, tolerance;
  uint32_t max_cg, lm_iterations, newton_steps, mode;
  double* out;
  size_t out_len;
  tcv_ba_report report;
  int status;
  char detail[256];
} Job;

static void execute(napi_env env, void* data) {
  (void)env;
  Job* j = data;
  if (j->kind == JOB_SOLVE)
    j->status = tcv_ba_solve(j->ctx, &j->control, j->scene, j->index, j->out, &j->report);
  else if (j->kind == JOB_BACKWARD)
    j->status = tcv_ba_backward(j->ctx, &j->control, j->scene, j->index, j->log_focal,
                                j->cotangent, j->max_cg, j->tolerance, j->mode, j->out,
                                j->out + 2 * (size_t)j->control.observations, NULL, &j->report);
  else if (j->kind == JOB_POLISH)
    j->status = tcv_ba_polish(j->ctx, &j->control, j->scene, j->index, j->log_focal,
                              j->lm_iterations, j->newton_steps, j->max_cg, j->tolerance, j->out,
                              &j->report);
  else
    j->status = tcv_ba_jacobian(j->ctx, &j->control, j->scene, j->index, j->log_focal, j->out,
                                &j->report);
  if (j->status != TCV_BA_OK) snprintf(j->detail, sizeof j->detail, "%s", tcv_ba_last_error());
}

static void complete(napi_env env, napi_status s, void* data) {
  Job* j = data;
  if (s == napi_ok && j->status == TCV_BA_OK) {
    napi_value buffer, array;
    void* bytes = NULL;
    napi_create_arraybuffer(env, j->out_len * sizeof(double), &bytes, &buffer);
    memcpy(bytes, j->out, j->out_len * sizeof(double));
    napi_create_typedarray(env, napi_float64_array, j->out_len, buffer, 0, &array);
    napi_resolve_deferred(env, j->deferred, array);
  } else {
    char text[512];
    snprintf(text, sizeof text, "%s (%s)",
             s == napi_ok ? tcv_ba_status_string(j->status) : "job cancelled", j->detail);
    napi_value msg, err;
    napi_create_string_utf8(env, text, NAPI_AUTO_LENGTH, &msg);
    if (j->status == TCV_BA_ERR_ARGUMENT)
      napi_create_range_error(env, NULL, msg, &err);
    else
      napi_create_error(env, NULL, msg, &err);
    napi_reject_deferred(env, j->deferred, err);
  }
  for (int k = 0; k < j->nrefs; k++) napi_delete_reference(env, j->refs[k]);
  napi_delete_async_work(env, j->work);
  free(j->out);
  free(j);
}

static napi_value queue(napi_env env, Job* j, const char* name) {
  napi_value promise, resource;
  if (napi_create_promise(env, &j->deferred, &promise) != napi_ok ||
      napi_create_string_utf8(env, name, NAPI_AUTO_LENGTH, &resource) != napi_ok ||
      napi_create_async_work(env, NULL, resource, execute, complete, j, &j->work) != napi_ok ||
      napi_queue_async_work(env, j->work) != napi_ok) {
    for (int k = 0; k < j->nrefs; k++) napi_delete_reference(env, j->refs[k]);
    free(j->out);
    free(j);
    napi_throw_error(env, NULL, "typecv_ba addon: could not queue the job");
    return NULL;
  }
  return promise;
}

/* Common arguments: context, control, scene, index, with the sizes the control implies. */
static Job* job_of(napi_env env, napi_value* argv, size_t argc, size_t need, int kind) {
  if (argc < need) {
    napi_throw_type_error(env, NULL, "missing arguments");
    return NULL;
  }
  Job* j = calloc(1, sizeof(Job));
  if (!j) {
    napi_throw_error(env, NULL, "out of host memory");
    return NULL;
  }
  j->kind = kind;
  j->ctx = context_of(env, argv[0]);
  size_t ns = 0, ni = 0;
  if (!j->ctx || !control_of(env, argv[1], &j->control) ||
      !(j->scene = typed_data(env, argv[2], napi_float64_array, &ns)) ||
      !(j->index = typed_data(env, argv[3], napi_uint32_array, &ni))) {
    free(j);
    return NULL;
  }
  if (ns != tcv_ba_scene_size(&j->control) || ni != tcv_ba_index_size(&j->control)) {
    napi_throw_range_error(env, NULL, "scene/index lengths do not match the control sizes");
    free(j);
    return NULL;
  }
  napi_create_reference(env, argv[0], 1, &j->refs[j->nrefs++]);
  napi_create_reference(env, argv[2], 1, &j->refs[j->nrefs++]);
  napi_create_reference(env, argv[3], 1, &j->refs[j->nrefs++]);
  return j;
}

static napi_value solve(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value argv[4];
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  Job* j = job_of(env, argv, argc, 4, JOB_SOLVE);
  if (!j) return NULL;
  j->out_len = tcv_ba_output_size(&j->control);
  j->out = malloc(j->out_len * sizeof(double));
  return queue(env, j, "typecv_ba.solve");
}

static napi_value backward(napi_env env, napi_callback_info info) {
  size_t argc = 9;
  napi_value argv[9];
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  Job* j = job_of(env, argv, argc, 8, JOB_BACKWARD);
  if (!j) return NULL;
  size_t nc = 0;
  double max_cg = 0, mode = 0;
  if (napi_get_value_double(env, argv[4], &j->log_focal) != napi_ok ||
      !(j->cotangent = typed_data(env, argv[5], napi_float64_array, &nc)) ||
      napi_get_value_double(env, argv[6], &max_cg) != napi_ok ||
      napi_get_value_double(env, argv[7], &j->tolerance) != napi_ok ||
      (argc > 8 && napi_get_value_double(env, argv[8], &mode) != napi_ok)) {
    for (int k = 0; k < j->nrefs; k++) napi_delete_reference(env, j->refs[k]);
    free(j);
    return NULL;
  }
  const size_t want = (size_t)j->control.cameras * 12 + (size_t)j->control.points * 3 +
                      (j->control.shared_focal ? 1 : 0);
  if (nc != want || !(max_cg >= 1 && max_cg <= 100000) || !(mode == 0 || mode == 1)) {
    for (int k = 0; k < j->nrefs; k++) napi_delete_reference(env, j->refs[k]);
    free(j);
    napi_throw_range_error(env, NULL, "cotangent length, maxCg or mode out of range");
    return NULL;
  }
  j->max_cg = (uint32_t)max_cg;
  j->mode = (uint32_t)mode;
  napi_create_reference(env, argv[5], 1, &j->refs[j->nrefs++]);
  j->out_len = 2 * (size_t)j->control.observations + j->control.cameras;
  j->out = malloc(j->out_len * sizeof(double));
  return queue(env, j, "typecv_ba.backward");
}

/* A JavaScript number that is an integer in [lo, hi], or 0 with a pending RangeError. */
static int integer_of(napi_env env, napi_value v, double lo, double hi, uint32_t* out) {
  double x = 0;
  if (napi_get_value_double(env, v, &x) != napi_ok || !(x >= lo && x <= hi) ||
      x != (double)(uint32_t)x) {
    napi_throw_range_error(env, NULL, "integer argument out of range");
    return 0;
  }
  *out = (uint32_t)x;
  return 1;
}

static napi_value polish(napi_env env, napi_callback_info info) {
  size_t argc = 9;
  napi_value argv[9];
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  Job* j = job_of(env, argv, argc, 9, JOB_POLISH);
  if (!j) return NULL;
  if (napi_get_value_double(env, argv[4], &j->log_focal) != napi_ok ||
      !integer_of(env, argv[5], 0, 1000, &j->lm_iterations) ||
      !integer_of(env, argv[6], 0, 100, &j->newton_steps) ||
      !integer_of(env, argv[7], 1, 100000, &j->max_cg) ||
      napi_get_value_double(env, argv[8], &j->tolerance) != napi_ok) {
    for (int k = 0; k < j->nrefs; k++) napi_delete_reference(env, j->refs[k]);
    free(j);
    return NULL;
  }
  j->out_len = tcv_ba_output_size(&j->control);
  j->out = malloc(j->out_len * sizeof(double));
  return queue(env, j, "typecv_ba.polish");
}

static napi_value jacobian(napi_env env, napi_callback_info info) {
  size_t argc = 5;
  napi_value argv[5];
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  Job* j = job_of(env, argv, argc, 5, JOB_JACOBIAN);
  if (!j) return NULL;
  if (napi_get_value_double(env, argv[4], &j->log_focal) != napi_ok) {
    for (int k = 0; k < j->nrefs; k++) napi_delete_reference(env, j->refs[k]);
    free(j);
    return NULL;
  }
  j->out_len = 23 * (size_t)j->control.observations;
  j->out = malloc((j->out_len ? j->out_len : 1) * sizeof(double));
  return queue(env, j, "typecv_ba.jacobian");
}

/* ---------------------------------------------------------------- batches */

typedef struct {
  napi_async_work work;
  napi_deferred deferred;
  tcv_ba_context* ctx;
  uint32_t count;
  tcv_ba_control* controls;
  const double** scenes;
  const uint32_t** indices;
  double** outs;
  size_t* out_lens;
  napi_ref* refs; /* context, then scene and index of each bundle */
  uint32_t nrefs;
  tcv_ba_report report;
  int status;
  char detail[256];
} BatchJob;

static void free_batch(napi_env env, BatchJob* j) {
  for (uint32_t k = 0; k < j->nrefs; k++) napi_delete_reference(env, j->refs[k]);
  if (j->outs)
    for (uint32_t b = 0; b < j->count; b++) free(j->outs[b]);
  free(j->controls);
  free(j->scenes);
  free(j->indices);
  free(j->outs);
  free(j->out_lens);
  free(j->refs);
  free(j);
}

static void execute_batch(napi_env env, void* data) {
  (void)env;
  BatchJob* j = data;
  j->status = tcv_ba_solve_batch(j->ctx, j->count, j->controls, j->scenes, j->indices, j->outs,
                                 &j->report, NULL);
  if (j->status != TCV_BA_OK) snprintf(j->detail, sizeof j->detail, "%s", tcv_ba_last_error());
}

static void complete_batch(napi_env env, napi_status s, void* data) {
  BatchJob* j = data;
  if (s == napi_ok && j->status == TCV_BA_OK) {
    napi_value list;
    napi_create_array_with_length(env, j->count, &list);
    for (uint32_t b = 0; b < j->count; b++) {
      napi_value buffer, array;
      void* bytes = NULL;
      napi_create_arraybuffer(env, j->out_lens[b] * sizeof(double), &bytes, &buffer);
      memcpy(bytes, j->outs[b], j->out_lens[b] * sizeof(double));
      napi_create_typedarray(env, napi_float64_array, j->out_lens[b], buffer, 0, &array);
      napi_set_element(env, list, b, array);
    }
    napi_resolve_deferred(env, j->deferred, list);
  } else {
    char text[512];
    snprintf(text, sizeof text, "%s (%s)",
             s == napi_ok ? tcv_ba_status_string(j->status) : "job cancelled", j->detail);
    napi_value msg, err;
    napi_create_string_utf8(env, text, NAPI_AUTO_LENGTH, &msg);
    if (j->status == TCV_BA_ERR_ARGUMENT)
      napi_create_range_error(env, NULL, msg, &err);
    else
      napi_create_error(env, NULL, msg, &err);
    napi_reject_deferred(env, j->deferred, err);
  }
  napi_delete_async_work(env, j->work);
  free_batch(env, j);
}

/* Element b of a JavaScript array, or NULL with a pending TypeError. */
static napi_value element_of(napi_env env, napi_value list, uint32_t b) {
  napi_value v = NULL;
  if (napi_get_element(env, list, b, &v) != napi_ok) {
    napi_throw_type_error(env, NULL, "expected arrays of typed arrays");
    return NULL;
  }
  return v;
}

static napi_value solve_batch(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value argv[4];
  CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 4) {
    napi_throw_type_error(env, NULL, "missing arguments");
    return NULL;
  }
  tcv_ba_context* ctx = context_of(env, argv[0]);
  if (!ctx) return NULL;
  uint32_t count = 0, ns = 0, ni = 0;
  bool arrays[3] = {false, false, false};
  for (int k = 0; k < 3; k++) napi_is_array(env, argv[1 + k], &arrays[k]);
  if (!arrays[0] || !arrays[1] || !arrays[2] ||
      napi_get_array_length(env, argv[1], &count) != napi_ok ||
      napi_get_array_length(env, argv[2], &ns) != napi_ok ||
      napi_get_array_length(env, argv[3], &ni) != napi_ok || count < 1 || ns != count ||
      ni != count) {
    napi_throw_range_error(env, NULL, "controls, scenes and indices must be arrays of one length");
    return NULL;
  }
  BatchJob* j = calloc(1, sizeof(BatchJob));
  if (!j) {
    napi_throw_error(env, NULL, "out of host memory");
    return NULL;
  }
  j->ctx = ctx;
  j->count = count;
  j->controls = calloc(count, sizeof(tcv_ba_control));
  j->scenes = calloc(count, sizeof(double*));
  j->indices = calloc(count, sizeof(uint32_t*));
  j->outs = calloc(count, sizeof(double*));
  j->out_lens = calloc(count, sizeof(size_t));
  j->refs = calloc(1 + 2 * (size_t)count, sizeof(napi_ref));
  if (!j->controls || !j->scenes || !j->indices || !j->outs || !j->out_lens || !j->refs) {
    free_batch(env, j);
    napi_throw_error(env, NULL, "out of host memory");
    return NULL;
  }
  napi_create_reference(env, argv[0], 1, &j->refs[j->nrefs++]);
  for (uint32_t b = 0; b < count; b++) {
    napi_value c = element_of(env, argv[1], b), sv = element_of(env, argv[2], b),
               iv = element_of(env, argv[3], b);
    size_t nsv = 0, niv = 0;
    if (!c || !sv || !iv || !control_of(env, c, &j->controls[b]) ||
        !(j->scenes[b] = typed_data(env, sv, napi_float64_array, &nsv)) ||
        !(j->indices[b] = typed_data(env, iv, napi_uint32_array, &niv))) {
      free_batch(env, j);
      return NULL;
    }
    if (nsv != tcv_ba_scene_size(&j->controls[b]) || niv != tcv_ba_index_size(&j->controls[b])) {
      free_batch(env, j);
      napi_throw_range_error(env, NULL, "scene/index lengths do not match the control sizes");
      return NULL;
    }
    napi_create_reference(env, sv, 1, &j->refs[j->nrefs++]);
    napi_create_reference(env, iv, 1, &j->refs[j->nrefs++]);
    j->out_lens[b] = tcv_ba_output_size(&j->controls[b]);
    j->outs[b] = malloc(j->out_lens[b] * sizeof(double));
    if (!j->outs[b]) {
      free_batch(env, j);
      napi_throw_error(env, NULL, "out of host memory");
      return NULL;
    }
  }
  napi_value promise, resource;
  if (napi_create_promise(env, &j->deferred, &promise) != napi_ok ||
      napi_create_string_utf8(env, "typecv_ba.solveBatch", NAPI_AUTO_LENGTH, &resource) != napi_ok ||
      napi_create_async_work(env, NULL, resource, execute_batch, complete_batch, j, &j->work) !=
          napi_ok ||
      napi_queue_async_work(env, j->work) != napi_ok) {
    free_batch(env, j);
    napi_throw_error(env, NULL, "typecv_ba addon: could not queue the job");
    return NULL;
  }
  return promise;
}

static napi_value init(napi_env env, napi_value exports) {
  const napi_property_descriptor props[] = {
      {"open", NULL, open_device, NULL, NULL, NULL, napi_enumerable, NULL},
      {"outputSize", NULL, output_size, NULL, NULL, NULL, napi_enumerable, NULL},
      {"deviceBytes", NULL, device_bytes, NULL, NULL, NULL, napi_enumerable, NULL},
      {"release", NULL, release, NULL, NULL, NULL, napi_enumerable, NULL},
      {"setWorkingBytes", NULL, set_working_bytes, NULL, NULL, NULL, napi_enumerable, NULL},
      {"polish", NULL, polish, NULL, NULL, NULL, napi_enumerable, NULL},
      {"jacobian", NULL, jacobian, NULL, NULL, NULL, napi_enumerable, NULL},
      {"solve", NULL, solve, NULL, NULL, NULL, napi_enumerable, NULL},
      {"backward", NULL, backward, NULL, NULL, NULL, napi_enumerable, NULL},
      {"solveBatch", NULL, solve_batch, NULL, NULL, NULL, napi_enumerable, NULL},
      {"evaluate", NULL, evaluate, NULL, NULL, NULL, napi_enumerable, NULL},
  };
  napi_define_properties(env, exports, sizeof props / sizeof props[0], props);
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
