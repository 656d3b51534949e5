'use strict';
/*
 * The TypeScript operator's side of the engine, in JavaScript for the installed Node: a device
 * with the reference's per-device FIFO of whole jobs (src/device/gpu/device.ts:592-598), the
 * control vector of tcv_ba_control, and the packed solve/backward as Promises, as
 * bundleAdjustGpu's enqueue body (src/sfm/bundle/gpu/bundle.ts:178-799) would call them.
 */
const path = require('path');
const addon = require(path.join(__dirname, 'typecv_ba.node'));

// BUNDLE_STATE, src/sfm/bundle/gpu/bundle-shared.ts:38-54 (the result words after the coordinates)
const BUNDLE_STATE = Object.freeze({
  lambda: 0, active: 1, baseline: 2, cost: 6, sse: 7, initialCost: 8, initialSse: 9,
  accepted: 10, rejected: 11, inputError: 14, cgSteps: 16, failedSolves: 17,
  initialRmsePx: 24, finalRmsePx: 25, focalScale: 26,
});

class HipDevice {
  /* `workingBytes`: ShadeDevice.open's option (device.ts:357-360), the device memory one bundle
   * adjustment may hold; a larger problem streams its cross blocks or is refused. 0: no bound. */
  constructor(index = 0, { workingBytes = 0 } = {}) {
    this.native = addon.open(index);
    this.tail = Promise.resolve();
    this.closed = false;
    if (workingBytes) addon.setWorkingBytes(this.native, workingBytes);
  }
  /* Jobs run one after another on this device, in the order they were enqueued. */
  enqueue(job) {
    if (this.closed) return Promise.reject(new Error('device is closed'));
    const run = this.tail.then(job, job);
    this.tail = run.catch(() => undefined);
    return run;
  }
  get statistics() {
    return { liveBytes: addon.deviceBytes(this.native) };
  }
  drop() {
    addon.release(this.native);
  }
  dispose() {
    this.closed = true;
    addon.release(this.native);
  }
}

/* tcv_ba_control as the addon's Float64Array(16). `schurCache` sets TCV_BA_FLAG_SCHUR_CACHE (4),
 * the reference's option (honoured inside its guard: shared focal, at most 16 cameras,
 * M (6C + 1) <= 262,144); `directSolve` sets TCV_BA_FLAG_DIRECT (8), the direct route for every
 * bundle of at most 2,048 cameras. */
function makeControl({ cameras, points, observations, directionalPoints = 0, anchor = 0,
  scaleView = 1, normalizeScale = true, sharedFocal = false, iterations = 20, maxCg = 24,
  precision = 'fp64', flags = 0, huber = 2, tolerance = 1e-6, stallTolerance = 1e-9,
  refinementTolerance = 0, schurCache = false, directSolve = false }) {
  const bits = flags | (schurCache ? 4 : 0) | (directSolve ? 8 : 0);
  return Float64Array.from([cameras, points, directionalPoints, observations, anchor, scaleView,
    normalizeScale ? 1 : 0, sharedFocal ? 1 : 0, iterations, maxCg,
    precision === 'mixed' ? 1 : 0, bits, huber, tolerance, stallTolerance, refinementTolerance]);
}

/* A packBundleResult array and its diagnostics, with checkBundleOutput's errors
 * (bundle-host.ts:278-283). */
function resultOf(control, values) {
  const base = control[0] * 12 + control[1] * 3;
  if (values[base + BUNDLE_STATE.inputError] === 1)
    throw new RangeError('GPU BA gauge requires a nonzero baseline');
  if (values[base + BUNDLE_STATE.inputError] === 2)
    throw new RangeError('GPU BA observations must initially project in front of their cameras');
  return {
    values,
    diagnostics: {
      initialCost: values[base + BUNDLE_STATE.initialCost],
      finalCost: values[base + BUNDLE_STATE.cost],
      initialRmsePx: values[base + BUNDLE_STATE.initialRmsePx],
      finalRmsePx: values[base + BUNDLE_STATE.finalRmsePx],
      acceptedSteps: values[base + BUNDLE_STATE.accepted],
      rejectedSteps: values[base + BUNDLE_STATE.rejected],
      observations: control[3],
      baseline: values[base + BUNDLE_STATE.baseline],
    },
    cgIterations: values[base + BUNDLE_STATE.cgSteps],
    linearSolverFailures: values[base + BUNDLE_STATE.failedSolves],
  };
}

/* One packed solve. */
function solvePacked(device, control, scene, index) {
  return device.enqueue(async () =>
    resultOf(control, await addon.solve(device.native, control, scene, index)));
}

/* Many independent packed problems in one batched engine call (tcv_ba_solve_batch), as one device
 * job: the results in order. The controls must share every option but sizes and gauge. */
function solveBatchPacked(device, controls, scenes, indices) {
  return device.enqueue(async () => {
    const values = await addon.solveBatch(device.native, controls, scenes, indices);
    return values.map((v, b) => resultOf(controls[b], v));
  });
}

/* The backward entry (tcv_ba_backward): for a cotangent on the solve's outputs, the pixel
 * cotangents (2M) and the input-focal cotangent of each camera (C); mode 'exact' or
 * 'gaussNewton'. */
function backwardPacked(device, control, scene, index, logFocal, cotangent,
  maxCg = 4000, tolerance = 1e-10, mode = 'exact') {
  const code = { exact: 0, gaussNewton: 1 }[mode];
  if (code === undefined) throw new RangeError(`adjoint mode must be 'exact' or 'gaussNewton'`);
  const M = control[3];
  return device.enqueue(async () => {
    const values = await addon.backward(device.native, control, scene, index, logFocal,
      cotangent, maxCg, tolerance, code);
    return { pixels: values.subarray(0, 2 * M), focal: values.subarray(2 * M) };
  });
}

/* The Newton polish of a solve to a stationary point (tcv_ba_polish): the solve's output layout. */
function polishPacked(device, control, scene, index, logFocal, lmIterations = 30, newtonSteps = 3,
  maxCg = 4000, tolerance = 1e-12) {
  return device.enqueue(() => addon.polish(device.native, control, scene, index, logFocal,
    lmIterations, newtonSteps, maxCg, tolerance));
}

/* The Jacobian kernels' output (tcv_ba_jacobian): 23 doubles per observation. */
function jacobianPacked(device, control, scene, index, logFocal = 0) {
  return device.enqueue(() => addon.jacobian(device.native, control, scene, index, logFocal));
}

/* The fixed-scene observation evaluation (evaluateObservationsGpu's device part) of a request
 * packed as observationRequest packs it (src/geometry/gpu/observations.ts:132-197): the flags
 * (observations [+ tracks]) and the summary [error sum, inlier count]. */
function evaluatePacked(device, { views, points, observations, tracks = 0, maxErrorPx = 3,
  maxRelativeDepthStdDev = 0, pixelNoiseStdDevPx = 1 }, cameras, xyz, pixels, indices,
offsets = null) {
  const control = Float64Array.from([views, points, observations, tracks, maxErrorPx,
    maxRelativeDepthStdDev, pixelNoiseStdDevPx]);
  return device.enqueue(() =>
    addon.evaluate(device.native, control, cameras, xyz, pixels, indices, offsets));
}

module.exports = {
  HipDevice, makeControl, solvePacked, solveBatchPacked, backwardPacked, polishPacked,
  jacobianPacked, evaluatePacked, BUNDLE_STATE, addon,
};
