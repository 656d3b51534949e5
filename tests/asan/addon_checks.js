'use strict';
/*
 * TEST INFRASTRUCTURE: the Node-API addon's host code on the CPU, for the AddressSanitizer test
 * (tests/test_host.py). `node addon_checks.js <addon.node>` loads the addon given, drives every
 * export through its argument checks (valid and invalid controls, wrong types and lengths, a
 * context that is not one, opening a device where there may be none) and prints a JSON summary.
 * No device work is queued.
 */
const addon = require(process.argv[2]);

const control = Float64Array.from([10, 100, 0, 1000, 0, 1, 1, 1, 20, 24, 0, 0, 2, 1e-6, 1e-9, 0]);
const results = { outputSize: addon.outputSize(control), threw: {} };

function expectThrow(name, fn) {
  try {
    fn();
    results.threw[name] = false;
  } catch (e) {
    results.threw[name] = e instanceof Error;
  }
}

const fake = {};
const scene = new Float64Array(180 + 300 + 2000);
const index = new Uint32Array(3000 + 100 + 22);
expectThrow('outputSize/short', () => addon.outputSize(new Float64Array(3)));
expectThrow('outputSize/type', () => addon.outputSize([1, 2, 3]));
expectThrow('outputSize/negative', () => addon.outputSize(Float64Array.from(control, (v, k) => (k === 0 ? -1 : v))));
expectThrow('solve/context', () => addon.solve(fake, control, scene, index));
expectThrow('solve/missing', () => addon.solve(fake, control));
expectThrow('backward/context', () => addon.backward(fake, control, scene, index, 0, new Float64Array(421), 10, 1e-10));
expectThrow('polish/context', () => addon.polish(fake, control, scene, index, 0, 1, 1, 10, 1e-12));
expectThrow('jacobian/context', () => addon.jacobian(fake, control, scene, index, 0));
expectThrow('solveBatch/context', () => addon.solveBatch(fake, [control], [scene], [index]));
expectThrow('evaluate/context', () => addon.evaluate(fake, new Float64Array(7), new Float64Array(18), new Float64Array(3), new Float64Array(2), new Uint32Array(2), null));
expectThrow('deviceBytes/context', () => addon.deviceBytes(fake));
expectThrow('release/context', () => addon.release(fake));
expectThrow('setWorkingBytes/context', () => addon.setWorkingBytes(fake, 4096));
let opened = false;
try {
  const ctx = addon.open(0);
  opened = true;
  // a device exists: the checks that need a context, then its release
  expectThrow('solve/lengths', () => addon.solve(ctx, control, new Float64Array(5), index));
  expectThrow('backward/cotangent', () => addon.backward(ctx, control, scene, index, 0, new Float64Array(3), 10, 1e-10));
  expectThrow('backward/mode', () => addon.backward(ctx, control, scene, index, 0, new Float64Array(421), 10, 1e-10, 7));
  addon.release(ctx);
} catch (e) {
  results.openError = String(e.message || e);
}
results.opened = opened;
console.log(JSON.stringify(results));
