'use strict';
/*
 * Command-line runner of the packed solve through the Node-API addon:
 *   node run.js <control.f64> <scene.f64> <index.u32> <out.f64>
 * reads the raw little-endian arrays, solves on device 0 and writes the packBundleResult array;
 *   node run.js --batch <out prefix> (<control.f64> <scene.f64> <index.u32>)...
 * solves the problems in one batched call and writes <out prefix>.<b>.f64 for each.
 *   node run.js --observations <control.f64> <cameras.f64> <points.f64> <pixels.f64>
 *       <indices.u32> <offsets.u32 | -> <out prefix>
 * evaluates a packed observation request (control: views, points, observations, tracks,
 * maxErrorPx, maxRelativeDepthStdDev, pixelNoiseStdDevPx) and writes <out prefix>.flags.u32 and
 * <out prefix>.summary.f64.
 *   node run.js --concurrent <control.f64> <scene.f64> <index.u32> <out prefix>
 * issues two solves on one context at once through the addon itself (no FIFO in between), then a
 * third after both, and writes <out prefix>.0.f64, .1.f64 and .seq.f64;
 *   node run.js --polish <control.f64> <scene.f64> <index.u32> <logFocal> <out.f64>
 *   node run.js --jacobian <control.f64> <scene.f64> <index.u32> <logFocal> <out.f64>
 * run tcv_ba_polish (30 damped steps, 3 Newton steps, PCG <= 4000 to 1e-12) or tcv_ba_jacobian;
 *   node run.js --backward <control.f64> <scene.f64> <index.u32> <logFocal> <cotangent.f64>
 *       <exact|gaussNewton> <out prefix>
 * runs tcv_ba_backward (PCG <= 4000 to 1e-10) and writes <out prefix>.pixels.f64 and
 * <out prefix>.focal.f64.
 * `node run.js --exports` prints the addon's exported names (no device is touched).
 */
const fs = require('fs');
const { HipDevice, solvePacked, solveBatchPacked, polishPacked, jacobianPacked, backwardPacked,
  evaluatePacked, addon } = require('./bundle.js');

function read(file, Type) {
  const bytes = fs.readFileSync(file);
  return new Type(bytes.buffer.slice(bytes.byteOffset, bytes.byteOffset + bytes.byteLength));
}

async function main(argv) {
  if (argv[0] === '--exports') {
    console.log(JSON.stringify(Object.keys(addon).sort()));
    return 0;
  }
  if (argv[0] === '--observations') {
    const [, controlFile, camerasFile, pointsFile, pixelsFile, indicesFile, offsetsFile, prefix] =
      argv;
    const c = read(controlFile, Float64Array);
    const offsets = offsetsFile === '-' ? null : read(offsetsFile, Uint32Array);
    const device = new HipDevice(0);
    try {
      const r = await evaluatePacked(device, {
        views: c[0], points: c[1], observations: c[2], tracks: c[3], maxErrorPx: c[4],
        maxRelativeDepthStdDev: c[5], pixelNoiseStdDevPx: c[6],
      }, read(camerasFile, Float64Array), read(pointsFile, Float64Array),
      read(pixelsFile, Float64Array), read(indicesFile, Uint32Array), offsets);
      fs.writeFileSync(`${prefix}.flags.u32`, Buffer.from(r.flags.buffer));
      fs.writeFileSync(`${prefix}.summary.f64`, Buffer.from(r.summary.buffer));
      console.log(JSON.stringify({ inliers: r.summary[1] }));
    } finally {
      device.dispose();
    }
    return 0;
  }
  if (argv[0] === '--concurrent') {
    const [, controlFile, sceneFile, indexFile, prefix] = argv;
    const control = read(controlFile, Float64Array);
    const scene = read(sceneFile, Float64Array), index = read(indexFile, Uint32Array);
    const device = new HipDevice(0);
    try {
      const both = await Promise.all([addon.solve(device.native, control, scene, index),
        addon.solve(device.native, control, scene, index)]);
      const after = await addon.solve(device.native, control, scene, index);
      both.forEach((v, k) => fs.writeFileSync(`${prefix}.${k}.f64`, Buffer.from(v.buffer)));
      fs.writeFileSync(`${prefix}.seq.f64`, Buffer.from(after.buffer));
      console.log(JSON.stringify({ solves: 3 }));
    } finally {
      device.dispose();
    }
    return 0;
  }
  if (argv[0] === '--polish' || argv[0] === '--jacobian') {
    const [mode, controlFile, sceneFile, indexFile, logFocal, outFile] = argv;
    const control = read(controlFile, Float64Array);
    const scene = read(sceneFile, Float64Array), index = read(indexFile, Uint32Array);
    const device = new HipDevice(0);
    try {
      const values = mode === '--polish'
        ? await polishPacked(device, control, scene, index, Number(logFocal), 30, 3, 4000, 1e-12)
        : await jacobianPacked(device, control, scene, index, Number(logFocal));
      fs.writeFileSync(outFile, Buffer.from(values.buffer));
      console.log(JSON.stringify({ length: values.length }));
    } finally {
      device.dispose();
    }
    return 0;
  }
  if (argv[0] === '--backward') {
    const [, controlFile, sceneFile, indexFile, logFocal, cotangentFile, mode, prefix] = argv;
    const control = read(controlFile, Float64Array);
    const scene = read(sceneFile, Float64Array), index = read(indexFile, Uint32Array);
    const device = new HipDevice(0);
    try {
      const r = await backwardPacked(device, control, scene, index, Number(logFocal),
        read(cotangentFile, Float64Array), 4000, 1e-10, mode);
      fs.writeFileSync(`${prefix}.pixels.f64`, Buffer.from(r.pixels.slice().buffer));
      fs.writeFileSync(`${prefix}.focal.f64`, Buffer.from(r.focal.slice().buffer));
      console.log(JSON.stringify({ pixels: r.pixels.length, focal: r.focal.length }));
    } finally {
      device.dispose();
    }
    return 0;
  }
  if (argv[0] === '--batch') {
    const prefix = argv[1], files = argv.slice(2);
    const controls = [], scenes = [], indices = [];
    for (let k = 0; k + 2 < files.length; k += 3) {
      controls.push(read(files[k], Float64Array));
      scenes.push(read(files[k + 1], Float64Array));
      indices.push(read(files[k + 2], Uint32Array));
    }
    const device = new HipDevice(0);
    try {
      const results = await solveBatchPacked(device, controls, scenes, indices);
      results.forEach((r, b) => fs.writeFileSync(`${prefix}.${b}.f64`, Buffer.from(r.values.buffer)));
      console.log(JSON.stringify(results.map((r) => r.diagnostics)));
    } finally {
      device.dispose();
    }
    return 0;
  }
  const [controlFile, sceneFile, indexFile, outFile] = argv;
  const control = read(controlFile, Float64Array);
  const scene = read(sceneFile, Float64Array);
  const index = read(indexFile, Uint32Array);
  const device = new HipDevice(0);
  try {
    const result = await solvePacked(device, control, scene, index);
    fs.writeFileSync(outFile, Buffer.from(result.values.buffer));
    console.log(JSON.stringify({ diagnostics: result.diagnostics, cgIterations: result.cgIterations }));
  } finally {
    device.dispose();
  }
  return 0;
}

main(process.argv.slice(2)).then((code) => process.exit(code), (error) => {
  console.error(String(error && error.stack || error));
  process.exit(1);
});
