// @ts-check
'use strict';
// Runs the tiny workload through the JavaScript host (MacSolver over
// the N-API addon) and prints the fields' checksums and step stats as JSON.
const path = require('path');
const root = path.join(__dirname, '..', '..', 'drop-of-blood_amd', 'js');
const { MacSolver } = require(path.join(root, 'MacSolver.js'));
const { workload } = require(path.join(root, 'scenarios.js'));

const steps = Number(process.argv[2] || 4);
const w = workload('tiny', 1);
const s = new MacSolver({ nx: w.nx, ny: w.ny, nz: w.nz, dx: w.dx, levels: w.levels,
  cycles: w.cycles, precision: w.precision });
s.initBlobs(w.blobs);
let bad = null;
try { s.step(NaN, 1); } catch (e) { bad = String(e.message); }
const st = s.step(w.dt, steps);
s.sync();
const sum = (a) => { let t = 0; for (let q = 0; q < a.length; q++) t += a[q] * a[q]; return Math.sqrt(t); };
const out = { stats: st, norms: { u: sum(s.u), v: sum(s.v), w: sum(s.w), p: sum(s.p), dye: sum(s.dye) },
  badDt: bad };
s.destroy();
s.destroy();
console.log(JSON.stringify(out));

// adjoint through the TypeScript host: gradient of the tiny run's dye
// against a zero target must be finite and match the C ABI's (checked by the
// pytest driver)
// the free-surface coupling through the same host: checksums of the fields
// and the wave state after the steps (compared with the C ABI's by pytest)
if (process.argv[3] === 'wave') {
  const s3 = new MacSolver({ nx: w.nx, ny: w.ny, nz: w.nz, dx: w.dx, levels: w.levels,
    cycles: w.cycles, precision: 0 });
  s3.initBlobs(w.blobs);
  s3.waveEnable({ n: 64, halfExtent: 0.06, halfXZ: 0.016, depth: 0.048, surfaceY: 0.0,
    marangoni: 0.8 });
  s3.setDrops([{ x: 0.004, y: -0.003, z: -0.002, radius: 0.02, velY: -0.4, drain: 0.55, seed: 1.3 },
    { x: -0.01, y: 0.01, z: 0.006, radius: 0.018, velY: -0.2, drain: 0.2, seed: 2.1 }]);
  s3.step(w.dt, steps);
  s3.sync();
  const wave = s3.read('WAVE', new Float64Array(64 * 64 * 4));
  const ptop = s3.read('PTOP', new Float64Array(w.nx * w.nz));
  console.log(JSON.stringify({ wave: { v: sum(s3.v), p: sum(s3.p), dye: sum(s3.dye),
    wave: sum(wave), ptop: sum(ptop) } }));
  s3.destroy();
}

// the parcel stage through the same host: 512 parcels laid out by formula
// after the coupled steps; checksums of the parcels and both splat counters
if (process.argv[3] === 'parcels') {
  const s4 = new MacSolver({ nx: w.nx, ny: w.ny, nz: w.nz, dx: w.dx, levels: w.levels,
    cycles: w.cycles, precision: 0 });
  s4.initBlobs(w.blobs);
  s4.waveEnable({ n: 64, halfExtent: 0.06, halfXZ: 0.016, depth: 0.048, surfaceY: 0.0 });
  s4.step(w.dt, steps);
  const n = 512;
  const P = new Float32Array(8 * n);
  for (let q = 0; q < n; q++) {
    P.set([-0.008 + 0.016 * ((q * 37) % 64) / 64, -0.04 + 0.035 * ((q * 11) % 128) / 128,
      -0.008 + 0.016 * ((q * 23) % 64) / 64, 0.1, q % 3 ? 0.002 : 0.006, q % 3 ? 0 : 0.004, 0,
      (q * 7) % 3000], 8 * q);
  }
  s4.parcelsUpload(P);
  s4.parcelsStep(w.dt);
  const out = s4.parcelsDownload(new Float32Array(8 * n));
  const cells = w.nx * w.ny * w.nz;
  const a = s4.splatDownload(0, new Uint32Array(cells)), b = s4.splatDownload(1, new Uint32Array(cells));
  console.log(JSON.stringify({ parcels: { p: sum(out), intact: a.reduce((x, y) => x + y, 0),
    lysed: b.reduce((x, y) => x + y, 0) } }));
  s4.destroy();
}

// the 4-channel concentration on the grid twice as fine, through the same
// host: coupled steps with emission; checksums of the fields, the
// concentration (4 per fine cell) and both fine-grid splat counters
if (process.argv[3] === 'conc2') {
  const s5 = new MacSolver({ nx: w.nx, ny: w.ny, nz: w.nz, dx: w.dx, levels: w.levels,
    cycles: w.cycles, precision: 0 });
  s5.initBlobs(w.blobs);
  s5.waveEnable({ n: 64, halfExtent: 0.25, halfXZ: 0.1, depth: 0.2, surfaceY: 0.0,
    marangoni: 0.8, scene: 1, physics: 1 });
  s5.setDrops([{ x: 0.004, y: -0.003, z: -0.002, radius: 0.02, velY: -0.4, drain: 0.85, seed: 1.3 }]);
  s5.concEnable(true, 2);
  s5.parcelsUpload(new Float32Array(8 * 4096));
  for (let q = 0; q < steps; q++) {
    s5.parcelsEmit(300 * q, 300);
    s5.step(w.dt, 1);
  }
  s5.sync();
  const nc = s5.concCells;
  const conc = s5.read('CONC', new Float64Array(4 * nc));
  const a = s5.splatDownload(0, new Uint32Array(nc)), b = s5.splatDownload(1, new Uint32Array(nc));
  console.log(JSON.stringify({ conc2: { cells: nc, v: sum(s5.v), conc: sum(conc),
    intact: a.reduce((x, y) => x + y, 0), lysed: b.reduce((x, y) => x + y, 0) } }));
  s5.destroy();
}

if (process.argv[3] === 'backward') {
  const s2 = new MacSolver({ nx: w.nx, ny: w.ny, nz: w.nz, dx: w.dx, levels: w.levels,
    cycles: w.cycles, precision: 2 });
  const target = new Float64Array(s2.cells);
  const g = s2.backward(target, { blobs: w.blobs, dt: w.dt, steps, checkpointEvery: 2 });
  s2.destroy();
  console.log(JSON.stringify({ backward: g }));
}
