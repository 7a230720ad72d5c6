// @ts-check
'use strict';
// Checkpoint round trip through the JavaScript host: the tiny
// workload (mixed precision when argv[3] is "mixed") steps 3 times, saves,
// steps 2 more, loads and steps the same 2 again.  Prints whether the two
// results are identical and writes the checkpoint and the final fields to
// argv[2] (a directory) for the C ABI side of the test.
const fs = require('fs');
const path = require('path');
const root = path.join(__dirname, '..', '..', 'drop-of-blood_amd', 'js');
const { MacSolver } = require(path.join(root, 'MacSolver.js'));
const { workload } = require(path.join(root, 'scenarios.js'));

const dir = process.argv[2];
const precision = process.argv[3] === 'mixed' ? 1 : 0;
const w = workload('tiny', 1);
const s = new MacSolver({ nx: w.nx, ny: w.ny, nz: w.nz, dx: w.dx, levels: w.levels,
  cycles: w.cycles, precision });
s.initBlobs(w.blobs);
s.step(w.dt, 3);
const ck = s.saveCheckpoint();
const snap = () => {
  s.sync();
  return [s.u, s.v, s.w, s.p, s.dye].map((a) => Float32Array.from(a));
};
s.step(w.dt, 2);
const a = snap();
s.loadCheckpoint(ck);
s.step(w.dt, 2);
const b = snap();
let same = true;
for (let f = 0; f < a.length; f++)
  for (let q = 0; q < a[f].length; q++)
    if (a[f][q] !== b[f][q] && !(Number.isNaN(a[f][q]) && Number.isNaN(b[f][q]))) same = false;
let bad = null;
try { s.loadCheckpoint(ck.subarray(0, 40)); } catch (e) { bad = String(e.message); }
fs.writeFileSync(path.join(dir, 'ck.bin'), Buffer.from(ck.buffer, ck.byteOffset, ck.byteLength));
const all = new Float32Array(b.reduce((n, x) => n + x.length, 0));
let o = 0;
for (const x of b) { all.set(x, o); o += x.length; }
fs.writeFileSync(path.join(dir, 'fields.bin'), Buffer.from(all.buffer));
s.destroy();
console.log(JSON.stringify({ identical: same, bytes: ck.length, badLoad: bad }));
