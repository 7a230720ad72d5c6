// @ts-check
'use strict';
// Runs the JavaScript FP64 oracle on a seeded workload and writes its fields
// for the Python side of the test: argv = workload name, steps, output
// directory [, "stages"].  Writes u, v, w, p, dye (Float64, cell order) to
// <dir>/fields.bin and prints the divergence norms as JSON.
const fs = require('fs');
const path = require('path');
const { MacOracle } = require(path.join(__dirname, '..', '..', 'oracle', 'js', 'mac3d_oracle.js'));
const { workload } = require(path.join(__dirname, '..', '..', 'drop-of-blood_amd', 'js',
  'scenarios.js'));

const name = process.argv[2] || 'tiny';
const steps = Number(process.argv[3] || 2);
const dir = process.argv[4];
const w = workload(name, 1);
const o = new MacOracle({ nx: w.nx, ny: w.ny, nz: w.nz, dx: w.dx, levels: w.levels,
  nu1: w.nu1, nu2: w.nu2, cycles: w.cycles });
o.initBlobs(w.blobs);
const t0 = process.hrtime.bigint();
o.step(w.dt, steps);
const secs = Number(process.hrtime.bigint() - t0) * 1e-9;
const parts = [o.u, o.v, o.w, o.p0, o.dye];
const all = new Float64Array(parts.reduce((n, a) => n + a.length, 0));
let off = 0;
for (const a of parts) { all.set(a, off); off += a.length; }
fs.writeFileSync(path.join(dir, 'fields.bin'), Buffer.from(all.buffer));
console.log(JSON.stringify(Object.assign({ steps, seconds: secs }, o.divNorms())));
