// @ts-check
'use strict';
// Drop-release gradient and fit through the JavaScript host: a 32x48x32 FP64
// run of argv[2] steps, the target dye of one release, then the release
// gradient at a guess and argv[3] iterations of the fit from it.  Prints the
// release blobs, the gradient and the fit as JSON for the C ABI side.
const path = require('path');
const root = path.join(__dirname, '..', '..', 'drop-of-blood_amd', 'js');
const { MacSolver, releaseBlobs } = require(path.join(root, 'MacSolver.js'));

const steps = Number(process.argv[2] || 6);
const iters = Number(process.argv[3] || 2);
const dt = 5e-4;
const s = new MacSolver({ nx: 32, ny: 48, nz: 32, dx: 1e-3, levels: 3, cycles: 8, precision: 2 });
const map = s.releaseMap({ velScale: 2.5 });
const truth = [{ x: 0.05, z: -0.03, radius: 0.09, startY: 0.495, velY: -0.075 }];
const guess = [{ x: 0.08, z: -0.05, radius: 0.1, startY: 0.495, velY: -0.15 }];
s.initBlobs(releaseBlobs(map, truth));
s.step(dt, steps);
const target = s.read('DYE', new Float64Array(s.cells));
const grad = s.backwardRelease(target, { releases: guess, dt, steps, checkpointEvery: 3, map });
const fit = s.fitRelease(target, { releases: guess, dt, steps, map, maxIters: iters,
  checkpointEvery: 3 });
let bad = null;
try { s.backwardRelease(target, { releases: [{ x: 0 }], dt, steps, map }); } catch (e) {
  bad = String(e.message);
}
s.destroy();
console.log(JSON.stringify({ map, blobs: releaseBlobs(map, guess), grad, fit, bad }));
