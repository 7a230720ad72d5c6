// @ts-check
'use strict';
/**
 * The JavaScript oracle's step on several host threads (Node worker_threads):
 * one MacOracle state in SharedArrayBuffers, each worker taking a contiguous
 * share of every level's z-planes and running the step in lockstep with the
 * others, an Atomics barrier between stages (and between a red-black sweep's
 * two colours), the small coarsest solve on worker 0.  The result equals the
 * single-threaded step's, which visits the same cells in the same order
 * within each plane (tests/test_js_oracle.py).  Test infrastructure: bench.py
 * times it as the CPU path.
 *
 * CLI: node parallel.js <workload> <steps> <threads> [<out fields.bin>]
 *   prints {"threads", "steps", "seconds", "per_step", "seconds_per_step" (the
 *   last step's), "maxDiv", "node"} as JSON.
 */
const path = require('path');
const fs = require('fs');
const { Worker, isMainThread, workerData, parentPort } = require('worker_threads');
const { MacOracle } = require('./mac3d_oracle.js');

/** sense-counting barrier over an Int32Array(2) in shared memory: [arrived, generation] */
class Barrier {
  /** @param {SharedArrayBuffer} sab @param {number} n */
  constructor(sab, n) {
    this.a = new Int32Array(sab);
    this.n = n;
  }
  wait() {
    const a = this.a;
    const gen = Atomics.load(a, 1);
    if (Atomics.add(a, 0, 1) === this.n - 1) {
      Atomics.store(a, 0, 0);
      Atomics.add(a, 1, 1);
      Atomics.notify(a, 1);
    } else {
      while (Atomics.load(a, 1) === gen) Atomics.wait(a, 1, gen);
    }
  }
}

/**
 * Names and lengths of the state's arrays (MacOracle's constructor reads the
 * same names).
 * @param {import('./mac3d_oracle.js').MacOracle} o
 */
function sharedLayout(o) {
  /** @type {[string, number][]} */
  const out = [];
  const n = o.d[0].n;
  for (const name of ['u', 'v', 'w', 'u2', 'v2', 'w2', 'p0', 'rhs', 'dye', 'dpred', 'cu', 'cv', 'cw'])
    out.push([name, n]);
  o.d.forEach((g, l) => {
    out.push(['mgx' + l, g.n]);
    out.push(['mgb' + l, g.n]);
    out.push(['mgr' + l, g.n]);
  });
  for (let q = 0; q < 4; ++q) out.push(['cg' + q, o.d[o.p.levels].n]);
  return out;
}

/**
 * nsteps steps of dt on `threads` workers from the given blobs; resolves to
 * {seconds, oracle} with the shared state viewed by the returned oracle.
 * @param {import('./mac3d_oracle.js').OracleParams} params
 * @param {import('./mac3d_oracle.js').Blob[]} blobs
 * @param {number} dt @param {number} nsteps @param {number} threads
 */
function runParallel(params, blobs, dt, nsteps, threads) {
  const probe = new MacOracle(params);
  /** @type {Object<string, SharedArrayBuffer>} */
  const shared = {};
  for (const [name, len] of sharedLayout(probe)) shared[name] = new SharedArrayBuffer(8 * len);
  const o = new MacOracle(params, shared);
  o.initBlobs(blobs);
  const bar = new SharedArrayBuffer(8);
  return new Promise((resolve, reject) => {
    let done = 0, secs = 0;
    /** @type {number[]} */
    let per = [];
    for (let r = 0; r < threads; ++r) {
      const w = new Worker(__filename, {
        workerData: { params, shared, bar, rank: r, nw: threads, dt, nsteps },
      });
      w.on('message', (m) => {
        secs = Math.max(secs, m.seconds);
        per = per.length ? per.map((x, q) => Math.max(x, m.per[q])) : m.per;
        if (++done === threads) resolve({ seconds: secs, per, oracle: o });
      });
      w.on('error', reject);
    }
  });
}

if (!isMainThread) {
  const { params, shared, bar, rank, nw, dt, nsteps } = workerData;
  const o = new MacOracle(params, shared);
  o.rank = rank;
  o.nw = nw;
  o.barrier = new Barrier(bar, nw);
  o.sync();
  const per = [];
  for (let s = 0; s < nsteps; ++s) {
    const t0 = process.hrtime.bigint();
    o.step(dt, 1);
    per.push(Number(process.hrtime.bigint() - t0) * 1e-9);
  }
  if (parentPort) parentPort.postMessage({ seconds: per.reduce((a, b) => a + b, 0), per });
} else if (require.main === module) {
  const { workload } = require(path.join(__dirname, '..', '..', 'drop-of-blood_amd', 'js',
    'scenarios.js'));
  const w = workload(process.argv[2] || 'tiny', 1);
  const steps = Number(process.argv[3] || 1);
  const threads = Number(process.argv[4] || 1);
  const outFile = process.argv[5];
  const params = { nx: w.nx, ny: w.ny, nz: w.nz, dx: w.dx, levels: w.levels, nu1: w.nu1,
    nu2: w.nu2, cycles: w.cycles };
  runParallel(params, w.blobs, w.dt, steps, threads).then(({ seconds, per, oracle }) => {
    if (outFile) {
      const parts = [oracle.u, oracle.v, oracle.w, oracle.p0, oracle.dye];
      const all = new Float64Array(parts.reduce((n, a) => n + a.length, 0));
      let off = 0;
      for (const a of parts) { all.set(a, off); off += a.length; }
      fs.writeFileSync(outFile, Buffer.from(all.buffer));
    }
    // per step: the slowest worker's time of each step (the first one
    // includes the compiler's warm-up)
    const cpus = require('os').cpus();
    console.log(JSON.stringify({ threads, steps, seconds, per_step: per,
      seconds_per_step: per[per.length - 1], maxDiv: oracle.divNorms().maxDiv,
      node: process.version, cpu_model: cpus.length ? cpus[0].model : 'unknown',
      cpus: cpus.length }));
  });
}

module.exports = { runParallel, Barrier, sharedLayout };
