// @ts-check
'use strict';
/**
 * Seeded workloads of the MAC step: splitmix64 blob placement, draw for
 * draw the same as mac3d/scenarios.py (checked by tests/test_host.py).
 * Blob positions in cell units, y vertical.
 */

const MASK = (1n << 64n) - 1n;

class SplitMix64 {
  /** @param {number|bigint} seed */
  constructor(seed) {
    this.state = BigInt(seed) & MASK;
  }
  /** @returns {bigint} */
  nextU64() {
    this.state = (this.state + 0x9E3779B97F4A7C15n) & MASK;
    let z = this.state;
    z = ((z ^ (z >> 30n)) * 0xBF58476D1CE4E5B9n) & MASK;
    z = ((z ^ (z >> 27n)) * 0x94D049BB133111EBn) & MASK;
    return z ^ (z >> 31n);
  }
  /** @returns {number} uniform in [0, 1) with 53 bits */
  uniform() {
    return Number(this.nextU64() >> 11n) * (1.0 / 9007199254740992.0);
  }
  /** @param {number} a @param {number} b */
  range(a, b) {
    return a + (b - a) * this.uniform();
  }
}

/**
 * @typedef {{cx:number, cy:number, cz:number, radius:number, w0:number}} Blob
 * @typedef {{name:string, nx:number, ny:number, nz:number, dx:number, dt:number,
 *   steps:number, levels:number, cycles:number, precision:number, blobs:Blob[],
 *   seed:number, nu1:number, nu2:number, ir_tol:number}} Workload
 */

/** @returns {Blob[]} */
function blobs(rng, nx, ny, nz, count, rLo, rHi, stagger, wLo, wHi) {
  const out = [];
  let y0 = null;
  for (let b = 0; b < count; b++) {
    const r = rHi > rLo ? rng.range(rLo, rHi) : rLo;
    const cx = rng.range(r + 2.0, nx - r - 2.0);
    const cz = rng.range(r + 2.0, nz - r - 2.0);
    if (y0 === null) y0 = rng.range(ny * 2.0 / 3.0, ny - r - 2.0);
    const cy = y0 - stagger * b;
    const w0 = wHi > wLo ? rng.range(wLo, wHi) : wLo;
    out.push({ cx, cy, cz, radius: r, w0 });
  }
  return out;
}

/**
 * @param {string} name tiny | ref | mixed | adjoint | large
 * @param {number} seed
 * @returns {Workload}
 */
function workload(name, seed = 1) {
  const rng = new SplitMix64(seed);
  /** @type {Workload} */
  let w;
  const base = (nx, ny, nz, dx, dt, steps, levels, cycles, precision) =>
    ({ name, nx, ny, nz, dx, dt, steps, levels, cycles, precision, blobs: [], seed,
       nu1: 2, nu2: 2, ir_tol: 1e-8 });
  if (name === 'tiny') {
    w = base(32, 48, 32, 1e-3, 5e-4, 20, 3, 8, 0);
    w.blobs = blobs(rng, w.nx, w.ny, w.nz, 1, 4.0, 4.0, 0.0, 0.5, 0.5);
  } else if (name === 'ref' || name === 'mixed') {
    w = base(192, 256, 192, 5e-4, 2.5e-4, 200, 3, 8, name === 'mixed' ? 1 : 0);
    w.blobs = blobs(rng, w.nx, w.ny, w.nz, 3, 6.0, 8.0, 40.0, 0.3, 0.8);
  } else if (name === 'adjoint') {
    w = base(64, 96, 64, 1e-3, 5e-4, 50, 3, 8, 0);
    w.blobs = blobs(rng, w.nx, w.ny, w.nz, 1, 6.0, 6.0, 0.0, 0.5, 0.5);
  } else if (name === 'large') {
    w = base(256, 384, 256, 3.75e-4, 1.875e-4, 200, 5, 8, 1);
    w.blobs = blobs(rng, w.nx, w.ny, w.nz, 5, 6.0, 10.0, 48.0, 0.3, 0.8);
  } else {
    throw new Error(`unknown workload ${name}`);
  }
  return w;
}

module.exports = { SplitMix64, workload };
