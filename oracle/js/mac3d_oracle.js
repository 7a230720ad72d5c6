// @ts-check
'use strict';
/**
 * FP64 JavaScript transcription of the MAC step's workload path, in the
 * reference's own language family (its engine is TypeScript on Node; this is
 * `@ts-check`ed JavaScript for Node 12, which has no TypeScript compiler
 * here).  Test infrastructure only: the tests hold it to the C oracle
 * (tests/test_js_oracle.py) and bench.py times it as the CPU path on the
 * box's host cores.  Nothing in the library or its hosts loads it.
 *
 * It restates, cell by cell, the kernels the workload runs
 * (src/shaders/fluid.ts, src/shaders/advect.ts) in the dispatch order of
 * BloodDropEngine.encodePhysicsStep (src/engine/BloodDropEngine.ts:1916-2069)
 * and pass B's dye transport (:2096-2108), with the workload's departures:
 * SI Casson law and linear Boussinesq buoyancy, no-slip walls and a free-slip
 * flat lid (p_top = 0), nu1/nu2 sweeps on every multigrid level with the
 * level-consistent lid coefficient, and an exact coarsest solve (here by
 * Jacobi-preconditioned conjugate gradients).  Layout: x fastest,
 * (k*ny + j)*nx + i, faces stored at +1/2.  FP64 throughout.
 */

/**
 * @typedef {{nx:number, ny:number, nz:number, dx:number, levels?:number,
 *   nu1?:number, nu2?:number, cycles?:number, rho_w?:number, mu_w?:number,
 *   rho_b?:number, tau_y?:number, mu_inf?:number, g?:number, gamma_reg?:number,
 *   shear_eps2?:number, alpha_max?:number, vel_damp?:number,
 *   maccormack?:number}} OracleParams
 * @typedef {{cx:number, cy:number, cz:number, radius:number, w0:number}} Blob
 */

const DEFAULTS = {
  levels: 3, nu1: 2, nu2: 2, cycles: 8, rho_w: 1000.0, mu_w: 1e-3, rho_b: 1060.0,
  tau_y: 0.005, mu_inf: 0.0035, g: 9.81, gamma_reg: 0.40, shear_eps2: 1e-6,
  alpha_max: 1.0 / 6.0, vel_damp: 1.0, maccormack: 0.45 * 0.995,
};

/** @param {number} x @param {number} lo @param {number} hi */
function clamp(x, lo, hi) {
  return x < lo ? lo : (x > hi ? hi : x);
}

/** WGSL mix @param {number} a @param {number} b @param {number} t */
function mix(a, b, t) {
  return a * (1.0 - t) + b * t;
}

/** WGSL smoothstep @param {number} e0 @param {number} e1 @param {number} x */
function smoothstep(e0, e1, x) {
  const t = clamp((x - e0) / (e1 - e0), 0.0, 1.0);
  return t * t * (3.0 - 2.0 * t);
}

/** one level's sizes */
class Grid {
  /** @param {number} nx @param {number} ny @param {number} nz */
  constructor(nx, ny, nz) {
    this.nx = nx;
    this.ny = ny;
    this.nz = nz;
    this.n = nx * ny * nz;
  }
}

/**
 * textureSampleLevel with a linear clamp-to-edge sampler (the sampler of
 * BloodDropEngine.ts:497-505, read by sampleMacU/V/W and traceBackUVW,
 * fluid.ts:154-174, and CONC_PREDICT, advect.ts:52-69): texel coordinate
 * uvw N - 1/2, floor and fraction, corner indices clamped to the grid, then
 * x, y and z linear blends.
 * @param {Float64Array} f @param {Grid} d
 * @param {number} ux @param {number} uy @param {number} uz
 */
function sample(f, d, ux, uy, uz) {
  const nx = d.nx, ny = d.ny, nz = d.nz;
  const tx = ux * nx - 0.5, ty = uy * ny - 0.5, tz = uz * nz - 0.5;
  const bx = Math.floor(tx), by = Math.floor(ty), bz = Math.floor(tz);
  const fx = tx - bx, fy = ty - by, fz = tz - bz;
  const i0 = clamp(bx, 0, nx - 1), i1 = clamp(bx + 1, 0, nx - 1);
  const j0 = clamp(by, 0, ny - 1), j1 = clamp(by + 1, 0, ny - 1);
  const k0 = clamp(bz, 0, nz - 1), k1 = clamp(bz + 1, 0, nz - 1);
  const r00 = (k0 * ny + j0) * nx, r10 = (k0 * ny + j1) * nx;
  const r01 = (k1 * ny + j0) * nx, r11 = (k1 * ny + j1) * nx;
  const x00 = mix(f[r00 + i0], f[r00 + i1], fx), x10 = mix(f[r10 + i0], f[r10 + i1], fx);
  const x01 = mix(f[r01 + i0], f[r01 + i1], fx), x11 = mix(f[r11 + i0], f[r11 + i1], fx);
  return mix(mix(x00, x10, fy), mix(x01, x11, fy), fz);
}

class MacOracle {
  /**
   * @param {OracleParams} params
   * @param {Object<string, SharedArrayBuffer>} [shared] the arrays' storage,
   *   named as sharedLayout lists them (the workers of parallel.js share one
   *   state); fresh private arrays when absent
   */
  constructor(params, shared) {
    /** @type {Required<OracleParams>} */
    this.p = Object.assign({}, DEFAULTS, params);
    const { nx, ny, nz, levels } = this.p;
    if (!(nx > 0 && ny > 0 && nz > 0) || nx % (1 << levels) || ny % (1 << levels) ||
        nz % (1 << levels)) {
      throw new RangeError('grid sizes must be positive and divide by 2^levels');
    }
    /** @type {Grid[]} */
    this.d = [];
    for (let l = 0; l <= levels; ++l) this.d.push(new Grid(nx >> l, ny >> l, nz >> l));
    const arr = (/** @type {string} */ name, /** @type {number} */ len) =>
      shared ? new Float64Array(shared[name]) : new Float64Array(len);
    const n = this.d[0].n;
    this.u = arr('u', n); this.v = arr('v', n); this.w = arr('w', n);
    this.u2 = arr('u2', n); this.v2 = arr('v2', n); this.w2 = arr('w2', n);
    this.p0 = arr('p0', n); this.rhs = arr('rhs', n); this.dye = arr('dye', n);
    this.dpred = arr('dpred', n);
    this.cu = arr('cu', n); this.cv = arr('cv', n); this.cw = arr('cw', n);
    /** multigrid level l: unknown, right-hand side, residual (level 0's unknown is p0) */
    this.mgx = this.d.map((g, l) => arr('mgx' + l, g.n));
    this.mgb = this.d.map((g, l) => arr('mgb' + l, g.n));
    this.mgr = this.d.map((g, l) => arr('mgr' + l, g.n));
    const nc = this.d[levels].n;
    this.cg = [0, 1, 2, 3].map((q) => arr('cg' + q, nc));
    /** this worker's share of each level's planes, and the barrier between stages */
    this.rank = 0;
    this.nw = 1;
    /** @type {{wait: () => void} | null} */
    this.barrier = null;
    /** scratch of trace(): the end point and the start velocity */
    this.bt = new Float64Array(3);
    this.v0 = new Float64Array(3);
    /** scratch of casson(): the neighbour face values per component */
    this.nb = new Float64Array(18);
  }

  /** the planes [k0, k1) of a level of nz planes this worker takes @param {number} nz */
  kr(nz) {
    return [Math.floor(nz * this.rank / this.nw), Math.floor(nz * (this.rank + 1) / this.nw)];
  }

  /** every worker past the previous stage before any starts the next */
  sync() {
    if (this.barrier) this.barrier.wait();
  }

  /**
   * Seeded blob initial condition (the workload's replacement of
   * FLUID_INIT_VEL, fluid.ts:80-121): dye = sum of
   * 1 - smoothstep(r - 1/2, r + 1/2, |x - c|) at cell centres, V on y-faces
   * (x = i + 1/2, y = j + 1, z = k + 1/2) = -sum w0 times the same profile
   * there, U = W = P = 0 (FLUID_CLEAR_R32, fluid.ts:123-138).
   * @param {Blob[]} blobs
   */
  initBlobs(blobs) {
    const { nx, ny, nz } = this.d[0];
    this.u.fill(0); this.w.fill(0); this.p0.fill(0);
    for (let k = 0; k < nz; ++k)
      for (let j = 0; j < ny; ++j)
        for (let i = 0; i < nx; ++i) {
          let dy = 0.0, vv = 0.0;
          for (const b of blobs) {
            const ex = i + 0.5 - b.cx, ey = j + 0.5 - b.cy, ez = k + 0.5 - b.cz;
            dy += 1.0 - smoothstep(b.radius - 0.5, b.radius + 0.5, Math.sqrt(ex * ex + ey * ey + ez * ez));
            const fy = j + 1.0 - b.cy;
            vv -= b.w0 * (1.0 - smoothstep(b.radius - 0.5, b.radius + 0.5,
                                           Math.sqrt(ex * ex + fy * fy + ez * ez)));
          }
          const c = (k * ny + j) * nx + i;
          this.dye[c] = dy;
          this.v[c] = vv;
        }
  }

  /**
   * FLUID_MAC_TO_CENTER (fluid.ts:917-943): centre velocity = mean of the
   * two faces of each axis, solid wall faces read as 0.
   */
  faceToCentre() {
    const { nx, ny, nz } = this.d[0];
    const sy = nx, sz = nx * ny, u = this.u, v = this.v, w = this.w;
    const [k0, k1] = this.kr(nz);
    for (let k = k0; k < k1; ++k)
      for (let j = 0; j < ny; ++j)
        for (let i = 0; i < nx; ++i) {
          const c = (k * ny + j) * nx + i;
          const uR = i >= nx - 1 ? 0.0 : u[c], uL = i === 0 ? 0.0 : u[c - 1];
          const vU = v[c], vD = j === 0 ? 0.0 : v[c - sy];
          const wF = k >= nz - 1 ? 0.0 : w[c], wB = k === 0 ? 0.0 : w[c - sz];
          this.cu[c] = 0.5 * (uL + uR);
          this.cv[c] = 0.5 * (vD + vU);
          this.cw[c] = 0.5 * (wB + wF);
        }
  }

  /**
   * traceBackUVW (fluid.ts:169-174): RK2 midpoint backtrace on the centred
   * velocity in normalised coordinates (ux, uy, uz); the start (when
   * clampStart), the midpoint and the end clamped to [lo, hi] (0.002/0.998
   * for the faces, 0.001/0.999 for the dye).  The end point goes to
   * this.bt[0..2] and the start velocity to this.v0[0..2].
   * @param {number} dt @param {number} ux @param {number} uy @param {number} uz
   * @param {number} lo @param {number} hi @param {boolean} clampStart
   */
  trace(dt, ux, uy, uz, lo, hi, clampStart) {
    const d = this.d[0], dx = this.p.dx;
    const px = clampStart ? clamp(ux, lo, hi) : ux, py = clampStart ? clamp(uy, lo, hi) : uy,
      pz = clampStart ? clamp(uz, lo, hi) : uz;
    const v0x = sample(this.cu, d, px, py, pz), v0y = sample(this.cv, d, px, py, pz),
      v0z = sample(this.cw, d, px, py, pz);
    const mx = clamp(ux - 0.5 * (v0x * dt) / (d.nx * dx), lo, hi),
      my = clamp(uy - 0.5 * (v0y * dt) / (d.ny * dx), lo, hi),
      mz = clamp(uz - 0.5 * (v0z * dt) / (d.nz * dx), lo, hi);
    const vmx = sample(this.cu, d, mx, my, mz), vmy = sample(this.cv, d, mx, my, mz),
      vmz = sample(this.cw, d, mx, my, mz);
    this.bt[0] = clamp(ux - (vmx * dt) / (d.nx * dx), lo, hi);
    this.bt[1] = clamp(uy - (vmy * dt) / (d.ny * dx), lo, hi);
    this.bt[2] = clamp(uz - (vmz * dt) / (d.nz * dx), lo, hi);
    this.v0[0] = v0x; this.v0[1] = v0y; this.v0[2] = v0z;
  }

  /**
   * FLUID_ADVECT_MAC (fluid.ts:144-192, 323-325): each staggered component
   * traced back from its own face and sampled with sampleMacU/V/W's -1/2
   * texel shift (fluid.ts:154-167), clamped to [0.001, 0.999], times
   * velDamp; the solid wall faces (U at i = nx-1, W at k = nz-1) stay 0.
   * @param {number} dt
   */
  advect(dt) {
    const d = this.d[0], { nx, ny, nz } = d, damp = this.p.vel_damp;
    const b = this.bt, L = 0.001, H = 0.999;
    const [k0, k1] = this.kr(nz);
    for (let k = k0; k < k1; ++k)
      for (let j = 0; j < ny; ++j)
        for (let i = 0; i < nx; ++i) {
          const c = (k * ny + j) * nx + i;
          this.trace(dt, (i + 1.0) / nx, (j + 0.5) / ny, (k + 0.5) / nz, 0.002, 0.998, true);
          let vx = sample(this.u, d, clamp(b[0] - 0.5 / nx, L, H), clamp(b[1], L, H),
                          clamp(b[2], L, H)) * damp;
          this.trace(dt, (i + 0.5) / nx, (j + 1.0) / ny, (k + 0.5) / nz, 0.002, 0.998, true);
          const vy = sample(this.v, d, clamp(b[0], L, H), clamp(b[1] - 0.5 / ny, L, H),
                            clamp(b[2], L, H)) * damp;
          this.trace(dt, (i + 0.5) / nx, (j + 0.5) / ny, (k + 1.0) / nz, 0.002, 0.998, true);
          let vz = sample(this.w, d, clamp(b[0], L, H), clamp(b[1], L, H),
                          clamp(b[2] - 0.5 / nz, L, H)) * damp;
          if (i >= nx - 1) vx = 0.0;
          if (k >= nz - 1) vz = 0.0;
          this.u2[c] = vx; this.v2[c] = vy; this.w2[c] = vz;
        }
    this.sync();
    this.swapVel();
  }

  swapVel() {
    let t = this.u; this.u = this.u2; this.u2 = t;
    t = this.v; this.v = this.v2; this.v2 = t;
    t = this.w; this.w = this.w2; this.w2 = t;
  }

  /**
   * Neighbour face values of the viscous stencil with the workload's walls
   * (the reference clamps to the edge, loadMac fluid.ts:387-392): no-slip
   * solid walls (tangential ghost = -value, normal wall face = 0) and a
   * free-slip lid (ghost = value).
   * @param {number} comp 0 U, 1 V, 2 W
   * @param {number} i @param {number} j @param {number} k
   */
  face(comp, i, j, k) {
    const { nx, ny, nz } = this.d[0];
    if (comp === 0) {
      const f = this.u;
      if (i < 0 || i >= nx) return 0.0;
      if (j < 0) return -f[(k * ny) * nx + i];
      if (j >= ny) return f[(k * ny + ny - 1) * nx + i];
      if (k < 0) return -f[j * nx + i];
      if (k >= nz) return -f[((nz - 1) * ny + j) * nx + i];
      return f[(k * ny + j) * nx + i];
    }
    if (comp === 1) {
      const f = this.v;
      if (j < 0) return 0.0;
      if (j >= ny) return f[(k * ny + ny - 1) * nx + i];
      if (i < 0) return -f[(k * ny + j) * nx];
      if (i >= nx) return -f[(k * ny + j) * nx + nx - 1];
      if (k < 0) return -f[j * nx + i];
      if (k >= nz) return -f[((nz - 1) * ny + j) * nx + i];
      return f[(k * ny + j) * nx + i];
    }
    const f = this.w;
    if (k < 0 || k >= nz) return 0.0;
    if (i < 0) return -f[(k * ny + j) * nx];
    if (i >= nx) return -f[(k * ny + j) * nx + nx - 1];
    if (j < 0) return -f[(k * ny) * nx + i];
    if (j >= ny) return f[(k * ny + ny - 1) * nx + i];
    return f[(k * ny + j) * nx + i];
  }

  /**
   * FLUID_CASSON_FORCES (fluid.ts:377-511) parts (1) and (3) with the
   * workload's SI law: strain proxy |+-1 differences| / (2h) with
   * shear_eps2 (fluid.ts:451-454), sqrt(mu_inf(c)) + sqrt(tau_y c^2 /
   * max(gamma + reg, reg)) squared (fluid.ts:458-462) over rho(c), times
   * dt/h^2, clamped to [0, alpha_max]; FTCS update mac += alpha (sum6 - 6
   * mac) (fluid.ts:464-465); linear Boussinesq on y-faces from the mean dye
   * of the cell and the one above (fluid.ts:477-486 restated); wall faces 0
   * (fluid.ts:506-507).
   * @param {number} dt
   */
  casson(dt) {
    const P = this.p, { nx, ny, nz } = this.d[0];
    const inv2h = 1.0 / (2.0 * P.dx), dtoh2 = dt / (P.dx * P.dx);
    const buoy = -P.g * (P.rho_b - P.rho_w) / P.rho_w * dt;
    const nb = this.nb;  // per component a: x+, x-, y+, y-, z+, z- at 6 a .. 6 a + 5
    const [k0, k1] = this.kr(nz);
    for (let k = k0; k < k1; ++k)
      for (let j = 0; j < ny; ++j)
        for (let i = 0; i < nx; ++i) {
          const c = (k * ny + j) * nx + i;
          let ss = 0.0;
          for (let a = 0; a < 3; ++a) {
            const o = 6 * a;
            nb[o] = this.face(a, i + 1, j, k); nb[o + 1] = this.face(a, i - 1, j, k);
            nb[o + 2] = this.face(a, i, j + 1, k); nb[o + 3] = this.face(a, i, j - 1, k);
            nb[o + 4] = this.face(a, i, j, k + 1); nb[o + 5] = this.face(a, i, j, k - 1);
            const du = nb[o] - nb[o + 1], dv = nb[o + 2] - nb[o + 3], dw = nb[o + 4] - nb[o + 5];
            ss += du * du + dv * dv + dw * dw;
          }
          const gam = inv2h * Math.sqrt(ss + P.shear_eps2);
          const cc = clamp(this.dye[c], 0.0, 1.0);
          const sqMu = Math.sqrt(P.mu_w + cc * (P.mu_inf - P.mu_w));
          const yterm = Math.sqrt(P.tau_y * cc * cc / Math.max(gam + P.gamma_reg, P.gamma_reg));
          const muEff = (sqMu + yterm) * (sqMu + yterm);
          const rho = P.rho_w + cc * (P.rho_b - P.rho_w);
          const alpha = clamp(muEff / rho * dtoh2, 0.0, P.alpha_max);
          const lap = (/** @type {number} */ a, /** @type {number} */ m) =>
            (nb[6 * a] + nb[6 * a + 1] + nb[6 * a + 2] + nb[6 * a + 3] + nb[6 * a + 4] +
             nb[6 * a + 5]) - 6.0 * m;
          const mu = this.u[c], mv = this.v[c], mw = this.w[c];
          let ou = mu + alpha * lap(0, mu);
          let ov = mv + alpha * lap(1, mv);
          let ow = mw + alpha * lap(2, mw);
          const jp = j + 1 < ny ? j + 1 : ny - 1;
          ov += buoy * (0.5 * (this.dye[c] + this.dye[(k * ny + jp) * nx + i]));
          if (i >= nx - 1) ou = 0.0;
          if (k >= nz - 1) ow = 0.0;
          this.u2[c] = ou; this.v2[c] = ov; this.w2[c] = ow;
        }
    this.sync();
    this.swapVel();
  }

  /**
   * FLUID_MAC_RHS (fluid.ts:517-555): MAC divergence with the solid wall
   * faces 0 and the flat lid's p_top = 0 (so no top-row term).
   * @param {Float64Array} out
   */
  divergence(out) {
    const { nx, ny, nz } = this.d[0], sy = nx, sz = nx * ny, u = this.u, v = this.v, w = this.w;
    const [k0, k1] = this.kr(nz);
    for (let k = k0; k < k1; ++k)
      for (let j = 0; j < ny; ++j)
        for (let i = 0; i < nx; ++i) {
          const c = (k * ny + j) * nx + i;
          const uR = i >= nx - 1 ? 0.0 : u[c], uL = i === 0 ? 0.0 : u[c - 1];
          const vT = v[c], vB = j === 0 ? 0.0 : v[c - sy];
          const wF = k >= nz - 1 ? 0.0 : w[c], wB = k === 0 ? 0.0 : w[c - sz];
          out[c] = (uR - uL) + (vT - vB) + (wF - wB);
        }
  }

  /** level l's lid coefficient: 1 on level 0, 2^(l+1)/(2^l + 1) below @param {number} l */
  static topc(l) {
    const s = 1 << l;
    return 2.0 * s / (s + 1.0);
  }

  /**
   * evalMgStencil (fluid.ts:561-603) at (i, j, k): returns the neighbour sum
   * and stores the diagonal in this._dg (Neumann walls drop the neighbour and
   * its diagonal term; the top face is always in the diagonal, Dirichlet 0
   * above the top row, with the level's lid coefficient).
   * @param {Float64Array} x @param {Grid} d @param {number} topc
   * @param {number} i @param {number} j @param {number} k
   */
  stencil(x, d, topc, i, j, k) {
    const nx = d.nx, ny = d.ny, nz = d.nz, c = (k * ny + j) * nx + i;
    let sn = 0.0, dg = j < ny - 1 ? 1.0 : topc;
    if (i > 0) { sn += x[c - 1]; dg += 1.0; }
    if (i < nx - 1) { sn += x[c + 1]; dg += 1.0; }
    if (j > 0) { sn += x[c - nx]; dg += 1.0; }
    if (j < ny - 1) sn += x[c + nx];
    if (k > 0) { sn += x[c - nx * ny]; dg += 1.0; }
    if (k < nz - 1) { sn += x[c + nx * ny]; dg += 1.0; }
    this._dg = dg;
    return sn;
  }

  /**
   * MG_RBGS_RED / MG_RBGS_BLACK (fluid.ts:645-681), sweeps times: colour
   * (i + j + k) & 1 = 0 then 1, x = (sum - h^2 b) / diag in place (a colour
   * reads only the other colour, so in place equals the reference's
   * ping-pong).
   * @param {Float64Array} x @param {Float64Array} b @param {number} l @param {number} sweeps
   */
  smooth(x, b, l, sweeps) {
    const d = this.d[l], { nx, ny, nz } = d, h2 = (1 << l) * (1 << l), topc = MacOracle.topc(l);
    const [k0, k1] = this.kr(nz);
    for (let s = 0; s < sweeps; ++s)
      for (let par = 0; par < 2; ++par) {
        for (let k = k0; k < k1; ++k)
          for (let j = 0; j < ny; ++j)
            for (let i = (par + j + k) & 1; i < nx; i += 2) {
              const sn = this.stencil(x, d, topc, i, j, k);
              const c = (k * ny + j) * nx + i;
              x[c] = (sn - h2 * b[c]) / this._dg;
            }
        this.sync();  // the next colour reads this one's planes
      }
  }

  /**
   * MG_RESIDUAL (fluid.ts:744-770) then MG_RESTRICT (fluid.ts:775-800):
   * r = b - (sum - diag x)/h^2 on level l, its 2x2x2 means into level l+1's
   * right-hand side.
   * @param {Float64Array} x @param {Float64Array} b @param {number} l
   */
  residualRestrict(x, b, l) {
    const d = this.d[l], { nx, ny, nz } = d, h2 = (1 << l) * (1 << l), topc = MacOracle.topc(l);
    const r = this.mgr[l];
    const [k0, k1] = this.kr(nz);
    for (let k = k0; k < k1; ++k)
      for (let j = 0; j < ny; ++j)
        for (let i = 0; i < nx; ++i) {
          const sn = this.stencil(x, d, topc, i, j, k), c = (k * ny + j) * nx + i;
          r[c] = b[c] - (sn - this._dg * x[c]) / h2;
        }
    this.sync();
    const dc = this.d[l + 1], bc = this.mgb[l + 1];
    const [c0, c1] = this.kr(dc.nz);
    for (let k = c0; k < c1; ++k)
      for (let j = 0; j < dc.ny; ++j)
        for (let i = 0; i < dc.nx; ++i) {
          let s = 0.0;
          for (let dz = 0; dz < 2; ++dz)
            for (let dy = 0; dy < 2; ++dy)
              for (let dx = 0; dx < 2; ++dx)
                s += r[((2 * k + dz) * ny + 2 * j + dy) * nx + 2 * i + dx];
          bc[(k * dc.ny + j) * dc.nx + i] = s * 0.125;
        }
    this.sync();
  }

  /**
   * MG_PROLONG_ADD (fluid.ts:806-855): x_l += cell-centred trilinear
   * interpolation of level l+1's correction, coarse indices clamped at the
   * walls and 0 above the top row (loadCoarseErr, :814-821).
   * @param {Float64Array} xf @param {number} l
   */
  prolong(xf, l) {
    const df = this.d[l], dc = this.d[l + 1], ec = this.mgx[l + 1];
    const at = (/** @type {number} */ i, /** @type {number} */ j, /** @type {number} */ k) => {
      if (j >= dc.ny) return 0.0;
      return ec[(clamp(k, 0, dc.nz - 1) * dc.ny + clamp(j, 0, dc.ny - 1)) * dc.nx +
                clamp(i, 0, dc.nx - 1)];
    };
    const [k0, k1] = this.kr(df.nz);
    for (let k = k0; k < k1; ++k)
      for (let j = 0; j < df.ny; ++j)
        for (let i = 0; i < df.nx; ++i) {
          const tx = (i + 0.5) / df.nx * dc.nx - 0.5, ty = (j + 0.5) / df.ny * dc.ny - 0.5,
            tz = (k + 0.5) / df.nz * dc.nz - 0.5;
          const bx = Math.floor(tx), by = Math.floor(ty), bz = Math.floor(tz);
          const fx = tx - bx, fy = ty - by, fz = tz - bz;
          const x00 = mix(at(bx, by, bz), at(bx + 1, by, bz), fx);
          const x10 = mix(at(bx, by + 1, bz), at(bx + 1, by + 1, bz), fx);
          const x01 = mix(at(bx, by, bz + 1), at(bx + 1, by, bz + 1), fx);
          const x11 = mix(at(bx, by + 1, bz + 1), at(bx + 1, by + 1, bz + 1), fx);
          xf[(k * df.ny + j) * df.nx + i] += mix(mix(x00, x10, fy), mix(x01, x11, fy), fz);
        }
    this.sync();
  }

  /**
   * The coarsest level solved exactly (the workload's replacement of
   * MG_COARSE_SOLVE_L3, fluid.ts:687-739, whose six sweeps do not
   * converge): (sum - diag x) = h^2 b, i.e. A x = -h^2 b with A = diag -
   * neighbours (symmetric positive definite with the lid), by
   * Jacobi-preconditioned conjugate gradients to a 1e-14 relative residual.
   */
  coarseSolve() {
    if (this.rank === 0) this.coarseSolveSerial();
    this.sync();
  }

  coarseSolveSerial() {
    const L = this.p.levels, d = this.d[L], n = d.n, h2 = (1 << L) * (1 << L);
    const topc = MacOracle.topc(L), x = this.mgx[L], b = this.mgb[L];
    const [r, z, q, Aq] = this.cg;
    let bn = 0.0;
    for (let c = 0; c < n; ++c) { x[c] = 0.0; r[c] = -h2 * b[c]; bn += r[c] * r[c]; }
    if (bn === 0.0) return;
    const diag = new Float64Array(n);
    for (let k = 0; k < d.nz; ++k)
      for (let j = 0; j < d.ny; ++j)
        for (let i = 0; i < d.nx; ++i) {
          this.stencil(r, d, topc, i, j, k);
          diag[(k * d.ny + j) * d.nx + i] = this._dg;
        }
    let rz = 0.0;
    for (let c = 0; c < n; ++c) { z[c] = r[c] / diag[c]; q[c] = z[c]; rz += r[c] * z[c]; }
    for (let it = 0; it < 20000; ++it) {
      let qAq = 0.0;
      for (let k = 0; k < d.nz; ++k)
        for (let j = 0; j < d.ny; ++j)
          for (let i = 0; i < d.nx; ++i) {
            const c = (k * d.ny + j) * d.nx + i;
            Aq[c] = diag[c] * q[c] - this.stencil(q, d, topc, i, j, k);
            qAq += q[c] * Aq[c];
          }
      const a = rz / qAq;
      let rn = 0.0;
      for (let c = 0; c < n; ++c) { x[c] += a * q[c]; r[c] -= a * Aq[c]; rn += r[c] * r[c]; }
      if (rn <= 1e-28 * bn) break;
      let rz2 = 0.0;
      for (let c = 0; c < n; ++c) { z[c] = r[c] / diag[c]; rz2 += r[c] * z[c]; }
      const beta = rz2 / rz;
      rz = rz2;
      for (let c = 0; c < n; ++c) q[c] = z[c] + beta * q[c];
    }
  }

  /**
   * One V-cycle in the schedule of BloodDropEngine.ts:1942-2059: smooth,
   * residual, restrict down (levels >= 1 from a zero guess, as
   * MG_RBGS_RED_ZERO, fluid.ts:605-641), the coarsest solve, prolong and
   * smooth up.
   */
  vcycle() {
    const L = this.p.levels, P = this.p;
    this.smooth(this.p0, this.rhs, 0, P.nu1);
    this.residualRestrict(this.p0, this.rhs, 0);
    for (let l = 1; l < L; ++l) {
      const g = this.d[l], [k0, k1] = this.kr(g.nz);
      this.mgx[l].fill(0, k0 * g.nx * g.ny, k1 * g.nx * g.ny);
      this.sync();
      this.smooth(this.mgx[l], this.mgb[l], l, P.nu1);
      this.residualRestrict(this.mgx[l], this.mgb[l], l);
    }
    this.coarseSolve();
    for (let l = L - 1; l >= 1; --l) {
      this.prolong(this.mgx[l], l);
      this.smooth(this.mgx[l], this.mgb[l], l, P.nu2);
    }
    this.prolong(this.p0, 0);
    this.smooth(this.p0, this.rhs, 0, P.nu2);
  }

  /**
   * FLUID_MAC_PROJECT (fluid.ts:863-911): faces -= p(+1) - p, the top face
   * against the flat lid's p_top = 0, solid wall faces 0.  (The +-2 clamp,
   * fluid.ts:908, is in lattice units and not part of the workload.)
   */
  project() {
    const { nx, ny, nz } = this.d[0], p = this.p0;
    const [k0, k1] = this.kr(nz);
    for (let k = k0; k < k1; ++k)
      for (let j = 0; j < ny; ++j)
        for (let i = 0; i < nx; ++i) {
          const c = (k * ny + j) * nx + i, pc = p[c];
          if (i < nx - 1) this.u[c] -= p[c + 1] - pc; else this.u[c] = 0.0;
          if (j < ny - 1) this.v[c] -= p[c + nx] - pc; else this.v[c] -= 0.0 - pc;
          if (k < nz - 1) this.w[c] -= p[c + nx * ny] - pc; else this.w[c] = 0.0;
        }
  }

  /**
   * CONC_PREDICT (advect.ts:28-72): RK2 backtrace from the cell centre on
   * the centred velocity, [0.001, 0.999], and the dye sampled there.
   * @param {number} dt
   */
  dyePredict(dt) {
    const d = this.d[0], { nx, ny, nz } = d, b = this.bt;
    const [k0, k1] = this.kr(nz);
    for (let k = k0; k < k1; ++k)
      for (let j = 0; j < ny; ++j)
        for (let i = 0; i < nx; ++i) {
          this.trace(dt, (i + 0.5) / nx, (j + 0.5) / ny, (k + 0.5) / nz, 0.001, 0.999, false);
          this.dpred[(k * ny + j) * nx + i] = sample(this.dye, d, b[0], b[1], b[2]);
        }
  }

  /**
   * CONC_RESOLVE's MacCormack core (advect.ts:196-222): the forward estimate
   * of the predictor at uvw + v0 dt, pred + strength (orig - estimate),
   * clamped to the range of the eight dye texels around the backtrace point.
   * @param {number} dt
   */
  dyeResolve(dt) {
    const d = this.d[0], { nx, ny, nz } = d, dx = this.p.dx;
    const mc = this.p.maccormack, out = this.u2, b = this.bt, v0 = this.v0;
    const [k0, k1] = this.kr(nz);
    for (let k = k0; k < k1; ++k)
      for (let j = 0; j < ny; ++j)
        for (let i = 0; i < nx; ++i) {
          const c = (k * ny + j) * nx + i;
          const ux = (i + 0.5) / nx, uy = (j + 0.5) / ny, uz = (k + 0.5) / nz;
          this.trace(dt, ux, uy, uz, 0.001, 0.999, false);
          const est = sample(this.dpred, d, clamp(ux + (v0[0] * dt) / (nx * dx), 0.001, 0.999),
                             clamp(uy + (v0[1] * dt) / (ny * dx), 0.001, 0.999),
                             clamp(uz + (v0[2] * dt) / (nz * dx), 0.001, 0.999));
          const corr = this.dpred[c] + mc * (this.dye[c] - est);
          const bi = Math.floor(b[0] * nx - 0.5), bj = Math.floor(b[1] * ny - 0.5),
            bk = Math.floor(b[2] * nz - 0.5);
          let lo = Infinity, hi = -Infinity;
          for (let dz = 0; dz < 2; ++dz)
            for (let dy = 0; dy < 2; ++dy)
              for (let ddx = 0; ddx < 2; ++ddx) {
                const sv = this.dye[(clamp(bk + dz, 0, nz - 1) * ny + clamp(bj + dy, 0, ny - 1)) * nx +
                                    clamp(bi + ddx, 0, nx - 1)];
                if (sv < lo) lo = sv;
                if (sv > hi) hi = sv;
              }
          out[c] = clamp(corr, lo, hi);
        }
    this.sync();  // (the bounds read the old dye of other planes)
    const pl = nx * ny;
    this.dye.set(out.subarray(k0 * pl, k1 * pl), k0 * pl);
  }

  /**
   * nsteps steps in the order of encodePhysicsStep and pass B: centre
   * velocity of the current faces, advect, Casson + buoyancy, divergence,
   * the V-cycles (warm-started from the previous pressure), project, centre
   * velocity, dye predict and resolve.
   * @param {number} dt @param {number} nsteps
   */
  step(dt, nsteps) {
    if (!Number.isFinite(dt) || !(dt > 0)) throw new RangeError('dt must be finite and > 0');
    for (let s = 0; s < nsteps; ++s) {
      this.faceToCentre();
      this.sync();
      this.advect(dt);
      this.casson(dt);
      this.divergence(this.rhs);
      this.sync();
      for (let c = 0; c < this.p.cycles; ++c) this.vcycle();
      this.project();
      this.sync();
      this.faceToCentre();
      this.sync();
      this.dyePredict(dt);
      this.sync();
      this.dyeResolve(dt);
      this.sync();
    }
  }

  /** max |D| and RMS D of the current faces */
  divNorms() {
    const D = this.mgr[0];
    this.divergence(D);
    let mx = 0.0, ss = 0.0;
    for (let c = 0; c < D.length; ++c) { mx = Math.max(mx, Math.abs(D[c])); ss += D[c] * D[c]; }
    return { maxDiv: mx, l2Div: Math.sqrt(ss / D.length) };
  }
}

module.exports = { MacOracle, sample, smoothstep };
