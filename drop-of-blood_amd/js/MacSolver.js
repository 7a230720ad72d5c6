// @ts-check
'use strict';
/**
 * JavaScript host of the MI355X MAC step (Node N-API over the C ABI), typed
 * by JSDoc under @ts-check and by MacSolver.d.ts.
 *
 * Stands where the reference's WebGPU compute pass stands: the engine used to
 * record the MAC kernels in encodePhysicsStep (src/engine/BloodDropEngine.ts:
 * 1855-2069) and step them per substep from renderFrame / seekToSimTime
 * (:2605, :2142-2179).  A MacSolver keeps the src/sim solver shape: a
 * `step(dt)` method like MultiDropSystem.step (src/sim/dropKinematics.ts:604)
 * and reusable Float32Array field buffers like SimUniformPacker's
 * (src/sim/uniformLayout.ts:52-54).
 *
 * Errors: the addon throws an Error carrying the library's status and
 * message; a non-finite or non-positive dt throws before any state changes.
 */
const path = require('path');

const addon = require(path.join(__dirname, '..', 'lib', 'mac3d.node'));

const FIELD = Object.freeze({ U: 0, V: 1, W: 2, P: 3, DYE: 4, CU: 5, CV: 6, CW: 7, RHS: 8,
  MGX: 9, MGB: 10, MGR: 11, DYE_PRED: 12, WAVE: 13, PTOP: 14, CURL: 15, CONC: 16,
  CONC_PRED: 17 });
const PRECISION = Object.freeze({ fp32: 0, mixed: 1, fp64: 2 });

/**
 * @typedef {{nx?:number, ny?:number, nz?:number, levels?:number, nu1?:number,
 *   nu2?:number, cycles?:number, precision?:number, ir_max_iters?:number,
 *   device?:number, dx?:number, rho_w?:number, mu_w?:number, rho_b?:number,
 *   tau_y?:number, mu_inf?:number, g?:number, gamma_reg?:number,
 *   shear_eps2?:number, alpha_max?:number, vel_damp?:number,
 *   maccormack?:number, ir_tol?:number, extrapolate?:number, sor_omega?:number}} MacConfig
 * @typedef {{x:number, z:number, radius:number, startY:number, velY:number}} Release
 *   spawnAerialDrop(offsetXZ = [x, z], radius, seed, startY, velY), scene metres and m/s
 * @typedef {{originX:number, originY:number, originZ:number, cell:number,
 *   surfaceY:number, gravity:number, velScale:number, submerge:number,
 *   maxOffset:number, rMin:number, rMax:number}} ReleaseMap
 * @typedef {{steps:number, vcycles:number, irIters:number, converged:number,
 *   maxDiv:number, l2Div:number, maxDivRate:number, relRes:number, ms:number}} StepStats
 */

class MacSolver {
  /** @param {MacConfig} [config] */
  constructor(config = {}) {
    /** @type {Required<MacConfig>} */
    this.config = Object.assign(addon.defaultConfig(), config);
    this._h = addon.create(this.config);
    const { nx, ny, nz } = this.config;
    this.cells = nx * ny * nz;
    /** field buffers, refreshed by sync(); x fastest, (k*ny + j)*nx + i */
    this.u = new Float32Array(this.cells);
    this.v = new Float32Array(this.cells);
    this.w = new Float32Array(this.cells);
    this.p = new Float32Array(this.cells);
    this.dye = new Float32Array(this.cells);
    /** the concentration grid per axis (concEnable), 0 while it is off */
    this.concRefine = 0;
    this.destroyed = false;
  }

  _live() {
    if (this.destroyed) throw new Error('MacSolver: destroyed');
    return this._h;
  }

  /** @param {{cx:number, cy:number, cz:number, radius:number, w0:number}[]} blobs */
  initBlobs(blobs) {
    addon.initBlobs(this._live(), blobs);
  }

  /**
   * Free-surface coupling: the wave solver and the lid pressure it feeds
   * (wave.ts:56-388, fluid.ts:16-78).  null disables it.
   * @param {{n?:number, halfExtent?:number, halfXZ?:number, depth?:number,
   *          surfaceY?:number, marangoni?:number, scene?:number, vortConf?:number,
   *          simTime?:number, schlieren?:number} | null} params
   */
  waveEnable(params) {
    addon.waveEnable(this._live(), params === undefined ? {} : params);
  }

  /**
   * Drop slots of the wave forcing (SimUniforms.drops, common.ts:26-43).
   * @param {{x:number, y:number, z:number, radius:number, velY:number, osc?:number,
   *          drain:number, seed:number}[]} drops
   */
  setDrops(drops) {
    addon.setDrops(this._live(), drops);
  }

  /**
   * The engine's packed SimUniforms block (SimUniformPacker.pack's
   * Float32Array(128)): surface height, Marangoni, domain bounds, drop slots.
   * @param {Float32Array} block
   */
  setUniforms(block) {
    addon.setUniforms(this._live(), block);
  }

  /**
   * Lagrangian parcels (RbcParcel, 8 floats each: x, y, z, mass, stretch x, y, z,
   * age ticks + lysis); parcelsStep(dt) advances and splats them.
   * @param {Float32Array} parcels
   */
  parcelsUpload(parcels) {
    addon.parcelsUpload(this._live(), parcels);
  }

  /** One parcel step and splat (MAC3D_STAGE_PARCELS) @param {number} dt */
  parcelsStep(dt) {
    if (!Number.isFinite(dt) || dt < 0) throw new RangeError('dt must be finite and >= 0');
    addon.stage(this._live(), 10, dt);
  }

  /** @param {Float32Array} out 8 floats per parcel to read back */
  parcelsDownload(out) {
    addon.parcelsDownload(this._live(), out);
    return out;
  }

  /** @param {0|1} which intact or lysed @param {Uint32Array} out one per cell of the splat grid (concCells) */
  splatDownload(which, out) {
    addon.splatDownload(this._live(), which, out);
    return out;
  }

  /**
   * Advance n steps of fixed dt (seconds); blocks until the GPU is done.
   * @param {number} dt @param {number} [n]
   * @returns {StepStats}
   */
  step(dt, n = 1) {
    if (!Number.isFinite(dt) || dt <= 0) throw new RangeError('dt must be finite and > 0');
    return addon.step(this._live(), dt, n);
  }

  /**
   * The reference's 4-channel concentration (field CONC, 4 per cell of its
   * grid): each step then runs pass B (parcel emission, advection and splat,
   * CONC_PREDICT, CONC_RESOLVE) after the projection.  refine 1 holds it on
   * the MAC grid, 2 on a grid twice as fine per axis (the reference's
   * CONC_GRID over its VEL_GRID); the parcels splat into that grid.
   * @param {boolean} on @param {1|2} [refine]
   */
  concEnable(on = true, refine = 1) {
    if (refine !== 1 && refine !== 2) throw new RangeError('refine is 1 or 2');
    addon.concEnable(this._live(), on ? refine : 0);
    this.concRefine = on ? refine : 0;
  }

  /** cells of the concentration grid (the splat grid; the MAC grid while it is off) */
  get concCells() {
    const r = Math.max(this.concRefine || 0, 1);
    return this.cells * r * r * r;
  }

  /**
   * The next parcel stage first emits `count` parcels from ring slot
   * `cursor` (the packer's spawnCursor / spawnCount).
   * @param {number} cursor @param {number} count
   */
  parcelsEmit(cursor, count) {
    addon.parcelsEmit(this._live(), cursor, count);
  }

  /** Copy the device fields into the reusable buffers. */
  sync() {
    const h = this._live();
    addon.download(h, FIELD.U, 0, this.u);
    addon.download(h, FIELD.V, 0, this.v);
    addon.download(h, FIELD.W, 0, this.w);
    addon.download(h, FIELD.P, 0, this.p);
    addon.download(h, FIELD.DYE, 0, this.dye);
    return this;
  }

  /**
   * @param {keyof typeof FIELD} name
   * @param {Float32Array|Float64Array} data
   */
  write(name, data) {
    addon.upload(this._live(), FIELD[name], 0, data);
  }

  /**
   * @param {keyof typeof FIELD} name
   * @param {Float32Array|Float64Array} out
   */
  read(name, out) {
    addon.download(this._live(), FIELD[name], 0, out);
    return out;
  }

  /**
   * Gradient of L = 1/2 ||dye_N - target||^2 for a run of `steps` steps of
   * `dt` from the blob initial condition, by the hand-written adjoint with a
   * checkpoint every `checkpointEvery` steps (recomputed in between).
   * Leaves the solver at an unspecified state of that run.
   * @param {Float32Array|Float64Array} target one dye value per cell
   * @param {{blobs:{cx:number, cy:number, cz:number, radius:number, w0:number}[],
   *   dt:number, steps:number, checkpointEvery?:number}} run
   * @returns {{loss:number, dRhoB:number,
   *   dBlobs:{cx:number, cy:number, cz:number, radius:number, w0:number}[]}}
   */
  backward(target, run) {
    if (!Number.isFinite(run.dt) || run.dt <= 0) throw new RangeError('dt must be finite and > 0');
    return addon.adjoint(this._live(), run.blobs, run.dt, run.steps, run.checkpointEvery || 10,
      target);
  }

  /**
   * The release map of this solver's grid (the reference tank's 0.96 m on nx
   * cells, the top face at the surface), with any fields overridden.
   * @param {Partial<ReleaseMap>} [overrides]
   * @returns {ReleaseMap}
   */
  releaseMap(overrides = {}) {
    return Object.assign(addon.releaseDefaultMap(this.config), overrides);
  }

  /**
   * Gradient of L = 1/2 ||dye_N - target||^2 with respect to the parameters
   * of MultiDropSystem.spawnAerialDrop (dropKinematics.ts:354-391) of each
   * release, through the blobs the release map makes of them.
   * @param {Float32Array|Float64Array} target one dye value per cell
   * @param {{releases:Release[], dt:number, steps:number, checkpointEvery?:number,
   *   map?:ReleaseMap}} run
   * @returns {{loss:number, dRhoB:number, dReleases:Release[],
   *   dBlobs:{cx:number, cy:number, cz:number, radius:number, w0:number}[]}}
   */
  backwardRelease(target, run) {
    if (!Number.isFinite(run.dt) || run.dt <= 0) throw new RangeError('dt must be finite and > 0');
    return addon.adjointRelease(this._live(), run.map || this.releaseMap(), run.releases, run.dt,
      run.steps, run.checkpointEvery || 10, target);
  }

  /**
   * Fits the releases to a target dye field from the given guess
   * (quasi-Newton on x, z, radius and velY by default; startY held).
   * @param {Float32Array|Float64Array} target
   * @param {{releases:Release[], dt:number, steps:number, map?:ReleaseMap,
   *   maxIters?:number, checkpointEvery?:number, mask?:number, tol?:number,
   *   gtol?:number}} run
   * @returns {{releases:Release[], iters:number, evals:number, converged:number,
   *   loss0:number, loss:number}}
   */
  fitRelease(target, run) {
    if (!Number.isFinite(run.dt) || run.dt <= 0) throw new RangeError('dt must be finite and > 0');
    return addon.fitRelease(this._live(), run.map || this.releaseMap(), run.releases, run.dt,
      run.steps, target, run);
  }

  divNorms() {
    return addon.divNorms(this._live());
  }

  /**
   * The evolving state (fields, pressure history, wave, parcels) as bytes the
   * caller keeps; loadCheckpoint resumes from them (the reference resumes by
   * replaying from a reset, BloodDropEngine.ts:2142-2179).
   * @returns {Uint8Array}
   */
  saveCheckpoint() {
    return new Uint8Array(addon.checkpointSave(this._live()));
  }

  /** @param {Uint8Array|ArrayBuffer} bytes a checkpoint of a solver of the same config */
  loadCheckpoint(bytes) {
    addon.checkpointLoad(this._live(), bytes);
  }

  /** Idempotent, like the reference engine's destroy() guard. */
  destroy() {
    if (this.destroyed) return;
    addon.destroy(this._h);
    this.destroyed = true;
  }
}

/**
 * The blobs of releases under a map (host arithmetic, no solver needed).
 * @param {ReleaseMap} map
 * @param {Release[]} releases
 * @returns {{cx:number, cy:number, cz:number, radius:number, w0:number}[]}
 */
function releaseBlobs(map, releases) {
  return addon.releaseBlobs(map, releases);
}

module.exports = { MacSolver, FIELD, PRECISION, releaseBlobs, version: addon.version };
