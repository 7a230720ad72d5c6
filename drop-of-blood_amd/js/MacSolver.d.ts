// Type declarations of the MAC step host (js/MacSolver.js) and its N-API addon.
export interface MacConfig {
  nx?: number; ny?: number; nz?: number;
  levels?: number; nu1?: number; nu2?: number; cycles?: number;
  precision?: 0 | 1 | 2; ir_max_iters?: number; device?: number;
  dx?: number; rho_w?: number; mu_w?: number; rho_b?: number;
  tau_y?: number; mu_inf?: number; g?: number; gamma_reg?: number;
  shear_eps2?: number; alpha_max?: number; vel_damp?: number;
  maccormack?: number; ir_tol?: number; extrapolate?: -1 | 0 | 1 | 2; sor_omega?: number;
}
export interface Blob { cx: number; cy: number; cz: number; radius: number; w0: number; }
export interface WaveParams {
  n?: number; halfExtent?: number; halfXZ?: number; depth?: number;
  surfaceY?: number; marangoni?: number; scene?: 0 | 1;
  vortConf?: number; simTime?: number; schlieren?: number; hemolysis?: number;
  physics?: 0 | 1; cassonScale?: number; buoyancy?: number;
}
export interface Drop {
  x: number; y: number; z: number; radius: number; velY: number; osc?: number;
  drain: number; seed: number;
}
export interface StepStats {
  steps: number; vcycles: number; irIters: number; converged: number;
  maxDiv: number; l2Div: number; maxDivRate: number; relRes: number; ms: number;
}
export declare const FIELD: Readonly<{ U: 0; V: 1; W: 2; P: 3; DYE: 4; CU: 5; CV: 6; CW: 7;
  RHS: 8; MGX: 9; MGB: 10; MGR: 11; DYE_PRED: 12; WAVE: 13; PTOP: 14; CURL: 15; CONC: 16;
  CONC_PRED: 17 }>;
export declare const PRECISION: Readonly<{ fp32: 0; mixed: 1; fp64: 2 }>;
export declare function version(): string;
/** MultiDropSystem.spawnAerialDrop's parameters (scene metres, m/s) */
export interface Release { x: number; z: number; radius: number; startY: number; velY: number; }
export interface ReleaseMap {
  originX: number; originY: number; originZ: number; cell: number; surfaceY: number;
  gravity: number; velScale: number; submerge: number; maxOffset: number; rMin: number;
  rMax: number;
}
export declare function releaseBlobs(map: ReleaseMap, releases: Release[]): Blob[];
export declare class MacSolver {
  constructor(config?: MacConfig);
  readonly config: Required<MacConfig>;
  readonly cells: number;
  readonly u: Float32Array; readonly v: Float32Array; readonly w: Float32Array;
  readonly p: Float32Array; readonly dye: Float32Array;
  destroyed: boolean;
  initBlobs(blobs: Blob[]): void;
  waveEnable(params?: WaveParams | null): void;
  setDrops(drops: Drop[]): void;
  setUniforms(block: Float32Array): void;
  parcelsUpload(parcels: Float32Array): void;
  parcelsEmit(cursor: number, count: number): void;
  concEnable(on?: boolean, refine?: 1 | 2): void;
  readonly concCells: number;
  concRefine: 0 | 1 | 2;
  parcelsStep(dt: number): void;
  parcelsDownload(out: Float32Array): Float32Array;
  splatDownload(which: 0 | 1, out: Uint32Array): Uint32Array;
  step(dt: number, n?: number): StepStats;
  sync(): this;
  write(name: keyof typeof FIELD, data: Float32Array | Float64Array): void;
  read<T extends Float32Array | Float64Array>(name: keyof typeof FIELD, out: T): T;
  backward(target: Float32Array | Float64Array,
           run: { blobs: Blob[]; dt: number; steps: number; checkpointEvery?: number }):
    { loss: number; dRhoB: number; dBlobs: Blob[] };
  releaseMap(overrides?: Partial<ReleaseMap>): ReleaseMap;
  backwardRelease(target: Float32Array | Float64Array,
                  run: { releases: Release[]; dt: number; steps: number;
                         checkpointEvery?: number; map?: ReleaseMap }):
    { loss: number; dRhoB: number; dReleases: Release[]; dBlobs: Blob[] };
  fitRelease(target: Float32Array | Float64Array,
             run: { releases: Release[]; dt: number; steps: number; map?: ReleaseMap;
                    maxIters?: number; checkpointEvery?: number; mask?: number; tol?: number;
                    gtol?: number }):
    { releases: Release[]; iters: number; evals: number; converged: number; loss0: number;
      loss: number };
  divNorms(): { maxDiv: number; l2Div: number };
  saveCheckpoint(): Uint8Array;
  loadCheckpoint(bytes: Uint8Array | ArrayBuffer): void;
  destroy(): void;
}
