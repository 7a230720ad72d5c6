/*
 * mac3d.h — C ABI of the MI355X staggered-MAC Navier–Stokes step.
 *
 * One context owns every device buffer of one simulation and one HIP stream.
 * All entry points return an int status (MAC3D_OK = 0); on failure the
 * context's last-error string says why (mac3d_last_error).  Host buffers
 * belong to the caller: upload/download copy synchronously and never keep
 * the pointer.  Calls on one context must be serialised by the caller.
 *
 * Grid convention (same as the reference's textures): cell (i,j,k) with x
 * fastest, flat index (k*ny + j)*nx + i, y is the vertical axis.  The face
 * components are stored per cell at the +1/2 face: U(i,j,k) is the x-face
 * between cells i and i+1, V the y-face above cell j, W the z-face in front
 * of cell k (src/shaders/fluid.ts:186-188, FLUID_INIT_VEL fluid.ts:95-97).
 * U at i = nx-1 and W at k = nz-1 are solid walls (always 0); V at j = ny-1
 * is the free-slip lid, left free (fluid.ts:323-325).
 *
 * Reference interface being replaced: the WebGPU compute pass recorded by
 * BloodDropEngine.encodePhysicsStep (src/engine/BloodDropEngine.ts:1855-2069)
 * for the MAC portion (:1916-2069), driven per substep by renderFrame
 * (:2605) and seekToSimTime (:2142-2179).
 */
#ifndef MAC3D_H
#define MAC3D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MAC3D_ABI_VERSION 2

/* status codes */
#define MAC3D_OK 0
#define MAC3D_EINVAL 1   /* bad argument: sizes, NaN/negative dt, unknown id */
#define MAC3D_EHIP 2     /* a HIP runtime call failed */
#define MAC3D_ENOMEM 3   /* device allocation failed */
#define MAC3D_ESTATE 4   /* call not valid in the context's current state */
#define MAC3D_ENONFINITE 5 /* the state went non-finite (NaN or inf): seen by the
                              divergence norms of mac3d_step's stats, or by the mixed
                              mode's residual norms; the state is left as computed, for
                              the caller to restore from a checkpoint */

/* precision modes of the step */
#define MAC3D_PREC_FP32 0  /* plain FP32: fixed V-cycle count per step */
#define MAC3D_PREC_MIXED 1 /* FP32 V-cycles, FP16 level>=2 storage, FP64 IR */
#define MAC3D_PREC_FP64 2  /* FP64 everywhere (gradient checks, references) */

/* host dtypes for upload/download */
#define MAC3D_F32 0
#define MAC3D_F64 1

/* field ids (level 0 unless noted) */
#define MAC3D_FIELD_U 0    /* x-face velocity (m/s) */
#define MAC3D_FIELD_V 1    /* y-face velocity (m/s), vertical */
#define MAC3D_FIELD_W 2    /* z-face velocity (m/s) */
#define MAC3D_FIELD_P 3    /* dimensionless pressure p~ = dt p/(rho h) */
#define MAC3D_FIELD_DYE 4  /* blood dye fraction */
#define MAC3D_FIELD_CU 5   /* cell-centred velocity x */
#define MAC3D_FIELD_CV 6   /* cell-centred velocity y */
#define MAC3D_FIELD_CW 7   /* cell-centred velocity z */
#define MAC3D_FIELD_RHS 8  /* Poisson right-hand side = MAC divergence */
#define MAC3D_FIELD_MGX 9  /* multigrid unknown of a level (level 0 = P) */
#define MAC3D_FIELD_MGB 10 /* multigrid right-hand side of a level */
#define MAC3D_FIELD_MGR 11 /* multigrid residual of a level */
#define MAC3D_FIELD_DYE_PRED 12 /* dye semi-Lagrangian predictor */
#define MAC3D_FIELD_WAVE 13 /* free-surface wave state, n*n texels x 4 (eta (m), d eta/dt
                               (m/s), 9-point Laplacian of eta, surfactant film), texel
                               (ix, iy) at 4*(iy*n + ix); needs mac3d_wave_enable */
#define MAC3D_FIELD_PTOP 14 /* top Dirichlet pressure of the lid (dimensionless like P),
                               nx*nz values at iz*nx + ix; needs mac3d_wave_enable */
#define MAC3D_FIELD_CURL 15 /* scene only: curl of the post-advection centre velocity,
                               4 per cell (omega x, y, z, |omega|; grid units, 1/s
                               times h), FLUID_CURL (fluid.ts:331-369) */
#define MAC3D_FIELD_CONC 16 /* the reference's 4-channel concentration (needs
                               mac3d_conc_enable): 4 per cell of the concentration
                               grid at 4*cell: intact RBC (.r), lysed Hb (.g),
                               oxygenation/shear state (.b), floor bed (.a), the
                               layout of its rgba texture (advect.ts:75-79) */
#define MAC3D_FIELD_CONC_PRED 17 /* the lysed channel's semi-Lagrangian predictor
                                    (CONC_PREDICT's concPredOut .g), one per cell of
                                    the concentration grid */
#define MAC3D_FIELD_COUNT 18

/* single stages of the step, for kernel-level parity tests */
#define MAC3D_STAGE_FACE_TO_CENTRE 0 /* CU,CV,CW <- U,V,W */
#define MAC3D_STAGE_ADVECT 1         /* U,V,W <- advect(U,V,W; CU,CV,CW) */
#define MAC3D_STAGE_CASSON 2         /* U,V,W <- Casson viscosity + buoyancy */
#define MAC3D_STAGE_DIVERGENCE 3     /* RHS <- D(U,V,W) */
#define MAC3D_STAGE_VCYCLE 4         /* P <- one V-cycle on (P, RHS) */
#define MAC3D_STAGE_PROJECT 5        /* U,V,W <- U,V,W - G(P) */
#define MAC3D_STAGE_DYE_PREDICT 6    /* DYE_PRED <- semi-Lagrangian(DYE) */
#define MAC3D_STAGE_DYE_RESOLVE 7    /* DYE <- MacCormack + limiter */
#define MAC3D_STAGE_WAVE 8           /* WAVE <- one free-surface wave step (needs V) */
#define MAC3D_STAGE_TOP_PRESSURE 9   /* PTOP <- top Dirichlet pressure from WAVE and drops */
#define MAC3D_STAGE_PARCELS 10       /* parcels advanced through CU,CV,CW; splat counters <- their
                                        mass (cleared first); needs the coupling and parcels */
#define MAC3D_STAGE_CONC_PREDICT 11  /* DYE_PRED <- CONC_PREDICT of CONC's lysed channel */
#define MAC3D_STAGE_CONC_RESOLVE 12  /* CONC <- CONC_RESOLVE from CONC, DYE_PRED, the splat
                                        counters and CU,CV,CW */
#define MAC3D_STAGE_COUNT 13

/* single multigrid operators on one level, for parity tests */
#define MAC3D_MG_SMOOTH 0   /* count RBGS sweeps on (MGX, MGB) */
#define MAC3D_MG_RESIDUAL 1 /* MGR <- MGB - L(MGX) */
#define MAC3D_MG_RESTRICT 2 /* MGB[l+1] <- R(MGR[l]) */
#define MAC3D_MG_PROLONG 3  /* MGX[l] += P(MGX[l+1]) */
#define MAC3D_MG_COARSE 4   /* MGX[L] <- exact solve of L(x) = MGB[L] */
#define MAC3D_MG_ZERO 5     /* MGX[l] <- 0 */
#define MAC3D_MG_SMOOTH_ZERO 6 /* count RBGS sweeps on MGB from a zero guess (MGX unread;
                                  MG_RBGS_RED_ZERO, src/shaders/fluid.ts:605-641) */
#define MAC3D_MG_SMOOTH2_MARCH 7 /* two RBGS sweeps on (MGX, MGB) by the z-marching two-sweep
                                    pass (level 0's) on any FP32/FP16 level; count 0: from
                                    MGX, 1: from a zero guess.  Coarse levels smooth with a
                                    tile pass, which must equal this one bit for bit */

/* kernel classes timed by mac3d_profile (HIP events on the stream) */
#define MAC3D_KC_FACE_TO_CENTRE 0
#define MAC3D_KC_ADVECT 1
#define MAC3D_KC_CASSON 2
#define MAC3D_KC_DIVERGENCE 3
#define MAC3D_KC_SMOOTH_L0 4       /* one RBGS sweep (both colours) on level 0 */
#define MAC3D_KC_RESTRICT_L0 5     /* level-0 residual and restriction */
#define MAC3D_KC_PROLONG_L0 6      /* prolongation onto level 0 */
#define MAC3D_KC_MG_COARSE_LEVELS 7 /* all multigrid work on levels >= 1 */
#define MAC3D_KC_COARSE_SOLVE 8    /* exact coarsest-level solve */
#define MAC3D_KC_PROJECT 9
#define MAC3D_KC_DYE 10            /* dye predictor + resolve */
#define MAC3D_KC_REFINE64 11       /* mixed: FP64 residual/update */
#define MAC3D_KC_SMOOTH2_L0 12     /* two RBGS sweeps in one pass on level 0 (FP32) */
#define MAC3D_KC_WAVE 13           /* free-surface wave step and top pressure */
#define MAC3D_KC_PARCELS 14        /* parcel step and splat (MAC3D_STAGE_PARCELS) */
#define MAC3D_KC_COUNT 15

typedef struct mac3d_config {
  int32_t nx, ny, nz;   /* level-0 cells; each must divide by 2^levels */
  int32_t levels;       /* coarsenings below level 0 (3 => L0..L3) */
  int32_t nu1, nu2;     /* pre-/post-smoothing RBGS sweeps, every level */
  int32_t cycles;       /* V-cycles per step (FP32 and FP64 modes) */
  int32_t precision;    /* MAC3D_PREC_* */
  int32_t ir_max_iters; /* mixed: cap on refinement iterations per step */
  int32_t device;       /* HIP device ordinal */
  int32_t extrapolate;  /* mixed: warm-start extrapolation order of the FP64
                           pressure (-1: mode default 2; 0, 1, 2) */
  double dx;            /* cell size (m) */
  double rho_w, mu_w;   /* ambient water density (kg/m^3), viscosity (Pa s) */
  double rho_b;         /* blood density (kg/m^3) */
  double tau_y, mu_inf; /* Casson yield stress (Pa), infinite-shear visc. */
  double g;             /* gravity (m/s^2), acts along -y */
  double gamma_reg;     /* Casson shear-rate regulariser (1/s) */
  double shear_eps2;    /* (m/s)^2 added under the strain-rate root */
  double alpha_max;     /* explicit viscous stability clamp on nu dt/h^2 */
  double vel_damp;      /* velocity multiplier after advection */
  double maccormack;    /* dye MacCormack correction strength */
  double ir_tol;        /* mixed: relative divergence target ||r||/||b|| */
  double sor_omega;     /* smoother over-relaxation (0: mode default, 1 in the
                           FP32/FP64 modes as the reference, 1.25 in mixed) */
} mac3d_config;

/* one spherical blood blob of the seeded initial condition (cell units:
 * cell (i,j,k) has its centre at (i+0.5, j+0.5, k+0.5)) */
typedef struct mac3d_blob {
  double cx, cy, cz; /* centre */
  double radius;     /* radius; dye edge is a smoothstep over one cell */
  double w0;         /* initial downward speed (m/s), applied to V faces */
} mac3d_blob;

typedef struct mac3d_stats {
  int32_t steps;      /* steps completed by the call */
  int32_t vcycles;    /* V-cycles run by the call, all steps */
  int32_t ir_iters;   /* mixed: refinement iterations of the last step */
  int32_t converged;  /* mixed: last step met ir_tol */
  double max_div;     /* max |D(u)| after the last projection (m/s) */
  double l2_div;      /* RMS of D(u) after the last projection (m/s) */
  double max_div_rate;/* max |D(u)|/h (1/s) */
  double rel_res;     /* ||r||_2/||b||_2 of the last pressure solve */
  double ms;          /* wall time of the call (ms) */
} mac3d_stats;

#define MAC3D_MAX_BLOBS 16

/* gradient of L = 1/2 ||dye_N - target||^2 (adjoint of the whole run) */
typedef struct mac3d_grad {
  double loss;
  double d_rho_b;                     /* dL/d rho_b (density excess; rho_w fixed) */
  int32_t nblobs;
  int32_t steps;
  double d_blob[MAC3D_MAX_BLOBS][5];  /* dL/d (cx, cy, cz, radius, w0) per blob */
} mac3d_grad;

/* replaces: constructor texture/pipeline setup, BloodDropEngine.ts:545-589,
 * 861-897, 1100-1260, and the SimUniforms parameter block (common.ts:26-43,
 * uniformLayout.ts:56-190) for the MAC step */
int mac3d_default_config(mac3d_config *cfg);
int mac3d_create(const mac3d_config *cfg, void **ctx_out);
/* replaces: BloodDropEngine.destroy (BloodDropEngine.ts:2671-2736);
 * idempotent on NULL */
int mac3d_destroy(void *ctx);

/* replaces: queue.writeBuffer / texture readback (BloodDropEngine.ts:2602,
 * verify.mjs:560-606); count = elements of that field at that level */
int mac3d_upload(void *ctx, int field, int level, const void *host, int dtype,
                 size_t count);
int mac3d_download(void *ctx, int field, int level, void *host, int dtype,
                   size_t count);

/* replaces: resetSimulation's FLUID_INIT_VEL + FLUID_CLEAR_R32 dispatches
 * (BloodDropEngine.ts:1734-1744, fluid.ts:80-138) with the seeded blob
 * initial condition: dye and V from the blobs, U = W = P = 0 */
int mac3d_init_blobs(void *ctx, const mac3d_blob *blobs, int nblobs);

/* replaces: encodePhysicsStep's MAC portion (BloodDropEngine.ts:1916-2069)
 * plus CONC_PREDICT/CONC_RESOLVE's dye core (advect.ts:28-72,196-222),
 * nsteps times at fixed dt, blocking until the stream completes.
 * dt must be finite and > 0 (EINVAL otherwise, state untouched).  With
 * stats, a non-finite divergence norm after the last step returns
 * MAC3D_ENONFINITE (stats still filled); the mixed mode returns it as soon
 * as a residual norm is non-finite, with or without stats.  (The reference
 * clamps dt and keeps stepping, dropKinematics.ts:605, fluid.ts:182.) */
int mac3d_step(void *ctx, double dt, int nsteps, mac3d_stats *stats);

/* Checkpoints of the evolving state, in a byte buffer the caller owns.
 * replaces: the reference resumes by deterministic replay, resetSimulation +
 * seekToSimTime (BloodDropEngine.ts:1682-1763, 2142-2179); these let a host
 * keep the state itself and resume from it instead of replaying.
 * Layout: an 88-byte header (magic "MAC3DCK1", ABI version, nx, ny, nz, levels,
 * precision, element size, wave n (0: coupling off), parcel count, the
 * mixed mode's pressure-history count and refinement-iteration log, and
 * whether the 4-channel concentration is on), then
 * U, V, W, DYE (element size each, n cells), the pressure (mixed: the FP64
 * pressures of the current and the two previous steps, which the warm start
 * extrapolates from; else P), the wave state (4 floats per texel), the
 * parcels (8 floats each) and the 4-channel concentration (4 planes of n
 * floats) when present.  A step after a load gives bit for
 * bit what the same step gives after the save.  Not in the checkpoint: the
 * per-step inputs (dt, the drop slots, the coupling's parameters and scene
 * clock), which the host sets each step as the reference packs them per
 * substep, and the splat counters, which each parcel stage clears. */
/* bytes a checkpoint of the context's current state takes */
int mac3d_checkpoint_size(void *ctx, size_t *bytes);
/* writes the checkpoint into host[0 .. size); EINVAL if bytes < size */
int mac3d_checkpoint_save(void *ctx, void *host, size_t bytes);
/* restores a checkpoint saved from a context of the same grid, levels and
 * precision, with the same wave n (enable the coupling first) */
int mac3d_checkpoint_load(void *ctx, const void *host, size_t bytes);

/* Adjoint of nsteps steps from the blob initial condition (no reference
 * counterpart: the reference has no differentiation).  Runs the forward
 * pass keeping the state every checkpoint_every steps, computes
 * L = 1/2 ||dye_N - target||^2, and sweeps back step by step, recomputing
 * each segment from its checkpoint; returns dL/d(blob parameters) and
 * dL/d rho_b.  target: level-0 dye field (count = cells, dtype F32/F64).
 * Leaves the context at an unspecified state of the run. */
int mac3d_adjoint(void *ctx, const mac3d_blob *blobs, int nblobs, double dt, int nsteps,
                  int checkpoint_every, const void *target, int dtype, mac3d_grad *out);

/* Drop releases.  The reference lets a drop go with
 * MultiDropSystem.spawnAerialDrop(offsetXZ, radius, seed, startY, velY)
 * (src/sim/dropKinematics.ts:354-391): the offset is clamped to the disk of
 * MAX_DROP_OFFSET_XZ, the radius to [PENDANT_MIN_RADIUS, 0.095], and the
 * aerial branch of MultiDropSystem.step (:651-707) lets it fall freely under
 * G_SLOW, solving each substep's contact time exactly, until its lowest point
 * reaches WATER_SURFACE_Y.  The release map below takes that contact state in
 * closed form and makes it a blob of the grid; the seed only shapes the
 * drop's oscillation in the reference, so it has no part here. */
typedef struct mac3d_release {
  double x, z;    /* offsetXZ (scene m) */
  double radius;  /* drop radius (scene m) */
  double start_y; /* height of the centre at release (scene m; Y_START 0.495) */
  double vel_y;   /* vertical velocity at release (scene m/s; V_START -0.075) */
} mac3d_release;

typedef struct mac3d_release_map {
  double origin_x, origin_y, origin_z; /* scene position (m) of the grid's corner (0, 0, 0) */
  double cell;       /* scene metres per grid cell */
  double surface_y;  /* WATER_SURFACE_Y (0.30) */
  double gravity;    /* G_SLOW (-0.155 m/s^2) */
  double vel_scale;  /* blob w0 (m/s) per scene m/s of downward contact speed */
  double submerge;   /* the blob's centre sits this many radii under the surface */
  double max_offset; /* MAX_DROP_OFFSET_XZ (0.26) */
  double r_min, r_max; /* radius clamp (PENDANT_MIN_RADIUS 0.018, 0.095) */
} mac3d_release_map;

/* dL/d(release parameters), chained from the blob gradient */
typedef struct mac3d_release_grad {
  double loss;
  double d_rho_b;
  int32_t n;
  int32_t steps;
  double d_release[MAC3D_MAX_BLOBS][5]; /* dL/d (x, z, radius, start_y, vel_y) */
  double d_blob[MAC3D_MAX_BLOBS][5];    /* dL/d (cx, cy, cz, radius, w0) it came from */
} mac3d_release_grad;

/* The reference tank's width (2 x DOMAIN_HALF_XZ = 0.96 m, uniformLayout.ts:9-11) on
 * nx cells, centred in x and z, the grid's top face at the surface, the
 * reference's constants, vel_scale 1, submerge 1 (the blob's top at the
 * surface). */
int mac3d_release_default_map(const mac3d_config *cfg, mac3d_release_map *map);
/* Blobs of n releases (host arithmetic only: no context, no device).  jac, if
 * not NULL, receives n 5x5 Jacobians, d blob[q].(cx, cy, cz, radius, w0)[a] /
 * d release[q].(x, z, radius, start_y, vel_y)[b] at jac[25 q + 5 a + b]
 * (zero across a clamp that binds). */
int mac3d_release_blobs(const mac3d_release_map *map, const mac3d_release *rel, int n,
                        mac3d_blob *blobs, double *jac);
/* mac3d_adjoint from the releases' blobs, with the gradient taken on to the
 * release parameters through the map's Jacobian */
int mac3d_adjoint_release(void *ctx, const mac3d_release_map *map, const mac3d_release *rel,
                          int n, double dt, int nsteps, int checkpoint_every,
                          const void *target, int dtype, mac3d_release_grad *out);

/* Fitting releases to a target dye field: quasi-Newton (BFGS with a
 * backtracking line search) on the release parameters, each evaluation one
 * mac3d_adjoint_release.  Parameters are scaled so that one unit moves the
 * blob by about a cell (x, z and radius by one cell; start_y and vel_y by
 * the change of contact speed that moves the blob a cell over the run). */
typedef struct mac3d_fit_opts {
  int32_t max_iters;        /* quasi-Newton iterations (20) */
  int32_t checkpoint_every; /* of each adjoint (10) */
  uint64_t mask;            /* bit 5 q + b fits parameter b of release q; 0: x, z, radius,
                               vel_y of every release (start_y and vel_y reach the dye only
                               through the contact speed, so one of them is held) */
  double tol;               /* stop when an iteration lowers the loss by less than
                               tol of itself (1e-6) */
  double gtol;              /* or when every scaled gradient component is below
                               gtol times the loss (1e-9) */
} mac3d_fit_opts;

typedef struct mac3d_fit_report {
  int32_t iters;     /* quasi-Newton iterations taken */
  int32_t evals;     /* adjoint evaluations */
  int32_t converged; /* 1: stopped by tol or gtol, 0: by max_iters or a failed line search */
  int32_t reserved;
  double loss0, loss; /* at the guess and at the result */
} mac3d_fit_report;

int mac3d_fit_default_opts(mac3d_fit_opts *opts);
/* rel: the guess in, the fit out (1..12 releases: five mask bits each).
 * Leaves the context at an unspecified state of the last evaluation's run. */
int mac3d_fit_release(void *ctx, const mac3d_release_map *map, mac3d_release *rel, int n,
                      double dt, int nsteps, const void *target, int dtype,
                      const mac3d_fit_opts *opts, mac3d_fit_report *report);

/* Free-surface coupling.  The wave solver of the
 * reference scene -- a gravity-capillary spectral propagator on an n x n grid
 * with drop impulses, surfactant film, Marangoni pull and the MAC top layer's
 * upward velocity as forcing (src/shaders/wave.ts:56-388) -- and the top
 * Dirichlet pressure it feeds into the MAC divergence and projection
 * (evalTopDirichletPressure, src/shaders/fluid.ts:16-78, used at
 * fluid.ts:548-551 and 896-898).  Enabled, every step first advances the
 * wave one step and evaluates the lid pressure, in the reference's order
 * (BloodDropEngine.ts:1886-1914, before the MAC advection).  The wave computes in FP32 in every mode.
 * Lengths in metres; the tank spans [-half_xz, half_xz] in x and z. */
typedef struct mac3d_wave_params {
  int32_t n;          /* wave texels per side, power of two in [16, 1024] (512) */
  int32_t scene;      /* 1: the step also takes the reference scene's terms: in
                         the advection the drop wetting-neck momentum, Marangoni
                         pull and kinematic top from the wave, wall/floor sponges
                         (fluid.ts:198-321); in the Casson pass the vorticity
                         confinement from the curl of the advected velocity, the
                         solenoidal unfurling forcing and the +-1.8 m/s clamp
                         (fluid.ts:331-509); the tank top at surface_y.
                         0 (default): not */
  double half_extent; /* half-width of the wave domain (WAVE_HALF_EXTENT, 1.60) */
  double half_xz;     /* tank half-width (domainBounds.x; 0: nx*dx/2) */
  double depth;       /* tank depth (domainBounds.z - domainBounds.y; 0: ny*dx) */
  double surface_y;   /* still-surface height (surfaceParams.x, 0.30) */
  double marangoni;   /* Marangoni strength (surfaceParams.y, 0) */
  double vort_conf;   /* scene: vorticity confinement scale (fluidParams.x, 1.02) */
  double sim_time;    /* scene: simulation time of the unfurling phase
                         (cameraPosAndTime.w, 0 s) */
  double schlieren;   /* scene: schlierenStr, the unfurling phase offset
                         (particleParams.x, 0.032) */
  double hemolysis;   /* parcels' osmotic lysis rate (lensAndRheology.w, 0.088) */
  int32_t physics;    /* 1: the reference's own lattice physics in place of the
                         workload's SI laws: the lattice Casson law with its
                         50 s^-1 strain scale and [1e-4, 0.0044] clamp, the
                         tanh-saturated buoyancy gated under the surface, the
                         viscous stencil's neighbours clamped to the edge
                         (fluid.ts:387-392, 451-486), the projection's +-2
                         clamp (fluid.ts:908) and the dye predictor's static /
                         one-stage branches with the walls and top row held at
                         0 (advect.ts:42-67, 100-103).  0 (default): not.
                         With mac3d_set_uniforms the block's buoyancy [45],
                         velocityDamping [46], concSharpness [47] and
                         cassonYield [58] then set these and the config's
                         vel_damp and maccormack (0.45 concSharpness) */
  int32_t reserved;
  double casson_scale;/* physics 1: Casson yield scale (lensAndRheology.z, 1.0) */
  double buoyancy;    /* physics 1: buoyancy strength (fluidParams.y, 0.48) */
} mac3d_wave_params;

/* one drop slot (SimUniforms.drops[d], common.ts:26-43, packed by
 * uniformLayout.ts:161-187): active when seed > 0 and drain > 0.0005 */
typedef struct mac3d_drop {
  double x, y, z, radius; /* posAndRadius (m) */
  double vel_y, osc;      /* velOscAndDrain.x (m/s), .y */
  double drain, seed;     /* velOscAndDrain.z, .w */
} mac3d_drop;
#define MAC3D_MAX_DROPS 8

/* no reference counterpart (the reference fills these uniforms itself) */
int mac3d_wave_default_params(mac3d_wave_params *p);
/* replaces: the wave textures and pipelines of the BloodDropEngine
 * constructor (BloodDropEngine.ts:545-589) and WAVE_CLEAR (wave.ts:15-30):
 * allocates the wave and resets it to a flat surface; p == NULL disables
 * the coupling (the lid pressure is 0 again) */
int mac3d_wave_enable(void *ctx, const mac3d_wave_params *p);
/* replaces: the drop slots of SimUniformPacker.pack (uniformLayout.ts:161-187);
 * ndrops in [0, 8], the other slots inactive; ESTATE unless enabled */
int mac3d_set_drops(void *ctx, const mac3d_drop *drops, int ndrops);
/* replaces: queue.writeBuffer of the packed 512-B SimUniforms block
 * (BloodDropEngine.ts:2602; layout common.ts:26-43, packed by
 * uniformLayout.ts:56-190) for what the coupling reads from it: count = 128
 * floats; cameraPosAndTime.w [35] (simTime), surfaceParams.x/.y [40, 41]
 * (surface height, Marangoni), fluidParams.x [44] (vorticity confinement),
 * domainBounds [52..55] (half_xz = .x, depth = .z - .y), particleParams.x
 * [60] (schlierenStr) and the 8 drop slots [64..127].  dt stays the step's
 * argument.  ESTATE unless enabled. */
int mac3d_set_uniforms(void *ctx, const float *block, size_t count);

/* Lagrangian blood parcels (RbcParcel, particles.ts:16-20): 8 floats each,
 * position (m), intact mass (0: inactive), stretch vector (m), age ticks +
 * lysis fraction.  MAC3D_STAGE_PARCELS advances them one step through the
 * centre velocity (PARTICLE_ADVECT_SPLAT, particles.ts:189-463) and splats
 * their mass into two 32-bit counters per cell, 65536 per unit mass (intact,
 * lysed; splatIntactBuffer/splatLysedBuffer).  The world box is the
 * coupling's: [-half_xz, half_xz] x [surface_y - depth, surface_y] x
 * [-half_xz, half_xz].
 * replaces: the parcel storage buffer's writeBuffer (BloodDropEngine.ts:1716-1720);
 * n parcels (n >= 0); ESTATE unless the coupling is enabled.  The stage runs in
 * the FP32 and mixed modes (EINVAL in FP64: the parcels sample the FP32
 * centre velocity) */
int mac3d_parcels_upload(void *ctx, const float *parcels, int32_t n);
/* replaces: the spawnCursor / spawnCount of the packed block
 * (particleParams.y/.z, BloodDropEngine.ts:1821-1844) consumed by
 * PARTICLE_EMIT (particles.ts:53-186, dispatched at BloodDropEngine.ts:2082-2087):
 * the next parcel stage (MAC3D_STAGE_PARCELS or the step's, with the 4-channel
 * concentration on) first emits `count` parcels (<= 8192) into the uploaded
 * parcel ring from slot cursor mod n, from the active drop slots in
 * proportion to their drain rates, moving a live parcel's mass at a written
 * slot into the lysed counter; the request is then consumed.  ESTATE without
 * the coupling or (count > 0) without uploaded parcels. */
int mac3d_parcels_emit(void *ctx, int32_t cursor, int32_t count);
/* no reference counterpart as a call (the reference always runs its pass B):
 * turns on or off the reference's 4-channel concentration (MAC3D_FIELD_CONC,
 * zero at first) on its own grid: enable 1, the MAC grid; enable 2, a grid
 * twice as fine per axis, as the reference holds it (CONC_GRID 192x256x192
 * over VEL_GRID 96x128x96, common.ts:4-15), which the kernels address in uvw
 * and sample the MAC grid's centre velocity at, and whose channels the Casson
 * pass reads as samples at the MAC cell centres and y-faces (fluid.ts:446-447,
 * 477-478); enable 0, off.  A change of grid starts from zero.  The parcels
 * splat into the concentration grid while it is on (else the MAC grid).
 * While on, each step runs the reference's
 * pass B after the projection (BloodDropEngine.ts:2073-2108): the parcel stage
 * (counters cleared, pending emission, advection and splat), CONC_PREDICT of
 * the lysed channel and CONC_RESOLVE (advect.ts:28-298) in place of the dye
 * transport, which then stands still; the Casson pass, the buoyancy and the
 * parcels' hemolysis read the intact (.r) and lysed (.g) channels in place of
 * the dye.  Needs the coupling and the FP32 or mixed step (ESTATE, EINVAL). */
int mac3d_conc_enable(void *ctx, int enable);
/* the parcels after MAC3D_STAGE_PARCELS (n <= the uploaded count) */
int mac3d_parcels_download(void *ctx, float *parcels, int32_t n);
/* the splat counters: which 0 intact, 1 lysed; count = the splat grid's cells
 * (the concentration grid while it is on, else the MAC grid) */
int mac3d_splat_download(void *ctx, int which, uint32_t *out, size_t count);

/* one stage of the step (MAC3D_STAGE_*), blocking */
int mac3d_stage(void *ctx, int stage, double dt);
/* one multigrid operator (MAC3D_MG_*) on a level, blocking */
int mac3d_mg_op(void *ctx, int op, int level, int count);
/* max|D|, RMS(D) of the current velocity, FP64 reduction, blocking */
int mac3d_div_norms(void *ctx, double *max_div, double *l2_div);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc: mac3d_div_norms on caller-owned device arrays, with no
 * context and no allocation.  u, v, w: device arrays of an nx*ny*nz face
 * velocity in the layout of MAC3D_FIELD_U/V/W, dtype MAC3D_F32 or MAC3D_F64;
 * out: device array of 2 doubles that receives max |D(u)| and RMS(D) (m/s),
 * the same reduction as mac3d_div_norms; stream: a hipStream_t (NULL: the null
 * stream).  Only enqueues work: the caller synchronises before reading out.
 * The per-block partial sums live in a static device array of 131072
 * (max, sum) pairs (2 MiB) on each GPU, kept until the process ends: grids
 * with ceil(nx/64) ceil(ny/4) nz > 131072 are refused with EINVAL.  Calls that
 * overlap on one GPU share that array and race; serialise them (one stream, or
 * events).  Errors: mac3d_last_error(NULL). */
int mac3d_div_norms_noalloc(const void *u, const void *v, const void *w, int dtype, int nx,
                            int ny, int nz, double *out, void *stream);

/* kernel timing (no reference counterpart; the reference reads pass-level
 * timestamp queries, BloodDropEngine.ts:411-439,2400-2456).  enable != 0
 * clears the totals and brackets every launch with HIP events on the
 * context's stream; read returns the class's total ms and launch count. */
int mac3d_profile(void *ctx, int enable);
int mac3d_profile_read(void *ctx, int kclass, double *ms_total, int64_t *launches);

const char *mac3d_last_error(void *ctx);
const char *mac3d_version(void);

#ifdef __cplusplus
}
#endif

#endif /* MAC3D_H */
