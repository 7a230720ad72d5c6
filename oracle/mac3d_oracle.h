/*
 * mac3d_oracle.h — FP64 CPU restatement of the reference's MAC step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against (tests, the smoke check, the CPU-baseline leg of the
 * benchmark).  Nothing in the product links, loads or calls it.
 *
 * Every function follows one WGSL kernel of the reference, cited at its
 * definition in mac3d_oracle.c, with the documented changes the MI355X
 * workload makes (SI Casson/Boussinesq, no-slip walls, exact coarse solve).
 * Parity unpinned: advect_scene's azimuthal modulation between the lines
 * through a drop's axis along x and z (its value there is pinned in closed
 * form; in between only the device parity checks it); parcel_step's
 * meniscus-hold, stretch and hemolysis heuristics beyond what the splat,
 * settling and mass-conservation pins fix (only the device parity checks
 * them).
 * Arrays are cell-major with x fastest: idx = (k*ny + j)*nx + i.
 */
#ifndef MAC3D_ORACLE_H
#define MAC3D_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
  int nx, ny, nz;
  int levels, nu1, nu2, cycles;
  int solve_to_tol; /* 0: fixed cycles; 1: cycles until ||r||/||b|| <= tol */
  int max_cycles;
  double dx, rho_w, mu_w, rho_b, tau_y, mu_inf, g;
  double gamma_reg, shear_eps2, alpha_max, vel_damp, maccormack, tol;
  double omega; /* smoother over-relaxation (0 or 1: plain Gauss-Seidel) */
} orc_params;

typedef struct orc_blob {
  double cx, cy, cz, radius, w0;
} orc_blob;

/* free-surface wave coupling (lengths in metres; 0 picks nx dx / 2 and
 * ny dx for half_xz and depth) and one drop slot, as in include/mac3d.h */
typedef struct orc_wave_params {
  int n, scene; /* scene: also the advection's and Casson pass's scene terms */
  double half_extent, half_xz, depth, surface_y, marangoni;
  double vort_conf, sim_time, schlieren; /* Casson scene terms */
  double hemolysis;                      /* parcels' lysis rate */
  int physics;  /* 1: the reference's own lattice physics (see casson_ref) */
  int reserved;
  double casson_scale, buoyancy;         /* lensAndRheology.z, fluidParams.y */
} orc_wave_params;

typedef struct orc_drop {
  double x, y, z, radius, vel_y, osc, drain, seed;
} orc_drop;

/* state of one simulation held by the oracle */
typedef struct orc_state orc_state;

orc_state *orc_create(const orc_params *p);
void orc_destroy(orc_state *s);
/* field ids match MAC3D_FIELD_* of include/mac3d.h */
double *orc_field(orc_state *s, int field, int level);
long orc_field_len(orc_state *s, int field, int level);
void orc_init_blobs(orc_state *s, const orc_blob *b, int nb);
/* returns V-cycles run; *rel_res gets the last ||r||/||b|| */
int orc_step(orc_state *s, double dt, int nsteps, double *rel_res);
void orc_stage(orc_state *s, int stage, double dt);
void orc_mg_op(orc_state *s, int op, int level, int count);
void orc_div_norms(orc_state *s, double *max_div, double *l2_div);
int orc_set_threads(int n);
/* enable (p != NULL: flat surface, no drops) or disable the wave coupling */
void orc_wave_enable(orc_state *s, const orc_wave_params *p);
void orc_set_drops(orc_state *s, const orc_drop *d, int n);
/* Lagrangian parcels: n x 8 doubles (x, y, z, mass, stretch x, y, z,
 * age ticks + lysis); orc_parcels returns them after ST_PARCELS, orc_splat
 * the intact (0) and lysed (1) counters, one per cell of the splat grid
 * (orc_splat_dims: the concentration grid while it is on, else the MAC grid) */
void orc_parcels_set(orc_state *s, const double *p, int n);
/* the next parcel stage first emits `count` parcels from ring slot `cursor`
 * (PARTICLE_EMIT); the 4-channel concentration (stages 11, 12 and pass B in
 * the step; field 16 = its planes r, g, b, a, field 17 = the lysed channel's
 * predictor): on = 0 off, 1 on the MAC grid,
 * 2 on a grid twice as fine per axis (the reference's CONC_GRID over its
 * VEL_GRID); a change of grid starts from zero */
void orc_parcels_emit(orc_state *s, int cursor, int count);
void orc_conc_enable(orc_state *s, int on);
double *orc_parcels(orc_state *s);
uint32_t *orc_splat(orc_state *s, int which);
void orc_splat_dims(orc_state *s, int out[3]);
/* unitary in-place complex FFT of length n (a power of two) over interleaved
 * (re, im) doubles, sign -1 forward / +1 inverse: exposed for its pinning */
void orc_fft(double *z, int n, int stride, int sign);

/* helpers exposed for the pinning tests */
double orc_sample_cell(const double *f, int nx, int ny, int nz, double ux,
                       double uy, double uz);
double orc_smoothstep(double e0, double e1, double x);
double orc_top_coeff(int level);

#ifdef __cplusplus
}
#endif
#endif
