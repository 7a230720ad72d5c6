/*
 * mac3d_oracle.c — FP64 CPU restatement of drop-of-blood's MAC step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Only tests, the
 * smoke check and the CPU-baseline leg of the benchmark load it.
 *
 * Each function cites the reference kernel it restates.  Changes the MI355X
 * workload makes to the reference scene are marked "WORKLOAD:" where they
 * occur; the restatement is otherwise operation-for-operation:
 *   - WebGPU textureSampleLevel with a linear clamp-to-edge sampler
 *     (BloodDropEngine.ts:497-505) is restated with exact FP64 weights;
 *   - velocity is held in FP64 here (rgba16float in the reference,
 *     BloodDropEngine.ts:550-553).
 * OpenMP parallelises the k loops; every loop body is independent (red-black
 * colour passes only read the other colour), so results do not depend on the
 * thread count.  Norm reductions sum per-plane partials in plane order.
 */
#include "mac3d_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXLEV 8

enum { F_U, F_V, F_W, F_P, F_DYE, F_CU, F_CV, F_CW, F_RHS, F_MGX, F_MGB,
       F_MGR, F_DYE_PRED, F_WAVE, F_PTOP, F_CURL, F_CONC, F_CONC_PRED, F_COUNT };
enum { ST_F2C, ST_ADVECT, ST_CASSON, ST_DIV, ST_VCYCLE, ST_PROJECT,
       ST_DYE_PRED, ST_DYE_RES, ST_WAVE, ST_TOP_PRESSURE, ST_PARCELS, ST_CONC_PRED,
       ST_CONC_RES };
#define MAX_DROPS 8
enum { MG_SMOOTH, MG_RESIDUAL, MG_RESTRICT, MG_PROLONG, MG_COARSE, MG_ZERO };

typedef struct { int nx, ny, nz; } dims3;

struct orc_state {
  orc_params p;
  dims3 d[MAXLEV + 1];
  long n[MAXLEV + 1];
  double *u, *v, *w, *pr, *dye, *cu, *cv, *cw, *rhs, *dpred;
  double *u2, *v2, *w2;
  double *mgx[MAXLEV + 1], *mgb[MAXLEV + 1], *mgr[MAXLEV + 1];
  double *cg[4]; /* coarse-solve CG vectors */
  /* wave coupling: state (4 per texel), complex work (2 per texel), film,
   * lid pressure (nx * nz) */
  int wave_on;
  orc_wave_params wp;
  double *wave, *wz, *wz2, *wfilm, *ptop;
  orc_drop drops[MAX_DROPS];
  double *curl; /* scene: curl of the advected centre velocity, 4 per cell */
  /* Lagrangian parcels (8 per parcel: position, mass, stretch, age + lysis)
   * and their two splat counters (intact, lysed; one per cell of the splat
   * grid: the concentration grid while it is on, else the MAC grid) */
  double *parcels;
  int np, np_cap;
  uint32_t *splat[2];
  long splat_cap;
  /* the reference's 4-channel concentration (planes: intact RBC, lysed Hb,
   * oxygenation/shear state, floor bed) and its twin on its own grid, conc_r
   * times the MAC grid per axis (the reference's CONC_GRID is twice its
   * VEL_GRID, common.ts:4-15), the lysed channel's predictor there; the next
   * parcel stage's emission (ring cursor, count) */
  int conc_on, conc_r;
  dims3 dc;
  long nc;
  double *conc, *conc2, *cpred;
  int emit_cursor, emit_count;
};

#define IDX(D, i, j, k) ((((long)(k)) * (D).ny + (j)) * (long)(D).nx + (i))

/* the intact-RBC concentration the Casson pass, the buoyancy and the parcels
 * read: the dye, or with the 4-channel concentration its .r plane; the lysed
 * one (.g) is 0 without it (NULL); conc_dims is the grid they live on */
static const double *intact_of(const orc_state *s) { return s->conc_on ? s->conc : s->dye; }
static const double *lysed_of(const orc_state *s) { return s->conc_on ? s->conc + s->nc : NULL; }
static dims3 conc_dims(const orc_state *s) { return s->conc_on ? s->dc : s->d[0]; }
static long splat_cells(const orc_state *s) { return s->conc_on ? s->nc : s->n[0]; }

static double clampd(double x, double lo, double hi) {
  return x < lo ? lo : (x > hi ? hi : x);
}
static int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }
static double mixd(double a, double b, double t) { return a * (1.0 - t) + b * t; }

double orc_smoothstep(double e0, double e1, double x) {
  /* WGSL smoothstep (used at fluid.ts:74, advect.ts and the blob edge) */
  double t = clampd((x - e0) / (e1 - e0), 0.0, 1.0);
  return t * t * (3.0 - 2.0 * t);
}

int orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

/* textureSampleLevel(tex, linear clamp-to-edge sampler, uvw): texel space
 * t = uvw*N - 0.5, floor/frac, clamped texel indices, x then y then z lerp.
 * Used by sampleMacU/V/W and traceBackUVW (fluid.ts:154-174) and the dye
 * predictor (advect.ts:52-69). */
static double sample_cell(const double *f, dims3 d, double ux, double uy, double uz) {
  double tx = ux * d.nx - 0.5, ty = uy * d.ny - 0.5, tz = uz * d.nz - 0.5;
  double bx = floor(tx), by = floor(ty), bz = floor(tz);
  double fx = tx - bx, fy = ty - by, fz = tz - bz;
  int i0 = clampi((int)bx, 0, d.nx - 1), i1 = clampi((int)bx + 1, 0, d.nx - 1);
  int j0 = clampi((int)by, 0, d.ny - 1), j1 = clampi((int)by + 1, 0, d.ny - 1);
  int k0 = clampi((int)bz, 0, d.nz - 1), k1 = clampi((int)bz + 1, 0, d.nz - 1);
  double c000 = f[IDX(d, i0, j0, k0)], c100 = f[IDX(d, i1, j0, k0)];
  double c010 = f[IDX(d, i0, j1, k0)], c110 = f[IDX(d, i1, j1, k0)];
  double c001 = f[IDX(d, i0, j0, k1)], c101 = f[IDX(d, i1, j0, k1)];
  double c011 = f[IDX(d, i0, j1, k1)], c111 = f[IDX(d, i1, j1, k1)];
  double x00 = mixd(c000, c100, fx), x10 = mixd(c010, c110, fx);
  double x01 = mixd(c001, c101, fx), x11 = mixd(c011, c111, fx);
  return mixd(mixd(x00, x10, fy), mixd(x01, x11, fy), fz);
}

double orc_sample_cell(const double *f, int nx, int ny, int nz, double ux,
                       double uy, double uz) {
  dims3 d = {nx, ny, nz};
  return sample_cell(f, d, ux, uy, uz);
}

/* A concentration channel f (intact_of / lysed_of) as a MAC cell's Casson,
 * buoyancy and unfurling terms read it: textureSampleLevel at the cell
 * centre (uvwCell, fluid.ts:446-447) and at the top y-face (uvwFaceY
 * clamped to [0.001, 0.999], :477-478).  On the MAC grid itself (the dye, or
 * the concentration at conc_r 1) these are the cell's value and the mean of
 * it and the cell above (the top row: its own); on a finer concentration
 * grid, trilinear samples of it. */
static double conc_at_cell(const orc_state *s, const double *f, int i, int j, int k) {
  dims3 d = s->d[0];
  if (!s->conc_on || s->conc_r == 1) return f[IDX(d, i, j, k)];
  return sample_cell(f, s->dc, (i + 0.5) / d.nx, (j + 0.5) / d.ny, (k + 0.5) / d.nz);
}
static double conc_at_yface(const orc_state *s, const double *f, int i, int j, int k) {
  dims3 d = s->d[0];
  if (!s->conc_on || s->conc_r == 1) {
    int jp = j + 1 < d.ny ? j + 1 : d.ny - 1;
    return 0.5 * (f[IDX(d, i, j, k)] + f[IDX(d, i, jp, k)]);
  }
  return sample_cell(f, s->dc, clampd((i + 0.5) / d.nx, 0.001, 0.999),
                     clampd((j + 1.0) / d.ny, 0.001, 0.999), clampd((k + 0.5) / d.nz, 0.001, 0.999));
}

static double top_stagnation_gate(const orc_state *s, int i, int j, int k);
static double sstep(double e0, double e1, double x);

/* ---- face -> centre: FLUID_MAC_TO_CENTER (fluid.ts:917-943); with the wave
 * coupling's scene flag, the top-row stagnation gate (fluid.ts:944-959,
 * top_stagnation_gate) on downward centre velocities of the top two rows. */
static void face_to_centre(orc_state *s) {
  dims3 d = s->d[0];
#pragma omp parallel for
  for (int k = 0; k < d.nz; ++k)
    for (int j = 0; j < d.ny; ++j)
      for (int i = 0; i < d.nx; ++i) {
        long c = IDX(d, i, j, k);
        double uR = (i >= d.nx - 1) ? 0.0 : s->u[c];
        double uL = (i == 0) ? 0.0 : s->u[IDX(d, i - 1, j, k)];
        double vU = s->v[c];
        double vD = (j == 0) ? 0.0 : s->v[IDX(d, i, j - 1, k)];
        double wF = (k >= d.nz - 1) ? 0.0 : s->w[c];
        double wB = (k == 0) ? 0.0 : s->w[IDX(d, i, j, k - 1)];
        s->cu[c] = 0.5 * (uL + uR);
        s->cv[c] = 0.5 * (vD + vU);
        s->cw[c] = 0.5 * (wB + wF);
        if (s->wave_on && s->wp.scene && j >= d.ny - 2 && s->cv[c] < 0.0)
          s->cv[c] *= top_stagnation_gate(s, i, j, k);
      }
}

/* traceBackUVW (fluid.ts:169-174): RK2 midpoint backtrace on the
 * cell-centred velocity, uvw clamped to [0.002, 0.998]. */
static void trace_back(const orc_state *s, double dt, const double uvw[3],
                       double out[3]) {
  dims3 d = s->d[0];
  double dom[3] = {d.nx * s->p.dx, d.ny * s->p.dx, d.nz * s->p.dx};
  double p0[3], mid[3], v0[3], vm[3];
  for (int a = 0; a < 3; ++a) p0[a] = clampd(uvw[a], 0.002, 0.998);
  v0[0] = sample_cell(s->cu, d, p0[0], p0[1], p0[2]);
  v0[1] = sample_cell(s->cv, d, p0[0], p0[1], p0[2]);
  v0[2] = sample_cell(s->cw, d, p0[0], p0[1], p0[2]);
  for (int a = 0; a < 3; ++a) mid[a] = clampd(uvw[a] - 0.5 * (v0[a] * dt) / dom[a], 0.002, 0.998);
  vm[0] = sample_cell(s->cu, d, mid[0], mid[1], mid[2]);
  vm[1] = sample_cell(s->cv, d, mid[0], mid[1], mid[2]);
  vm[2] = sample_cell(s->cw, d, mid[0], mid[1], mid[2]);
  for (int a = 0; a < 3; ++a) out[a] = clampd(uvw[a] - (vm[a] * dt) / dom[a], 0.002, 0.998);
}

static void advect_scene(const orc_state *s, int i, int j, int k, double dt, double *vx,
                         double *vy, double *vz);

/* ---- FLUID_ADVECT_MAC (fluid.ts:144-192, 323-325): semi-Lagrangian RK2
 * advection of each staggered component, sampled with the -1/2 texel shift
 * of sampleMacU/V/W (fluid.ts:154-167), times velDamp.  WORKLOAD: the scene
 * forcing blocks (drop necks :198-285, Marangoni/kinematic top :287-310,
 * sponges :312-321) apply only with the wave coupling's scene flag
 * (advect_scene). */
static void advect_mac(orc_state *s, double dt) {
  dims3 d = s->d[0];
  double damp = s->p.vel_damp;
#pragma omp parallel for
  for (int k = 0; k < d.nz; ++k)
    for (int j = 0; j < d.ny; ++j)
      for (int i = 0; i < d.nx; ++i) {
        long c = IDX(d, i, j, k);
        double fX[3] = {(i + 1.0) / d.nx, (j + 0.5) / d.ny, (k + 0.5) / d.nz};
        double fY[3] = {(i + 0.5) / d.nx, (j + 1.0) / d.ny, (k + 0.5) / d.nz};
        double fZ[3] = {(i + 0.5) / d.nx, (j + 0.5) / d.ny, (k + 1.0) / d.nz};
        double b[3];
        trace_back(s, dt, fX, b);
        double vx = sample_cell(s->u, d, clampd(b[0] - 0.5 / d.nx, 0.001, 0.999),
                                clampd(b[1], 0.001, 0.999), clampd(b[2], 0.001, 0.999)) * damp;
        trace_back(s, dt, fY, b);
        double vy = sample_cell(s->v, d, clampd(b[0], 0.001, 0.999),
                                clampd(b[1] - 0.5 / d.ny, 0.001, 0.999), clampd(b[2], 0.001, 0.999)) * damp;
        trace_back(s, dt, fZ, b);
        double vz = sample_cell(s->w, d, clampd(b[0], 0.001, 0.999), clampd(b[1], 0.001, 0.999),
                                clampd(b[2] - 0.5 / d.nz, 0.001, 0.999)) * damp;
        if (s->wave_on && s->wp.scene) advect_scene(s, i, j, k, dt, &vx, &vy, &vz);
        if (i >= d.nx - 1) vx = 0.0;
        if (k >= d.nz - 1) vz = 0.0;
        s->u2[c] = vx; s->v2[c] = vy; s->w2[c] = vz;
      }
  double *t;
  t = s->u; s->u = s->u2; s->u2 = t;
  t = s->v; s->v = s->v2; s->v2 = t;
  t = s->w; s->w = s->w2; s->w2 = t;
}

/* Neighbour face values with the workload's wall rules (used by the viscous
 * stencil).  WORKLOAD: the reference clamps indices to the edge
 * (loadMac, fluid.ts:387-392), i.e. zero-gradient walls; the workload has
 * no-slip walls (tangential ghost = -value, normal wall face = 0) and a
 * free-slip lid (ghost = value). */
static double U_at(const orc_state *s, int i, int j, int k) {
  dims3 d = s->d[0];
  if (i < 0 || i >= d.nx) return 0.0;
  if (j < 0) return -s->u[IDX(d, i, 0, k)];
  if (j >= d.ny) return s->u[IDX(d, i, d.ny - 1, k)];
  if (k < 0) return -s->u[IDX(d, i, j, 0)];
  if (k >= d.nz) return -s->u[IDX(d, i, j, d.nz - 1)];
  return s->u[IDX(d, i, j, k)];
}
static double V_at(const orc_state *s, int i, int j, int k) {
  dims3 d = s->d[0];
  if (j < 0) return 0.0;
  if (j >= d.ny) return s->v[IDX(d, i, d.ny - 1, k)];
  if (i < 0) return -s->v[IDX(d, 0, j, k)];
  if (i >= d.nx) return -s->v[IDX(d, d.nx - 1, j, k)];
  if (k < 0) return -s->v[IDX(d, i, j, 0)];
  if (k >= d.nz) return -s->v[IDX(d, i, j, d.nz - 1)];
  return s->v[IDX(d, i, j, k)];
}
static double W_at(const orc_state *s, int i, int j, int k) {
  dims3 d = s->d[0];
  if (k < 0 || k >= d.nz) return 0.0;
  if (i < 0) return -s->w[IDX(d, 0, j, k)];
  if (i >= d.nx) return -s->w[IDX(d, d.nx - 1, j, k)];
  if (j < 0) return -s->w[IDX(d, i, 0, k)];
  if (j >= d.ny) return s->w[IDX(d, i, d.ny - 1, k)];
  return s->w[IDX(d, i, j, k)];
}

/* ---- FLUID_CASSON_FORCES (fluid.ts:377-511), parts (1) and (3):
 * strain-rate proxy from +-1 neighbour differences (fluid.ts:451-454),
 * Casson effective viscosity with the max(gamma+reg, reg) floor
 * (fluid.ts:458-462), explicit update mac += alpha*(sum6 - 6 mac)
 * (fluid.ts:464-465), buoyancy on y-faces from the face-sampled dye
 * (fluid.ts:477-486), wall faces zeroed (fluid.ts:506-507).
 * WORKLOAD: SI Casson parameters blended by dye c in [0,1]:
 *   mu_inf(c) = mu_w + c (mu_inf - mu_w), tau_y(c) = tau_y c^2 (the
 *   reference's c^2 yield law, fluid.ts:460), rho(c) = rho_w + c (rho_b -
 *   rho_w), gamma = |diff|/(2h) (the reference's 50 = 1/(2*0.01 m)),
 *   alpha = clamp(mu_eff/rho dt/h^2, 0, alpha_max);
 *   linear Boussinesq a_y = -g (rho_b - rho_w)/rho_w * dye_face in place of
 *   the tanh-saturated surface-gated term; vorticity confinement (2) and
 *   unfurling forcing (4) and the +-1.8 clamp are not part of the workload. */
static void casson_buoyancy(orc_state *s, double dt) {
  dims3 d = s->d[0];
  const orc_params *P = &s->p;
  double h = P->dx;
  double inv2h = 1.0 / (2.0 * h);
  double dtoh2 = dt / (h * h);
  double buoy = -P->g * (P->rho_b - P->rho_w) / P->rho_w * dt;
#pragma omp parallel for
  for (int k = 0; k < d.nz; ++k)
    for (int j = 0; j < d.ny; ++j)
      for (int i = 0; i < d.nx; ++i) {
        long c = IDX(d, i, j, k);
        double m[3] = {s->u[c], s->v[c], s->w[c]};
        double xp[3] = {U_at(s, i + 1, j, k), V_at(s, i + 1, j, k), W_at(s, i + 1, j, k)};
        double xn[3] = {U_at(s, i - 1, j, k), V_at(s, i - 1, j, k), W_at(s, i - 1, j, k)};
        double yp[3] = {U_at(s, i, j + 1, k), V_at(s, i, j + 1, k), W_at(s, i, j + 1, k)};
        double yn[3] = {U_at(s, i, j - 1, k), V_at(s, i, j - 1, k), W_at(s, i, j - 1, k)};
        double zp[3] = {U_at(s, i, j, k + 1), V_at(s, i, j, k + 1), W_at(s, i, j, k + 1)};
        double zn[3] = {U_at(s, i, j, k - 1), V_at(s, i, j, k - 1), W_at(s, i, j, k - 1)};
        double ss = 0.0;
        for (int a = 0; a < 3; ++a) {
          double du = xp[a] - xn[a], dv = yp[a] - yn[a], dw = zp[a] - zn[a];
          ss += du * du + dv * dv + dw * dw;
        }
        double gam = inv2h * sqrt(ss + P->shear_eps2);
        const double *cr = intact_of(s);
        double cc = clampd(conc_at_cell(s, cr, i, j, k), 0.0, 1.0);
        double sq_muinf = sqrt(P->mu_w + cc * (P->mu_inf - P->mu_w));
        double tau = P->tau_y * cc * cc;
        double yterm = sqrt(tau / fmax(gam + P->gamma_reg, P->gamma_reg));
        double mu_eff = (sq_muinf + yterm) * (sq_muinf + yterm);
        double rho = P->rho_w + cc * (P->rho_b - P->rho_w);
        double alpha = clampd(mu_eff / rho * dtoh2, 0.0, P->alpha_max);
        double out[3];
        for (int a = 0; a < 3; ++a) {
          double lap = (xp[a] + xn[a] + yp[a] + yn[a] + zp[a] + zn[a]) - 6.0 * m[a];
          out[a] = m[a] + alpha * lap;
        }
        out[1] += buoy * conc_at_yface(s, cr, i, j, k);
        if (i >= d.nx - 1) out[0] = 0.0;
        if (k >= d.nz - 1) out[2] = 0.0;
        s->u2[c] = out[0]; s->v2[c] = out[1]; s->w2[c] = out[2];
      }
  double *t;
  t = s->u; s->u = s->u2; s->u2 = t;
  t = s->v; s->v = s->v2; s->v2 = t;
  t = s->w; s->w = s->w2; s->w2 = t;
}

/* ---- FLUID_CASSON_FORCES (fluid.ts:377-511), parts (1) and (3), with the
 * reference's own physics (the coupling's physics flag):
 *   neighbours clamped to the edge (loadMac, fluid.ts:387-392);
 *   gamma = 50 sqrt(|du|^2 + |dv|^2 + |dw|^2 + 1e-6), du = mac(i+1) - mac(i-1)
 *     (all three components) etc. (fluid.ts:451-454);
 *   c = clamp(intact concentration, 0, 1.5) (fluid.ts:447-448; the dye, or
 *     the 4-channel concentration's .r);
 *   alpha = clamp((sqrt(1e-4 (1 + 1.35 c)) + sqrt(0.0036 scale c^2 /
 *     max(gamma + 0.4, 0.4)))^2, 1e-4, 0.0044) (fluid.ts:458-462);
 *   mac += alpha (sum of 6 neighbours - 6 mac) (fluid.ts:464-465);
 *   y-faces: rho' = 1.24 c_face + 0.26 lysed_face (0 without the 4-channel
 *     concentration) sampled at the face;
 *     if rho' > 0.0008: v += -buoyancy 0.345 1.24 tanh(rho'/1.24)
 *     smoothstep(0.002, 0.024, max(0, surface - y_face)) dt (fluid.ts:477-486);
 *   wall faces zeroed (fluid.ts:506-507).
 * The scene's terms (2), (4) and the clamp follow in casson_scene. */
static double mac_clamped(const double *f, dims3 d, int i, int j, int k) {
  return f[IDX(d, clampi(i, 0, d.nx - 1), clampi(j, 0, d.ny - 1), clampi(k, 0, d.nz - 1))];
}
static void casson_ref(orc_state *s, double dt) {
  dims3 d = s->d[0];
  const double scale = s->wp.casson_scale, buoy = s->wp.buoyancy;
  const double ytop = s->wp.surface_y, y0 = ytop - s->wp.depth;
  dt = clampd(dt, 0.0, 0.025);
#pragma omp parallel for
  for (int k = 0; k < d.nz; ++k)
    for (int j = 0; j < d.ny; ++j)
      for (int i = 0; i < d.nx; ++i) {
        long c = IDX(d, i, j, k);
        const double *F[3] = {s->u, s->v, s->w};
        double m[3], xp[3], xn[3], yp[3], yn[3], zp[3], zn[3];
        for (int a = 0; a < 3; ++a) {
          m[a] = F[a][c];
          xp[a] = mac_clamped(F[a], d, i + 1, j, k);
          xn[a] = mac_clamped(F[a], d, i - 1, j, k);
          yp[a] = mac_clamped(F[a], d, i, j + 1, k);
          yn[a] = mac_clamped(F[a], d, i, j - 1, k);
          zp[a] = mac_clamped(F[a], d, i, j, k + 1);
          zn[a] = mac_clamped(F[a], d, i, j, k - 1);
        }
        double ss = 0.0;
        for (int a = 0; a < 3; ++a) {
          double du = xp[a] - xn[a], dv = yp[a] - yn[a], dw = zp[a] - zn[a];
          ss += du * du + dv * dv + dw * dw;
        }
        double gam = 50.0 * sqrt(ss + 1e-6);
        const double *ci = intact_of(s), *cl = lysed_of(s);
        double cr = clampd(conc_at_cell(s, ci, i, j, k), 0.0, 1.5);
        double sq_muinf = sqrt(0.00010 * (1.0 + 1.35 * cr));
        double tau = 0.0036 * scale * cr * cr;
        double yterm = sqrt(tau / fmax(gam + 0.40, 0.40));
        double alpha = clampd((sq_muinf + yterm) * (sq_muinf + yterm), 0.00010, 0.0044);
        double out[3];
        for (int a = 0; a < 3; ++a) {
          double lap = (xp[a] + xn[a] + yp[a] + yn[a] + zp[a] + zn[a]) - 6.0 * m[a];
          out[a] = m[a] + alpha * lap;
        }
        double rho = 1.24 * conc_at_yface(s, ci, i, j, k) +
                     (cl ? 0.26 * conc_at_yface(s, cl, i, j, k) : 0.0);
        if (rho > 0.0008) {
          double yf = y0 + (j + 1.0) / d.ny * s->wp.depth;
          double gate = sstep(0.002, 0.024, fmax(0.0, ytop - yf));
          out[1] += -buoy * 0.345 * (1.24 * tanh(rho / 1.24)) * gate * dt;
        }
        if (i >= d.nx - 1) out[0] = 0.0;
        if (k >= d.nz - 1) out[2] = 0.0;
        s->u2[c] = out[0]; s->v2[c] = out[1]; s->w2[c] = out[2];
      }
  double *t;
  t = s->u; s->u = s->u2; s->u2 = t;
  t = s->v; s->v = s->v2; s->v2 = t;
  t = s->w; s->w = s->w2; s->w2 = t;
}

/* ---- FLUID_MAC_RHS (fluid.ts:517-555): D = (uR-uL)+(vT-vB)+(wF-wB),
 * solid wall faces 0, top face free; on the top row D -= p_top
 * (fluid.ts:548-551), the wave coupling's lid pressure, else the flat lid's
 * p_top = 0. */
static void divergence(const orc_state *s, const double *u, const double *v,
                       const double *w, double *out) {
  dims3 d = s->d[0];
#pragma omp parallel for
  for (int k = 0; k < d.nz; ++k)
    for (int j = 0; j < d.ny; ++j)
      for (int i = 0; i < d.nx; ++i) {
        long c = IDX(d, i, j, k);
        double uR = (i >= d.nx - 1) ? 0.0 : u[c];
        double uL = (i == 0) ? 0.0 : u[IDX(d, i - 1, j, k)];
        double vT = v[c];
        double vB = (j == 0) ? 0.0 : v[IDX(d, i, j - 1, k)];
        double wF = (k >= d.nz - 1) ? 0.0 : w[c];
        double wB = (k == 0) ? 0.0 : w[IDX(d, i, j, k - 1)];
        double rhs = (uR - uL) + (vT - vB) + (wF - wB);
        if (s->wave_on && j == d.ny - 1) rhs -= s->ptop[(long)k * d.nx + i];
        out[c] = rhs;
      }
}

/* ---- evalMgStencil (fluid.ts:561-603): neighbour sum with Neumann walls
 * (neighbour and diagonal term dropped) and the top face always in the
 * diagonal (Dirichlet 0 above the top row).  WORKLOAD: on level l the top
 * row's diagonal term is topc(l) = 2^(l+1)/(2^l + 1) instead of 1, which puts
 * the zero-pressure plane half a level-0 cell above the top row on every
 * level (linear ghost extrapolation); level 0 keeps the reference's 1.  With
 * the reference's 1 on every level the plane rises by one coarse cell per
 * level and the V-cycle stalls near 0.7 per cycle. */
static double level_h2(int l) { double s = (double)(1 << l); return s * s; }

double orc_top_coeff(int level) {
  double s = (double)(1 << level);
  return 2.0 * s / (s + 1.0);
}

static void stencil(const double *x, dims3 d, double topc, int i, int j, int k,
                    double *sum, double *diag) {
  double sn = 0.0, dg = (j < d.ny - 1) ? 1.0 : topc;
  if (i > 0) { sn += x[IDX(d, i - 1, j, k)]; dg += 1.0; }
  if (i < d.nx - 1) { sn += x[IDX(d, i + 1, j, k)]; dg += 1.0; }
  if (j > 0) { sn += x[IDX(d, i, j - 1, k)]; dg += 1.0; }
  if (j < d.ny - 1) { sn += x[IDX(d, i, j + 1, k)]; }
  if (k > 0) { sn += x[IDX(d, i, j, k - 1)]; dg += 1.0; }
  if (k < d.nz - 1) { sn += x[IDX(d, i, j, k + 1)]; dg += 1.0; }
  *sum = sn;
  *diag = dg;
}


/* ---- MG_RBGS_RED / MG_RBGS_BLACK (fluid.ts:645-681): colour
 * (i+j+k)&1 == parity, x <- (sum - h^2 b)/diag.  The reference ping-pongs
 * two textures per pass, copying the other colour; updating in place is the
 * same map because a colour only reads the other colour. */
static void rbgs_pass(double *x, const double *b, dims3 d, double h2, double topc,
                      double omega, int parity) {
#pragma omp parallel for
  for (int k = 0; k < d.nz; ++k)
    for (int j = 0; j < d.ny; ++j)
      for (int i = ((parity + j + k) & 1); i < d.nx; i += 2) {
        double sn, dg;
        stencil(x, d, topc, i, j, k, &sn, &dg);
        double gs = (sn - h2 * b[IDX(d, i, j, k)]) / dg;
        /* WORKLOAD (mixed mode only): over-relaxed update x + omega (gs - x) */
        if (omega != 1.0) gs = x[IDX(d, i, j, k)] + omega * (gs - x[IDX(d, i, j, k)]);
        x[IDX(d, i, j, k)] = gs;
      }
}

static double smoother_omega(const orc_state *s) { return s->p.omega > 0.0 ? s->p.omega : 1.0; }

static void smooth(double *x, const double *b, dims3 d, int l, int sweeps, double omega) {
  double h2 = level_h2(l), topc = orc_top_coeff(l);
  for (int s = 0; s < sweeps; ++s) {
    rbgs_pass(x, b, d, h2, topc, omega, 0);
    rbgs_pass(x, b, d, h2, topc, omega, 1);
  }
}

/* ---- MG_RESIDUAL (fluid.ts:744-770): r = b - (sum - diag x)/h^2 */
static void residual(const double *x, const double *b, double *r, dims3 d, int l) {
  double h2 = level_h2(l), topc = orc_top_coeff(l);
#pragma omp parallel for
  for (int k = 0; k < d.nz; ++k)
    for (int j = 0; j < d.ny; ++j)
      for (int i = 0; i < d.nx; ++i) {
        double sn, dg;
        stencil(x, d, topc, i, j, k, &sn, &dg);
        long c = IDX(d, i, j, k);
        r[c] = b[c] - (sn - dg * x[c]) / h2;
      }
}

/* ---- MG_RESTRICT (fluid.ts:775-800): coarse = 1/8 sum of the 2x2x2
 * children */
static void restrict_op(const double *rf, dims3 df, double *bc, dims3 dc) {
#pragma omp parallel for
  for (int k = 0; k < dc.nz; ++k)
    for (int j = 0; j < dc.ny; ++j)
      for (int i = 0; i < dc.nx; ++i) {
        double sum = 0.0;
        for (int dz = 0; dz < 2; ++dz)
          for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx)
              sum += rf[IDX(df, 2 * i + dx, 2 * j + dy, 2 * k + dz)];
        bc[IDX(dc, i, j, k)] = sum * 0.125;
      }
}

/* ---- MG_PROLONG_ADD (fluid.ts:806-855): cell-centred trilinear
 * interpolation of the coarse error, indices clamped at the walls and 0
 * above the top row (loadCoarseErr :814-821). */
static double coarse_err(const double *e, dims3 dc, int i, int j, int k) {
  if (j >= dc.ny) return 0.0;
  i = clampi(i, 0, dc.nx - 1);
  j = clampi(j, 0, dc.ny - 1);
  k = clampi(k, 0, dc.nz - 1);
  return e[IDX(dc, i, j, k)];
}

static void prolong_add(double *xf, dims3 df, const double *ec, dims3 dc) {
#pragma omp parallel for
  for (int k = 0; k < df.nz; ++k)
    for (int j = 0; j < df.ny; ++j)
      for (int i = 0; i < df.nx; ++i) {
        double t[3] = {(i + 0.5) / df.nx * dc.nx - 0.5, (j + 0.5) / df.ny * dc.ny - 0.5,
                       (k + 0.5) / df.nz * dc.nz - 0.5};
        int b[3];
        double f[3];
        for (int a = 0; a < 3; ++a) {
          double fl = floor(t[a]);
          b[a] = (int)fl;
          f[a] = t[a] - fl;
        }
        double c000 = coarse_err(ec, dc, b[0], b[1], b[2]);
        double c100 = coarse_err(ec, dc, b[0] + 1, b[1], b[2]);
        double c010 = coarse_err(ec, dc, b[0], b[1] + 1, b[2]);
        double c110 = coarse_err(ec, dc, b[0] + 1, b[1] + 1, b[2]);
        double c001 = coarse_err(ec, dc, b[0], b[1], b[2] + 1);
        double c101 = coarse_err(ec, dc, b[0] + 1, b[1], b[2] + 1);
        double c011 = coarse_err(ec, dc, b[0], b[1] + 1, b[2] + 1);
        double c111 = coarse_err(ec, dc, b[0] + 1, b[1] + 1, b[2] + 1);
        double x00 = mixd(c000, c100, f[0]), x10 = mixd(c010, c110, f[0]);
        double x01 = mixd(c001, c101, f[0]), x11 = mixd(c011, c111, f[0]);
        double corr = mixd(mixd(x00, x10, f[1]), mixd(x01, x11, f[1]), f[2]);
        xf[IDX(df, i, j, k)] += corr;
      }
}

/* ---- coarsest-level solve.  The reference runs 6 RBGS sweeps from zero in
 * one workgroup (MG_COARSE_SOLVE_L3, fluid.ts:687-739).  WORKLOAD: an exact
 * FP64 solve of (sum - diag x) = h^2 b.  The oracle solves it with
 * Jacobi-preconditioned conjugate gradients to a 1e-14 relative residual,
 * a method independent of the device's eigen-decomposition solve. */
static void apply_A(const double *x, double *y, dims3 d, double topc) {
  /* A = diag - neighbours (symmetric positive definite) */
#pragma omp parallel for
  for (int k = 0; k < d.nz; ++k)
    for (int j = 0; j < d.ny; ++j)
      for (int i = 0; i < d.nx; ++i) {
        double sn, dg;
        stencil(x, d, topc, i, j, k, &sn, &dg);
        y[IDX(d, i, j, k)] = dg * x[IDX(d, i, j, k)] - sn;
      }
}

static void coarse_solve(orc_state *s, int L) {
  dims3 d = s->d[L];
  long n = s->n[L];
  double h2 = level_h2(L), topc = orc_top_coeff(L);
  double *x = s->mgx[L], *r = s->cg[0], *z = s->cg[1], *q = s->cg[2], *Aq = s->cg[3];
  const double *b = s->mgb[L];
  /* A x = -h^2 b */
  double bn = 0.0;
  for (long c = 0; c < n; ++c) { x[c] = 0.0; r[c] = -h2 * b[c]; bn += r[c] * r[c]; }
  if (bn == 0.0) return;
  for (int k = 0; k < d.nz; ++k)
    for (int j = 0; j < d.ny; ++j)
      for (int i = 0; i < d.nx; ++i) {
        double sn, dg;
        stencil(r, d, topc, i, j, k, &sn, &dg);
        long c = IDX(d, i, j, k);
        z[c] = r[c] / dg;
      }
  memcpy(q, z, sizeof(double) * n);
  double rz = 0.0;
  for (long c = 0; c < n; ++c) rz += r[c] * z[c];
  for (int it = 0; it < 20000; ++it) {
    apply_A(q, Aq, d, topc);
    double qAq = 0.0;
    for (long c = 0; c < n; ++c) qAq += q[c] * Aq[c];
    double a = rz / qAq;
    double rn = 0.0;
    for (long c = 0; c < n; ++c) { x[c] += a * q[c]; r[c] -= a * Aq[c]; rn += r[c] * r[c]; }
    if (rn <= 1e-28 * bn) break;
    double rz2 = 0.0;
    for (int k = 0; k < d.nz; ++k)
      for (int j = 0; j < d.ny; ++j)
        for (int i = 0; i < d.nx; ++i) {
          double sn, dg;
          stencil(r, d, topc, i, j, k, &sn, &dg);
          long c = IDX(d, i, j, k);
          z[c] = r[c] / dg;
          rz2 += r[c] * z[c];
        }
    double beta = rz2 / rz;
    rz = rz2;
    for (long c = 0; c < n; ++c) q[c] = z[c] + beta * q[c];
  }
}

/* ---- one V-cycle: the dispatch schedule of BloodDropEngine.ts:1942-2059
 * (smooth, residual, restrict down; coarse solve; prolong, smooth up; the
 * level>=1 unknowns start from zero as MG_RBGS_RED_ZERO, fluid.ts:605-641).
 * WORKLOAD: nu1/nu2 sweeps on every level (reference: 1/1 on L0, 2/2 below),
 * exact coarse solve. */
static void vcycle(orc_state *s) {
  int L = s->p.levels;
  smooth(s->pr, s->rhs, s->d[0], 0, s->p.nu1, smoother_omega(s));
  residual(s->pr, s->rhs, s->mgr[0], s->d[0], 0);
  restrict_op(s->mgr[0], s->d[0], s->mgb[1], s->d[1]);
  for (int l = 1; l < L; ++l) {
    memset(s->mgx[l], 0, sizeof(double) * s->n[l]);
    smooth(s->mgx[l], s->mgb[l], s->d[l], l, s->p.nu1, smoother_omega(s));
    residual(s->mgx[l], s->mgb[l], s->mgr[l], s->d[l], l);
    restrict_op(s->mgr[l], s->d[l], s->mgb[l + 1], s->d[l + 1]);
  }
  coarse_solve(s, L);
  for (int l = L - 1; l >= 1; --l) {
    prolong_add(s->mgx[l], s->d[l], s->mgx[l + 1], s->d[l + 1]);
    smooth(s->mgx[l], s->mgb[l], s->d[l], l, s->p.nu2, smoother_omega(s));
  }
  prolong_add(s->pr, s->d[0], s->mgx[1], s->d[1]);
  smooth(s->pr, s->rhs, s->d[0], 0, s->p.nu2, smoother_omega(s));
}

static double rel_residual(orc_state *s) {
  residual(s->pr, s->rhs, s->mgr[0], s->d[0], 0);
  double rn = 0.0, bn = 0.0;
  for (long c = 0; c < s->n[0]; ++c) {
    rn += s->mgr[0][c] * s->mgr[0][c];
    bn += s->rhs[c] * s->rhs[c];
  }
  return bn > 0.0 ? sqrt(rn / bn) : 0.0;
}

/* ---- FLUID_MAC_PROJECT (fluid.ts:863-911): face -= p(+1) - p, wall faces
 * 0, top face against the lid's ghost p_top (fluid.ts:893-898; the wave
 * coupling's, else 0 for the flat lid).  The reference's +-2 clamp (lattice
 * units, fluid.ts:908) applies with the coupling's physics flag only: it is
 * not part of the workload. */
static void project(orc_state *s) {
  dims3 d = s->d[0];
  const int clamp2 = s->wave_on && s->wp.physics;
#pragma omp parallel for
  for (int k = 0; k < d.nz; ++k)
    for (int j = 0; j < d.ny; ++j)
      for (int i = 0; i < d.nx; ++i) {
        long c = IDX(d, i, j, k);
        double pc = s->pr[c];
        if (i < d.nx - 1) s->u[c] -= s->pr[IDX(d, i + 1, j, k)] - pc; else s->u[c] = 0.0;
        double ptop = s->wave_on ? s->ptop[(long)k * d.nx + i] : 0.0;
        if (j < d.ny - 1) s->v[c] -= s->pr[IDX(d, i, j + 1, k)] - pc; else s->v[c] -= ptop - pc;
        if (k < d.nz - 1) s->w[c] -= s->pr[IDX(d, i, j, k + 1)] - pc; else s->w[c] = 0.0;
        if (clamp2) {  /* the reference physics' +-2 clamp (fluid.ts:908) */
          s->u[c] = clampd(s->u[c], -2.0, 2.0);
          s->v[c] = clampd(s->v[c], -2.0, 2.0);
          s->w[c] = clampd(s->w[c], -2.0, 2.0);
        }
      }
}

/* ---- CONC_PREDICT (advect.ts:28-72): RK2 semi-Lagrangian predictor on the
 * centred velocity, uvw clamped to [0.001, 0.999].  WORKLOAD: the dye is a
 * cell field of the MAC grid; the static / single-stage speed shortcuts
 * (advect.ts:54-63) and the wall/top zeroing (advect.ts:42-45) are not part
 * of the workload: they apply with the coupling's physics flag only, as does
 * the resolve's zeroing of the same cells (advect.ts:100-103). */
static int conc_wall(dims3 d, int i, int j, int k) {
  return i == 0 || i >= d.nx - 1 || k == 0 || k >= d.nz - 1 || j >= d.ny - 1;
}
static void dye_predict(orc_state *s, double dt) {
  dims3 d = s->d[0];
  double dom[3] = {d.nx * s->p.dx, d.ny * s->p.dx, d.nz * s->p.dx};
  const int ref = s->wave_on && s->wp.physics;
#pragma omp parallel for
  for (int k = 0; k < d.nz; ++k)
    for (int j = 0; j < d.ny; ++j)
      for (int i = 0; i < d.nx; ++i) {
        double uvw[3] = {(i + 0.5) / d.nx, (j + 0.5) / d.ny, (k + 0.5) / d.nz};
        double v0[3] = {sample_cell(s->cu, d, uvw[0], uvw[1], uvw[2]),
                        sample_cell(s->cv, d, uvw[0], uvw[1], uvw[2]),
                        sample_cell(s->cw, d, uvw[0], uvw[1], uvw[2])};
        double mid[3], back[3];
        if (ref) {
          /* the reference's branches: lateral walls and the top row zeroed
           * (advect.ts:42-45); a static copy below speed^2 2.5e-5
           * (advect.ts:54-58); a single-stage trace below 0.012 (:60-62) */
          long c = IDX(d, i, j, k);
          double sp2 = v0[0] * v0[0] + v0[1] * v0[1] + v0[2] * v0[2];
          if (conc_wall(d, i, j, k)) { s->dpred[c] = 0.0; continue; }
          if (sp2 < 2.5e-5) { s->dpred[c] = s->dye[c]; continue; }
          if (sp2 < 0.012) {
            for (int a = 0; a < 3; ++a) back[a] = clampd(uvw[a] - (v0[a] * dt) / dom[a], 0.001, 0.999);
            s->dpred[c] = sample_cell(s->dye, d, back[0], back[1], back[2]);
            continue;
          }
        }
        for (int a = 0; a < 3; ++a) mid[a] = clampd(uvw[a] - 0.5 * (v0[a] * dt) / dom[a], 0.001, 0.999);
        double vm[3] = {sample_cell(s->cu, d, mid[0], mid[1], mid[2]),
                        sample_cell(s->cv, d, mid[0], mid[1], mid[2]),
                        sample_cell(s->cw, d, mid[0], mid[1], mid[2])};
        for (int a = 0; a < 3; ++a) back[a] = clampd(uvw[a] - (vm[a] * dt) / dom[a], 0.001, 0.999);
        s->dpred[IDX(d, i, j, k)] = sample_cell(s->dye, d, back[0], back[1], back[2]);
      }
}

/* ---- CONC_RESOLVE's MacCormack core (advect.ts:196-222): forward
 * estimate from the predictor at uvw + v0 dt, correction
 * pred + strength (orig - estimate), clamped to the min/max of the 8 source
 * texels around the backtrace point.  WORKLOAD: the diffusion/dilution and
 * parcel/bed/oxygenation parts (advect.ts:224-296) are not part of it. */
static void dye_resolve(orc_state *s, double dt) {
  dims3 d = s->d[0];
  double dom[3] = {d.nx * s->p.dx, d.ny * s->p.dx, d.nz * s->p.dx};
  double mc = s->p.maccormack;
  double *out = s->u2; /* scratch */
  const int ref = s->wave_on && s->wp.physics;
#pragma omp parallel for
  for (int k = 0; k < d.nz; ++k)
    for (int j = 0; j < d.ny; ++j)
      for (int i = 0; i < d.nx; ++i) {
        long c = IDX(d, i, j, k);
        if (ref && conc_wall(d, i, j, k)) { out[c] = 0.0; continue; } /* advect.ts:100-103 */
        double uvw[3] = {(i + 0.5) / d.nx, (j + 0.5) / d.ny, (k + 0.5) / d.nz};
        double v0[3] = {sample_cell(s->cu, d, uvw[0], uvw[1], uvw[2]),
                        sample_cell(s->cv, d, uvw[0], uvw[1], uvw[2]),
                        sample_cell(s->cw, d, uvw[0], uvw[1], uvw[2])};
        double mid[3], back[3], fwd[3];
        for (int a = 0; a < 3; ++a) mid[a] = clampd(uvw[a] - 0.5 * (v0[a] * dt) / dom[a], 0.001, 0.999);
        double vm[3] = {sample_cell(s->cu, d, mid[0], mid[1], mid[2]),
                        sample_cell(s->cv, d, mid[0], mid[1], mid[2]),
                        sample_cell(s->cw, d, mid[0], mid[1], mid[2])};
        for (int a = 0; a < 3; ++a) {
          back[a] = clampd(uvw[a] - (vm[a] * dt) / dom[a], 0.001, 0.999);
          fwd[a] = clampd(uvw[a] + (v0[a] * dt) / dom[a], 0.001, 0.999);
        }
        double est = sample_cell(s->dpred, d, fwd[0], fwd[1], fwd[2]);
        double corr = s->dpred[c] + mc * (s->dye[c] - est);
        int bi = (int)floor(back[0] * d.nx - 0.5);
        int bj = (int)floor(back[1] * d.ny - 0.5);
        int bk = (int)floor(back[2] * d.nz - 0.5);
        double lo = 1e300, hi = -1e300;
        for (int dz = 0; dz < 2; ++dz)
          for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
              double sv = s->dye[IDX(d, clampi(bi + dx, 0, d.nx - 1), clampi(bj + dy, 0, d.ny - 1),
                                     clampi(bk + dz, 0, d.nz - 1))];
              lo = fmin(lo, sv);
              hi = fmax(hi, sv);
            }
        out[c] = clampd(corr, lo, hi);
      }
  memcpy(s->dye, out, sizeof(double) * s->n[0]);
}

/* ==== free-surface wave coupling (wave.ts:15-388, fluid.ts:16-78) ==== */

static double sstep(double e0, double e1, double x) { return orc_smoothstep(e0, e1, x); }
static const double PI_D = 3.14159265358979323846;

/* loadWaveClamped (wave.ts:63-67): texel (x, y) channel ch, indices clamped */
static double wv_load(const double *w, int n, int x, int y, int ch) {
  x = clampi(x, 0, n - 1);
  y = clampi(y, 0, n - 1);
  return w[4 * ((long)y * n + x) + ch];
}

/* textureSampleLevel of the wave state with the linear clamp-to-edge sampler
 * (BloodDropEngine.ts:497-505): bilinear in texel space t = uv n - 1/2 */
static double wv_sample(const double *w, int n, double u, double v, int ch) {
  double tx = u * n - 0.5, ty = v * n - 0.5;
  double fx0 = floor(tx), fy0 = floor(ty);
  int x0 = (int)fx0, y0 = (int)fy0;
  double fx = tx - fx0, fy = ty - fy0;
  double a = mixd(wv_load(w, n, x0, y0, ch), wv_load(w, n, x0 + 1, y0, ch), fx);
  double b = mixd(wv_load(w, n, x0, y0 + 1, ch), wv_load(w, n, x0 + 1, y0 + 1, ch), fx);
  return mixd(a, b, fy);
}

/* WAVE_SPATIAL_FORCE (wave.ts:56-178) into wz (eta + i deta/dt) and wfilm */
static void wave_force(orc_state *s, double dt) {
  const int n = s->wp.n;
  const double E = s->wp.half_extent, hx = s->wp.half_xz;
  const double time_scale = dt / 0.0166667;
  const double lap_scale = ((double)n / 256.0) * ((double)n / 256.0);
#pragma omp parallel for
  for (int iy = 0; iy < n; ++iy)
    for (int ix = 0; ix < n; ++ix) {
      const double *W = s->wave;
      double eta = wv_load(W, n, ix, iy, 0), vel = wv_load(W, n, ix, iy, 1);
      double film = wv_load(W, n, ix, iy, 3);
      double raw_lap = 0.20 * (wv_load(W, n, ix, iy + 1, 3) + wv_load(W, n, ix, iy - 1, 3) +
                               wv_load(W, n, ix + 1, iy, 3) + wv_load(W, n, ix - 1, iy, 3)) +
                       0.05 * (wv_load(W, n, ix + 1, iy + 1, 3) + wv_load(W, n, ix - 1, iy + 1, 3) +
                               wv_load(W, n, ix + 1, iy - 1, 3) + wv_load(W, n, ix - 1, iy - 1, 3)) -
                       film;
      double lap_film = raw_lap * lap_scale;
      /* waveUVToWorldXZ (common.ts:69-71) at the texel centre */
      double wx = ((ix + 0.5) / n) * (2.0 * E) - E, wz = ((iy + 0.5) / n) * (2.0 * E) - E;
      double force_drops = 0.0, active = 0.0, max_drain = 0.0;
      for (int q = 0; q < MAX_DROPS; ++q) {
        const orc_drop *dr = &s->drops[q];
        if (!(dr->seed > 0.0 && dr->drain > 0.0005)) continue;
        max_drain = fmax(max_drain, dr->drain);
        double rd = fmax(dr->radius, 0.018);
        double speed = clampd(-dr->vel_y, 0.06, 0.65);
        double r = hypot(wx - dr->x, wz - dr->z);
        double rn = r / clampd(rd * 1.05, 0.024, 0.082);
        double emerg = (dr->y - s->wp.surface_y) / fmax(rd, 0.018);
        double excav = sstep(-0.78, -0.04, emerg);
        double bell = sin(clampd((1.05 - emerg) / 1.83, 0.0, 1.0) * 3.14159265) * excav;
        double shape = exp(-rn * rn) * (1.0 - rn * rn);
        double dyn = (6.8 * bell * (rd / 0.056)) * (0.72 + 2.65 * speed);
        double rnc = r / clampd(rd * 0.44, 0.013, 0.030);
        double cap_win = sstep(-0.35, 0.28, emerg) * (1.0 - sstep(0.68, 1.06, emerg));
        double cap_shape = exp(-rnc * rnc) * (1.0 - rnc * rnc);
        double cap = (4.6 * cap_win * (0.60 + 0.40 * (rd / 0.056))) * (0.65 + 2.20 * speed);
        force_drops = force_drops - dyn * shape - cap * cap_shape;
        active = fmax(active, excav * exp(-0.5 * rn * rn));
        double r_stain = fmax(0.018, rd * 0.60);
        double stain = exp(-(r * r) / (r_stain * r_stain)) * clampd(dr->drain * 0.42, 0.0, 1.0);
        film = fmax(film, stain * 0.65);
      }
      double m = s->wp.marangoni;
      double eff_lap = mixd(fmax(lap_film, -0.0015), lap_film, active);
      double force_mar = m * 0.28 * eff_lap;
      film = clampd(film + m * 0.065 * raw_lap * time_scale, 0.0, 1.0);
      film = film * pow(0.88, time_scale);
      double force_sub = 0.0;
      double gate = clampd(active * 1.4 + sstep(0.0, 0.06, max_drain), 0.0, 1.0);
      double rw = hypot(wx, wz);
      if (gate > 0.001 && rw < hx * 0.94) {
        /* velMac .y at uvw (tankUV.x, 0.988, tankUV.y), trilinear clamp-to-edge */
        double vsub = sample_cell(s->v, s->d[0], (wx + hx) / (2.0 * hx), 0.988, (wz + hx) / (2.0 * hx));
        double taper = 1.0 - sstep(hx * 0.62, hx * 0.92, rw);
        force_sub = clampd(fmax(vsub, 0.0), 0.0, 0.48) * 0.42 * taper * gate;
      }
      vel = vel + (force_drops + force_mar + force_sub) * dt;
      double absorb = sstep(0.90, 1.54, rw);
      double rdamp = exp(-dt * (24.0 * absorb * absorb));
      double border = 1.0 - sstep(1.22, 1.56, rw);
      eta = clampd(eta * rdamp * border, -0.14, 0.14);
      vel = clampd(vel * rdamp * border, -1.8, 1.8);
      long t = (long)iy * n + ix;
      s->wz[2 * t] = eta;
      s->wz[2 * t + 1] = vel;
      s->wfilm[t] = film * border;
    }
}

/* Unitary radix-2 FFT (the reference's createRadix2Fft1DShader, wave.ts:180-
 * 252, is a decimation-in-time radix-2 with 1/sqrt(n) on every 1-D
 * transform): bit-reversed load, log2 n butterfly stages with exact
 * twiddles exp(sign 2 pi i k / span), natural-order result. */
void orc_fft(double *z, int n, int stride, int sign) {
  double *t = (double *)malloc(sizeof(double) * 2 * n);
  int logn = 0;
  while ((1 << logn) < n) ++logn;
  for (int i = 0; i < n; ++i) {
    int r = 0;
    for (int b = 0; b < logn; ++b) r |= ((i >> b) & 1) << (logn - 1 - b);
    t[2 * i] = z[2 * (long)r * stride];
    t[2 * i + 1] = z[2 * (long)r * stride + 1];
  }
  for (int s = 0; s < logn; ++s) {
    int half = 1 << s, span = half << 1;
    for (int blk = 0; blk < n; blk += span)
      for (int k = 0; k < half; ++k) {
        double ang = sign * 2.0 * PI_D * k / span;
        double c = cos(ang), sn = sin(ang);
        int i0 = blk + k, i1 = i0 + half;
        double br = t[2 * i1] * c - t[2 * i1 + 1] * sn, bi = t[2 * i1] * sn + t[2 * i1 + 1] * c;
        double ar = t[2 * i0], ai = t[2 * i0 + 1];
        t[2 * i0] = ar + br;
        t[2 * i0 + 1] = ai + bi;
        t[2 * i1] = ar - br;
        t[2 * i1 + 1] = ai - bi;
      }
  }
  double sc = 1.0 / sqrt((double)n);
  for (int i = 0; i < n; ++i) {
    z[2 * (long)i * stride] = t[2 * i] * sc;
    z[2 * (long)i * stride + 1] = t[2 * i + 1] * sc;
  }
  free(t);
}

/* WAVE_SPECTRAL_PROPAGATE (wave.ts:254-331) from wz into wz2 */
static void wave_propagate(orc_state *s, double dt) {
  const int n = s->wp.n, hn = n / 2;
  const double L = 2.0 * s->wp.half_extent, H = s->wp.depth;
#pragma omp parallel for
  for (int uy = 0; uy < n; ++uy)
    for (int ux = 0; ux < n; ++ux) {
      long t = (long)uy * n + ux;
      if ((ux == 0 && uy == 0) || ux == hn || uy == hn) {
        s->wz2[2 * t] = s->wz2[2 * t + 1] = 0.0;
        continue;
      }
      long tn = (long)((n - uy) & (n - 1)) * n + ((n - ux) & (n - 1));
      double pr = s->wz[2 * t], pi = s->wz[2 * t + 1], mr = s->wz[2 * tn], mi = s->wz[2 * tn + 1];
      double er = 0.5 * (pr + mr), ei = 0.5 * (pi - mi);   /* eta_hat */
      double vr = 0.5 * (pi + mi), vi = 0.5 * (-pr + mr);  /* w_hat */
      double kx = (ux > hn ? ux - n : ux) * (2.0 * PI_D / L);
      double kz = (uy > hn ? uy - n : uy) * (2.0 * PI_D / L);
      double k2 = kx * kx + kz * kz, k = sqrt(k2);
      double th = tanh(fmin(k * H, 12.0));
      double omega = sqrt(fmax((0.84 * k + 0.0024 * k * k2) * th, 1e-8));
      double damp = exp(-(0.36 + 0.00018 * k2) * dt) * (1.0 - sstep(335.0, 465.0, k));
      double c = cos(omega * dt), sn = sin(omega * dt);
      double enr = damp * (er * c + vr * (sn / omega)), eni = damp * (ei * c + vi * (sn / omega));
      double vnr = damp * (-er * (omega * sn) + vr * c), vni = damp * (-ei * (omega * sn) + vi * c);
      s->wz2[2 * t] = enr - vni;
      s->wz2[2 * t + 1] = eni + vnr;
    }
}

/* WAVE_FINALIZE (wave.ts:333-388) from wz2 and wfilm into the wave state */
static void wave_finalize(orc_state *s) {
  const int n = s->wp.n;
  const double E = s->wp.half_extent;
  const double lap_scale = ((double)n / 256.0) * ((double)n / 256.0);
#define ETA(x, y) s->wz2[2 * ((long)clampi((y), 0, n - 1) * n + clampi((x), 0, n - 1))]
#pragma omp parallel for
  for (int iy = 0; iy < n; ++iy)
    for (int ix = 0; ix < n; ++ix) {
      long t = (long)iy * n + ix;
      double px = s->wz2[2 * t], py = s->wz2[2 * t + 1];
      double lap = (0.20 * (ETA(ix, iy + 1) + ETA(ix, iy - 1) + ETA(ix + 1, iy) + ETA(ix - 1, iy)) +
                    0.05 * (ETA(ix + 1, iy + 1) + ETA(ix - 1, iy + 1) + ETA(ix + 1, iy - 1) +
                            ETA(ix - 1, iy - 1)) - px) * lap_scale;
      double wx = ((ix + 0.5) / n) * (2.0 * E) - E, wz = ((iy + 0.5) / n) * (2.0 * E) - E;
      double border = 1.0 - sstep(1.22, 1.56, hypot(wx, wz));
      s->wave[4 * t] = clampd(px * border, -0.14, 0.14);
      s->wave[4 * t + 1] = clampd(py * border, -1.8, 1.8);
      s->wave[4 * t + 2] = lap * border;
      s->wave[4 * t + 3] = s->wfilm[t] * border;
    }
#undef ETA
}

/* one wave step in the pass order of BloodDropEngine.ts:1886-1914 */
static void wave_step(orc_state *s, double dt) {
  const int n = s->wp.n;
  dt = clampd(dt, 0.0, 0.025); /* viewportAndDt.z clamp (wave.ts:83) */
  wave_force(s, dt);
#pragma omp parallel for
  for (int y = 0; y < n; ++y) orc_fft(s->wz + 2 * (long)y * n, n, 1, -1);
#pragma omp parallel for
  for (int x = 0; x < n; ++x) orc_fft(s->wz + 2 * (long)x, n, n, -1);
  wave_propagate(s, dt);
#pragma omp parallel for
  for (int x = 0; x < n; ++x) orc_fft(s->wz2 + 2 * (long)x, n, n, +1);
#pragma omp parallel for
  for (int y = 0; y < n; ++y) orc_fft(s->wz2 + 2 * (long)y * n, n, 1, +1);
  wave_finalize(s);
}

/* evalTopDirichletPressure (fluid.ts:16-78) over the lid plane */
static void top_pressure(orc_state *s, double dt) {
  dims3 d = s->d[0];
  const double hx = s->wp.half_xz, E = s->wp.half_extent;
  dt = clampd(dt, 0.0, 0.025);
  const double dt_h = dt / ((2.0 * hx) / d.nx);
#pragma omp parallel for
  for (int iz = 0; iz < d.nz; ++iz)
    for (int ix = 0; ix < d.nx; ++ix) {
      double wx = mixd(-hx, hx, (ix + 0.5) / d.nx), wz = mixd(-hx, hx, (iz + 0.5) / d.nz);
      double su = clampd((wx + E) / (2.0 * E), 0.002, 0.998);
      double sv = clampd((wz + E) / (2.0 * E), 0.002, 0.998);
      double eta = wv_sample(s->wave, s->wp.n, su, sv, 0);
      double lap = wv_sample(s->wave, s->wp.n, su, sv, 2);
      double pk = 2.45 * (0.155 * eta + 0.00095 * (-18.4 * lap));
      for (int q = 0; q < MAX_DROPS; ++q) {
        const orc_drop *dr = &s->drops[q];
        if (!(dr->seed > 0.0 && dr->drain > 0.0005)) continue;
        double ex = wx - dr->x, ez = wz - dr->z;
        double r = hypot(ex, ez), th = atan2(ez, ex);
        double azim = 1.0 + 0.022 * cos(4.0 * th + dr->seed * 1.73) +
                      0.016 * cos(6.0 * th - dr->seed * 2.31) + 0.010 * sin(8.0 * th + dr->seed * 0.89);
        double slug = sstep(0.32, 0.92, dr->drain);
        double r_or = clampd(dr->radius * 1.12 + mixd(0.018, 0.032, slug), 0.030, 0.090) * azim;
        double rn = r / fmax(r_or, 1e-4);
        double speed = clampd(-dr->vel_y, 0.08, 0.65);
        double neck = (0.28 + 1.05 * rn * rn) * exp(-0.92 * rn * rn) * (1.0 - 0.38 * rn * rn);
        pk = pk + dr->drain * (0.014 + 0.048 * speed) * neck;
      }
      double taper = 1.0 - sstep(hx * 0.72, hx * 0.96, fmax(fabs(wx), fabs(wz)));
      s->ptop[(long)iz * d.nx + ix] = pk * dt_h * taper;
    }
}


/* The advection's scene terms (fluid.ts:198-321) on cell (i, j, k)'s damped
 * face values: the drops' wetting-neck momentum (:198-285: slug/stem draft on
 * the y-face, radial flare on the x- and z-faces, each with the m = 4, 6, 8
 * azimuthal modulation), the film's Marangoni pull and the kinematic top from
 * the wave (:287-310), the circular wall and floor sponges (:312-321).  World
 * positions by uvwToWorld (common.ts:45-64) over the tank [-hx, hx] x
 * [ytop - depth, ytop] x [-hx, hx], ytop the surface height. */
static double azim3(double th, double seed) {
  return 1.0 + 0.036 * cos(4.0 * th + seed * 1.73) + 0.026 * cos(6.0 * th - seed * 2.31) +
         0.018 * sin(8.0 * th + seed * 0.89);
}
static double neck_radial(double rr, double slug) {
  double sl = (0.14 + 1.00 * fmin(rr * rr, 1.20)) * (1.0 - sstep(0.78, 1.15, rr));
  double st = (0.68 + 0.32 * fmin(rr * rr, 1.0)) * (1.0 - sstep(0.74, 1.10, rr));
  return mixd(st, sl, slug);
}
static void advect_scene(const orc_state *s, int i, int j, int k, double dt, double *vx,
                         double *vy, double *vz) {
  dims3 d = s->d[0];
  const double hx = s->wp.half_xz, ytop = s->wp.surface_y, y0 = ytop - s->wp.depth;
  dt = clampd(dt, 0.0, 0.025);
  double cx = -hx + (i + 0.5) / d.nx * 2.0 * hx, cz = -hx + (k + 0.5) / d.nz * 2.0 * hx;
  double cy = y0 + (j + 0.5) / d.ny * s->wp.depth, fy = y0 + (j + 1.0) / d.ny * s->wp.depth;
  double xf = -hx + (i + 1.0) / d.nx * 2.0 * hx, zf = -hx + (k + 1.0) / d.nz * 2.0 * hx;
  double surf = s->wp.surface_y, active = 0.0;
  for (int q = 0; q < MAX_DROPS; ++q) {
    const orc_drop *dr = &s->drops[q];
    if (!(dr->seed > 0.0 && dr->drain > 0.0005)) continue;
    double ex = cx - dr->x, ez = cz - dr->z, r = hypot(ex, ez), th = atan2(ez, ex);
    double slug = sstep(0.14, 0.76, dr->drain);
    double rbase = clampd(dr->radius * mixd(0.80, 1.08, slug) + mixd(0.012, 0.030, slug), 0.018, 0.088);
    double radial = neck_radial(r / fmax(rbase * azim3(th, dr->seed), 1e-4), slug);
    double mask = radial * sstep(surf - 0.058, surf - 0.014, fy);
    double shear = mixd(1.0, 0.36, sstep(surf - 0.024, surf - 0.002, fy));
    active = fmax(active, radial);
    if (mask > 0.001) {
      double lobe = 1.0 + 0.036 * cos(4.0 * th + dr->seed * 1.73) + 0.026 * cos(6.0 * th - dr->seed * 2.31);
      double impact = clampd(-dr->vel_y, 0.10, 0.65);
      double slug_drive = clampd(pow(fmax(dr->drain, 0.0), 0.66) * 0.84, 0.0, 1.24);
      double stem_drive = clampd(0.045 + dr->drain * 0.22, 0.0, 0.20) * sstep(0.0, 0.05, dr->drain);
      double target = -clampd(0.14 + 0.64 * impact, 0.15, 0.39) * mixd(stem_drive, slug_drive, slug) *
                      lobe * shear;
      *vy = mixd(*vy, target, clampd(dt * mixd(14.0, 22.0, slug) * mask, 0.0, 1.0));
    }
    double vmxz = sstep(surf - 0.058, surf - 0.014, cy);
    if (vmxz > 0.001) {
      double flare = mixd(0.104 * (0.35 + 0.65 * slug), -0.020, sstep(surf - 0.066, surf - 0.012, cy)) *
                     clampd(dr->drain, 0.0, 1.3);
      double fxs[2] = {xf - dr->x, cx - dr->x}, fzs[2] = {ez, zf - dr->z};
      for (int f = 0; f < 2; ++f) {
        double rf = hypot(fxs[f], fzs[f]);
        if (!(rf > 1e-4 && rf < rbase * 1.34)) continue;
        double rrf = rf / fmax(rbase * azim3(atan2(fzs[f], fxs[f]), dr->seed), 1e-4);
        if (!(rrf < 1.22)) continue;
        double bf = clampd(dt * mixd(14.0, 22.0, slug) * neck_radial(rrf, slug) * vmxz, 0.0, 1.0);
        if (f == 0) *vx = mixd(*vx, (fxs[f] / rf) * flare, bf * 0.65);
        else *vz = mixd(*vz, (fzs[f] / rf) * flare, bf * 0.65);
      }
    }
  }
  if (j >= d.ny - 3) {
    double fade = (double)(j - (d.ny - 3) + 1) / 3.0, eps = 0.015, E = s->wp.half_extent;
    int n = s->wp.n;
#define UV(x) clampd(((x) + E) / (2.0 * E), 0.002, 0.998)
    double film = wv_sample(s->wave, n, UV(cx), UV(cz), 3), wvel = wv_sample(s->wave, n, UV(cx), UV(cz), 1);
    double gx = (wv_sample(s->wave, n, UV(cx + eps), UV(cz), 3) - wv_sample(s->wave, n, UV(cx - eps), UV(cz), 3)) / (2.0 * eps);
    double gz = (wv_sample(s->wave, n, UV(cx), UV(cz + eps), 3) - wv_sample(s->wave, n, UV(cx), UV(cz - eps), 3)) / (2.0 * eps);
#undef UV
    double cmc = 1.0 - sstep(0.45, 0.88, film);
    *vx += -s->wp.marangoni * 0.026 * cmc * gx * fade * dt;
    *vz += -s->wp.marangoni * 0.026 * cmc * gz * fade * dt;
    if (j >= d.ny - 2 && active < 0.02) {
      double calm = 1.0 - sstep(0.020, 0.14, fabs(wvel));
      double layer = (j == d.ny - 1) ? 0.72 : 0.26;
      *vy = mixd(*vy, clampd(wvel * 0.35, -0.045, 0.045), calm * layer);
    }
  }
  double wf = (1.0 - sstep(hx * 0.68, hx * 0.94, hypot(cx, cz))) * sstep(y0 + 0.004, y0 + 0.052, cy);
  double sponge = mixd(0.84, 1.0, wf);
  *vx *= sponge;
  *vy *= sponge;
  *vz *= sponge;
}

/* the stagnation factor of FLUID_MAC_TO_CENTER (fluid.ts:944-959):
 * mix(stagGate, 1, activeNeck), stagGate = mix(0.45, 1, smoothstep(0.002,
 * 0.018, surface - y)), activeNeck the largest over drops with seed > 0 and
 * drain > 0.008 of (1 - smoothstep(0.78 ra, 1.20 ra, r_xz)) smoothstep(0.008,
 * 0.065, drain), ra = clamp(1.15 radius + 0.032, 0.046, 0.092). */
static double top_stagnation_gate(const orc_state *s, int i, int j, int k) {
  dims3 d = s->d[0];
  const double hx = s->wp.half_xz, ytop = s->wp.surface_y, y0 = ytop - s->wp.depth;
  double px = -hx + (i + 0.5) / d.nx * 2.0 * hx, pz = -hx + (k + 0.5) / d.nz * 2.0 * hx;
  double py = y0 + (j + 0.5) / d.ny * s->wp.depth;
  double act = 0.0;
  for (int q = 0; q < MAX_DROPS; ++q) {
    const orc_drop *dr = &s->drops[q];
    if (!(dr->seed > 0.0 && dr->drain > 0.008)) continue;
    double ra = clampd(dr->radius * 1.15 + 0.032, 0.046, 0.092);
    double rd = hypot(px - dr->x, pz - dr->z);
    act = fmax(act, (1.0 - sstep(ra * 0.78, ra * 1.20, rd)) * sstep(0.008, 0.065, dr->drain));
  }
  double stag = mixd(0.45, 1.0, sstep(0.002, 0.018, fmax(0.0, ytop - py)));
  return mixd(stag, 1.0, act);
}

/* ---- FLUID_CURL (fluid.ts:331-369): omega = 0.5 (central differences of
 * the cell-centred velocity), neighbour indices clamped to the grid (grid
 * units, no 1/h), |omega| in the fourth slot. */
static void curl_pass(orc_state *s) {
  dims3 d = s->d[0];
#pragma omp parallel for
  for (int k = 0; k < d.nz; ++k)
    for (int j = 0; j < d.ny; ++j)
      for (int i = 0; i < d.nx; ++i) {
        long R = IDX(d, clampi(i + 1, 0, d.nx - 1), j, k), L = IDX(d, clampi(i - 1, 0, d.nx - 1), j, k);
        long U = IDX(d, i, clampi(j + 1, 0, d.ny - 1), k), D = IDX(d, i, clampi(j - 1, 0, d.ny - 1), k);
        long F = IDX(d, i, j, clampi(k + 1, 0, d.nz - 1)), B = IDX(d, i, j, clampi(k - 1, 0, d.nz - 1));
        double *o = s->curl + 4 * IDX(d, i, j, k);
        o[0] = 0.5 * ((s->cw[U] - s->cw[D]) - (s->cv[F] - s->cv[B]));
        o[1] = 0.5 * ((s->cu[F] - s->cu[B]) - (s->cw[R] - s->cw[L]));
        o[2] = 0.5 * ((s->cv[R] - s->cv[L]) - (s->cu[U] - s->cu[D]));
        o[3] = sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]);
      }
}

/* evalVortConf (fluid.ts:394-423): 0 within two cells of the side walls and
 * floor and three of the lid, or where |omega| <= 0.18; else
 * scale * smoothstep(0.18, 0.95, |omega|) / (1 + 0.28 |omega|) * (N x omega),
 * N the normalised central-difference gradient of |omega| (0 below 1e-5). */
static void vort_conf(const orc_state *s, int i, int j, int k, double scale, double f[3]) {
  dims3 d = s->d[0];
  f[0] = f[1] = f[2] = 0.0;
  if (i <= 1 || i >= d.nx - 2 || j <= 1 || j >= d.ny - 3 || k <= 1 || k >= d.nz - 2) return;
  const double *om = s->curl + 4 * IDX(d, i, j, k);
  if (om[3] <= 0.18) return;
  double gate = orc_smoothstep(0.18, 0.95, om[3]) / (1.0 + 0.28 * om[3]);
  double e[3] = {0.5 * (s->curl[4 * IDX(d, i + 1, j, k) + 3] - s->curl[4 * IDX(d, i - 1, j, k) + 3]),
                 0.5 * (s->curl[4 * IDX(d, i, j + 1, k) + 3] - s->curl[4 * IDX(d, i, j - 1, k) + 3]),
                 0.5 * (s->curl[4 * IDX(d, i, j, k + 1) + 3] - s->curl[4 * IDX(d, i, j, k - 1) + 3])};
  double el = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
  if (el <= 1e-5) return;
  for (int a = 0; a < 3; ++a) e[a] /= el;
  f[0] = scale * gate * (e[1] * om[2] - e[2] * om[1]);
  f[1] = scale * gate * (e[2] * om[0] - e[0] * om[2]);
  f[2] = scale * gate * (e[0] * om[1] - e[1] * om[0]);
}

/* ---- FLUID_CASSON_FORCES scene terms (fluid.ts:467-509), after the
 * viscous update and buoyancy: the confinement force averaged onto each face
 * with the +1 neighbour cell's (:467-472), the deep-column solenoidal
 * unfurling forcing at the cell centre (:488-503; the lysed-Hb channel of its
 * gate is 0 with one dye species), the +-1.8 clamp and the wall faces
 * zeroed (:505-507).  Positions as in advect_scene; dt in [0, 0.025]. */
static void casson_scene(orc_state *s, double dt) {
  dims3 d = s->d[0];
  const double hx = s->wp.half_xz, ytop = s->wp.surface_y, y0 = ytop - s->wp.depth;
  const double scale = s->wp.vort_conf;
  const double ts = s->wp.sim_time * 0.36 + s->wp.schlieren * 1.7;
  dt = clampd(dt, 0.0, 0.025);
#pragma omp parallel for
  for (int k = 0; k < d.nz; ++k)
    for (int j = 0; j < d.ny; ++j)
      for (int i = 0; i < d.nx; ++i) {
        long c = IDX(d, i, j, k);
        double f0[3], fx[3], fy[3], fz[3];
        vort_conf(s, i, j, k, scale, f0);
        vort_conf(s, clampi(i + 1, 0, d.nx - 1), j, k, scale, fx);
        vort_conf(s, i, clampi(j + 1, 0, d.ny - 1), k, scale, fy);
        vort_conf(s, i, j, clampi(k + 1, 0, d.nz - 1), scale, fz);
        double m[3] = {s->u[c] + 0.5 * (f0[0] + fx[0]) * dt, s->v[c] + 0.5 * (f0[1] + fy[1]) * dt,
                       s->w[c] + 0.5 * (f0[2] + fz[2]) * dt};
        double px = -hx + (i + 0.5) / d.nx * 2.0 * hx, pz = -hx + (k + 0.5) / d.nz * 2.0 * hx;
        double py = y0 + (j + 0.5) / d.ny * s->wp.depth;
        double below = ytop - py;
        const double *cl = lysed_of(s);
        double gate = sstep(0.11, 0.36, below) *
                      sstep(0.0008, 0.015, clampd(conc_at_cell(s, intact_of(s), i, j, k), 0.0, 1.5) +
                                               (cl ? 0.45 * conc_at_cell(s, cl, i, j, k) : 0.0));
        if (gate > 0.001) {
          double q1 = 6.4 * px + 4.0 * py + 5.2 * pz + ts;
          double q2 = -5.6 * px + 4.8 * py + 7.2 * pz - ts * 0.77;
          double q3 = 9.0 * px - 6.0 * py - 6.4 * pz + ts * 1.13;
          double k1[3] = {-6.25, 2.85, 5.50}, k2[3] = {-6.30, -1.26, -4.06}, k3[3] = {6.40, -0.96, 9.90};
          double a = 0.0028 * gate * dt;
          for (int e = 0; e < 3; ++e) {
            double sf = k1[e] * cos(q1) + k2[e] * cos(q2) + k3[e] * (0.42 * sin(q3));
            if (e == 1) sf *= sstep(0.16, 0.42, below);
            m[e] += sf * a;
          }
        }
        for (int e = 0; e < 3; ++e) m[e] = clampd(m[e], -1.8, 1.8);
        if (i >= d.nx - 1) m[0] = 0.0;
        if (k >= d.nz - 1) m[2] = 0.0;
        s->u[c] = m[0]; s->v[c] = m[1]; s->w[c] = m[2];
      }
}

/* the Casson pass: with the scene, the centre velocity of the advected faces
 * and its curl first (BloodDropEngine.ts:1922-1930), the scene terms last */
static void casson_stage(orc_state *s, double dt) {
  int scene = s->wave_on && s->wp.scene;
  if (scene) {
    face_to_centre(s);
    curl_pass(s);
  }
  if (s->wave_on && s->wp.physics) casson_ref(s, dt);
  else casson_buoyancy(s, dt);
  if (scene) casson_scene(s, dt);
}

/* ---- PARTICLE_ADVECT_SPLAT (particles.ts:189-463) ----
 * World box [-hx, hx] x [y_min, y_max] x [-hx, hx], y_max the surface height
 * and y_min = y_max - depth (worldToUVW, common.ts:56-58).  The two per-parcel
 * hash ranks fract(g 0.7548776662) and fract(g 0.6180339887 + 0.17) are formed
 * in single precision as the reference's f32 does (for g ~ 4e5 the FP64 value
 * differs by up to 0.02); the rest is FP64. */
typedef struct { double hx, ymin, ymax, t, kl, dt; } parcel_scene;

static void pc_uvw(const parcel_scene *S, const double p[3], double u[3]) {
  u[0] = (p[0] + S->hx) / (2.0 * S->hx);
  u[1] = (p[1] - S->ymin) / (S->ymax - S->ymin);
  u[2] = (p[2] + S->hx) / (2.0 * S->hx);
}

/* sampleVelWorld (particles.ts:201-236) */
static void pc_velocity(const orc_state *s, const parcel_scene *S, const double p[3],
                        double mature, double hold, double v[3]) {
  dims3 d = s->d[0];
  double below = S->ymax - p[1], u[3];
  pc_uvw(S, p, u);
  for (int a = 0; a < 3; ++a) u[a] = clampd(u[a], 0.002, 0.998);
  double gate = mixd(0.38, 1.0, sstep(0.002, 0.055, below));
  double mono = mixd(0.020, 0.046, sstep(0.04, 0.42, below)) * gate;
  double stack = mixd(0.032, 0.062, sstep(0.04, 0.48, below));
  double tail = hold * (1.0 - sstep(0.003, 0.034, fmax(0.0, below)));
  double settle = mixd(mono, stack, mature) * mixd(1.0, 0.035, tail);
  double sub[3] = {0.0, 0.0, 0.0};
  double deep = sstep(0.035, 0.24, below);
  if (deep > 0.001) {
    double tp = S->t * 0.42;
    double a1 = 8.4 * p[0] + 5.2 * p[1] + 6.6 * p[2] + tp;
    double a2 = -7.2 * p[0] + 6.0 * p[1] + 9.4 * p[2] - tp * 0.81;
    double a3 = 11.5 * p[0] - 7.5 * p[1] - 8.5 * p[2] + tp * 1.13;
    double k1[3] = {-8.16, 3.84, 7.36}, k2[3] = {-8.20, -1.38, -5.40}, k3[3] = {8.30, -1.10, 12.20};
    for (int a = 0; a < 3; ++a)
      sub[a] = k1[a] * cos(a1) + k2[a] * cos(a2) + k3[a] * (0.45 * sin(a3));
    sub[1] *= sstep(0.06, 0.24, below);
    for (int a = 0; a < 3; ++a) sub[a] *= 0.00028 * deep;
  }
  v[0] = sample_cell(s->cu, d, u[0], u[1], u[2]) + sub[0];
  v[1] = sample_cell(s->cv, d, u[0], u[1], u[2]) * mixd(1.0, 0.045, tail) - settle + sub[1];
  v[2] = sample_cell(s->cw, d, u[0], u[1], u[2]) + sub[2];
}

/* splatTrilinearIntact/Lysed (particles.ts:238-301): round-half-even of
 * ((mass 65536) w_x) w_y w_z added to the eight clamped cells of the splat
 * grid (CONC_GRID: the concentration's) */
static void pc_splat(orc_state *s, int which, const double u[3], double mass) {
  dims3 d = conc_dims(s);
  int n[3] = {d.nx, d.ny, d.nz}, b0[3], b1[3];
  double f[3];
  for (int a = 0; a < 3; ++a) {
    double c = u[a] * n[a] - 0.5, b = floor(c);
    f[a] = clampd(c - b, 0.0, 1.0);
    b0[a] = clampi((int)b, 0, n[a] - 1);
    b1[a] = clampi((int)b + 1, 0, n[a] - 1);
  }
  double sc = mass * 65536.0;
  for (int c = 0; c < 8; ++c) {
    int a = c & 1, b = (c >> 1) & 1, e = c >> 2;
    double w = sc * (a ? f[0] : 1.0 - f[0]) * (b ? f[1] : 1.0 - f[1]) * (e ? f[2] : 1.0 - f[2]);
    uint32_t q = (uint32_t)nearbyint(w);
    if (q > 0) s->splat[which][IDX(d, a ? b1[0] : b0[0], b ? b1[1] : b0[1], e ? b1[2] : b0[2])] += q;
  }
}

static void parcel_step(orc_state *s, int g, const parcel_scene *S) {
  double *P = s->parcels + 8L * g;
  double mass = P[3];
  if (mass <= 1e-6) return;
  const double dt = S->dt, hx = S->hx, ymin = S->ymin;
  double pos[3] = {P[0], P[1], P[2]}, st[3] = {P[4], P[5], P[6]};
  double raw = fmax(0.0, P[7]), ticks = floor(raw);
  double lysis = clampd(raw - ticks, 0.0, 0.98), age = ticks / 240.0;
  double mature = sstep(8.5, 18.0, age);
  float tr = (float)g * 0.7548776662f;
  double tail_rank = (double)(tr - floorf(tr));
  double hold = pow(sstep(0.36, 0.995, tail_rank), 1.65) * (1.0 - sstep(5.2, 10.2, age));
  double v1[3], v2[3], v3[3], va[3], q[3];
  pc_velocity(s, S, pos, mature, hold, v1);
  double sp2 = v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2];
  if (dt > 0.0) {
    ticks = fmin(8191.0, ticks + fmax(1.0, nearbyint(dt * 240.0)));
    for (int a = 0; a < 3; ++a) q[a] = pos[a] + 0.5 * v1[a] * dt;
    pc_velocity(s, S, q, mature, hold, v2);
    for (int a = 0; a < 3; ++a) va[a] = v2[a];
    if (sp2 > 0.0025) {
      for (int a = 0; a < 3; ++a) q[a] = pos[a] + 0.75 * v2[a] * dt;
      pc_velocity(s, S, q, mature, hold, v3);
      for (int a = 0; a < 3; ++a) va[a] = (2.0 / 9.0) * v1[a] + (3.0 / 9.0) * v2[a] + (4.0 / 9.0) * v3[a];
    }
    double slip = sstep(ymin + 0.005, ymin + 0.042, pos[1]);
    double rxz = hypot(pos[0], pos[2]), o[2] = {0.0, 0.0};
    if (rxz > 1e-4) {
      double k = 0.016 * (1.0 - slip) * sstep(0.01, 0.22, rxz);
      o[0] = pos[0] / fmax(rxz, 1e-4) * k;
      o[1] = pos[2] / fmax(rxz, 1e-4) * k;
    }
    va[0] = va[0] * mixd(0.45, 1.0, slip) + o[0];
    va[1] = va[1] * mixd(0.15, 1.0, slip);
    va[2] = va[2] * mixd(0.45, 1.0, slip) + o[1];
    for (int a = 0; a < 3; ++a) pos[a] += va[a] * dt;
    pos[1] = clampd(pos[1], ymin + 0.0055, S->ymax - 0.0018);
    if (sp2 > 2.5e-7 && slip > 0.15) {
      double sl = fmax(sqrt(st[0] * st[0] + st[1] * st[1] + st[2] * st[2]), 1e-5), vp[3], vm[3];
      double sd[3] = {st[0] / sl, st[1] / sl, st[2] / sl}, pr = 0.0065;
      for (int a = 0; a < 3; ++a) q[a] = pos[a] + sd[a] * pr;
      pc_velocity(s, S, q, mature, hold, vp);
      for (int a = 0; a < 3; ++a) q[a] = pos[a] - sd[a] * pr;
      pc_velocity(s, S, q, mature, hold, vm);
      for (int a = 0; a < 3; ++a) st[a] += (vp[a] - vm[a]) * (sl / (2.0 * pr)) * dt;
      double vmag = sqrt(sp2);
      if (vmag > 0.001 && sl < 0.006)
        for (int a = 0; a < 3; ++a) st[a] += (v1[a] / vmag) * (0.0022 * dt);
      double maxlen = mixd(0.0270, 0.0420, hold * (1.0 - sstep(0.005, 0.065, fmax(0.0, S->ymax - pos[1]))));
      double len = sqrt(st[0] * st[0] + st[1] * st[1] + st[2] * st[2]);
      double nl = clampd(len, 0.0026, maxlen);
      for (int a = 0; a < 3; ++a) st[a] = st[a] / fmax(len, 1e-5) * nl;
    } else if (slip <= 0.15) {
      st[1] *= exp(-dt * 5.0);
    }
  }
  double u[3], uc[3];
  pc_uvw(S, pos, u);
  double rh = hypot(pos[0], pos[2]);
  mass *= mixd(1.0, 1.0 - sstep(hx * 0.68, hx * 0.94, rh), clampd(dt * 6.0, 0.0, 1.0));
  if (mass <= 1e-5 || rh >= hx * 0.95 || u[0] <= 0.002 || u[0] >= 0.998 || u[2] <= 0.002 ||
      u[2] >= 0.998 || u[1] >= 0.998) {
    P[3] = 0.0; /* retired where it stood at the start of the step */
    return;
  }
  for (int a = 0; a < 3; ++a) uc[a] = clampd(u[a], 0.002, 0.998);
  if (dt > 0.0 && pos[1] < ymin + 0.026) {
    double w = 1.0 - sstep(ymin + 0.005, ymin + 0.026, pos[1]);
    double dm = mass * clampd(w * dt * 1.65, 0.0, 0.095);
    if (dm > 1e-7) {
      mass = fmax(0.0, mass - dm);
      if (dm * 0.68 > 1e-7) pc_splat(s, 1, uc, dm * 0.68);
    }
    if (mass <= 1e-5) {
      P[3] = 0.0;
      return;
    }
  }
  float crf = (float)g * 0.6180339887f + 0.17f;
  double cell_res = (double)(crf - floorf(crf));
  /* cSample.r + cSample.g (particles.ts:403-404): the dye alone without the
   * 4-channel concentration */
  double cn = sample_cell(intact_of(s), conc_dims(s), uc[0], uc[1], uc[2]);
  if (lysed_of(s)) cn += sample_cell(lysed_of(s), conc_dims(s), uc[0], uc[1], uc[2]);
  double dilute = 1.0 - sstep(0.012, 0.22, cn);
  double disp = sstep(1.1, 7.8, age + dilute * 2.2 + lysis * 2.5);
  double late = sstep(7.2, 16.0, age) * dilute;
  double min_intact = 0.13400 * mixd(mixd(0.46, 0.64, cell_res),
                                     mixd(mixd(0.07, 0.20, cell_res), mixd(0.014, 0.052, cell_res), late), disp);
  double headroom = mass - min_intact;
  double shelter = sstep(ymin + 0.024, ymin + 0.062, pos[1]);
  if (dt > 0.0 && headroom > 1e-7 && shelter > 0.02) {
    double shear = mixd(0.22, 1.0, sstep(0.0005, 0.0120, sp2));
    double water = (1.0 - 0.72 * sstep(0.016, 0.30, cn)) * shelter * shear;
    lysis = clampd(lysis + S->kl * (1.85 * water) * dt, 0.0, 0.98);
    double frag = sstep(0.04, 0.48, lysis) * water;
    double dm = fmin(headroom, mass * clampd(frag * S->kl * 1.95 * dt, 0.0, 0.16));
    if (dm > 1e-7) {
      mass = fmax(0.0, mass - dm);
      pc_splat(s, 1, uc, dm);
    }
  }
  P[0] = pos[0]; P[1] = pos[1]; P[2] = pos[2]; P[3] = mass;
  P[4] = st[0]; P[5] = st[1]; P[6] = st[2]; P[7] = ticks + clampd(lysis, 0.0, 0.98);
  double sl = sqrt(st[0] * st[0] + st[1] * st[1] + st[2] * st[2]);
  if (sl < 0.0038) {
    if (mass > 1e-6) pc_splat(s, 0, uc, mass);
    return;
  }
  double sxz = hypot(st[0], st[2]), ht[3] = {1.0, 0.0, 0.0};
  if (sxz > 1e-4) {
    ht[0] = -st[2] / fmax(sxz, 1e-5);
    ht[2] = st[0] / fmax(sxz, 1e-5);
  }
  double sc[3] = {st[1] * ht[2] - st[2] * ht[1], st[2] * ht[0] - st[0] * ht[2], st[0] * ht[1] - st[1] * ht[0]};
  double scl = sqrt(sc[0] * sc[0] + sc[1] * sc[1] + sc[2] * sc[2]), cc[3];
  for (int a = 0; a < 3; ++a) cc[a] = scl > 1e-4 ? sc[a] / fmax(scl, 1e-5) : ht[a];
  double hm = sstep(0.35, 0.85, fabs(st[1]) / fmax(sl, 1e-5)), nn[3];
  for (int a = 0; a < 3; ++a) nn[a] = mixd(cc[a], ht[a], hm);
  double nl = sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
  double span = clampd(sl * 0.62, 0.0032, 0.0085);
  double al[4] = {-0.42, -0.14, 0.14, 0.42}, bl[4] = {0.0, 1.0, -1.0, 0.0}, wt[4] = {0.21, 0.29, 0.29, 0.21};
  for (int k = 0; k < 4; ++k) {
    double pq[3], uq[3];
    for (int a = 0; a < 3; ++a) pq[a] = pos[a] + al[k] * st[a] + bl[k] * nn[a] / nl * span;
    pc_uvw(S, pq, uq);
    for (int a = 0; a < 3; ++a) uq[a] = clampd(uq[a], 0.002, 0.998);
    if (wt[k] * mass > 1e-6) pc_splat(s, 0, uq, wt[k] * mass);
  }
}

/* ---- PARTICLE_EMIT (particles.ts:53-186): emission `g` of the pending
 * count writes a new parcel at ring slot (cursor + g) mod np, in the drop slot
 * its share of the summed drain rates picks, after moving a live parcel's
 * mass at that slot into the lysed counter.  The R3 sequence, hash33 and the
 * fract() of the slot number are formed in single precision without fused
 * multiply-adds, as the reference's f32 does (for slot numbers ~4e5 an FP64
 * fract differs by up to 0.03); the rest is FP64. */
static float fr32(float x) { return x - floorf(x); }
static void hash33_f(float px, float py, float pz, float out[3]) {
  float a = fr32(px * 0.1031f), b = fr32(py * 0.1030f), c = fr32(pz * 0.0973f);
  /* p3 += dot(p3, p3.yxz + 33.33) */
  float dt = a * (b + 33.33f) + b * (a + 33.33f) + c * (c + 33.33f);
  a += dt; b += dt; c += dt;
  /* fract((p3.xxy + p3.yxx) * p3.zyx) */
  out[0] = fr32((a + b) * c);
  out[1] = fr32((a + a) * b);
  out[2] = fr32((b + a) * a);
}
static void parcel_emit(orc_state *s, int g, int count) {
  const double hx = s->wp.half_xz, ytop = s->wp.surface_y, y0 = ytop - s->wp.depth;
  dims3 d = conc_dims(s);
  double total = 0.0;
  for (int q = 0; q < MAX_DROPS; ++q)
    if (s->drops[q].seed > 0.0 && s->drops[q].drain > 0.0001) total += s->drops[q].drain;
  if (total <= 1e-6) return;
  double target = (g + 0.5) / (double)(count > 1 ? count : 1) * total, acc = 0.0;
  int pick = 0;
  for (int q = 0; q < MAX_DROPS; ++q)
    if (s->drops[q].seed > 0.0 && s->drops[q].drain > 0.0001) {
      acc += s->drops[q].drain;
      pick = q;
      if (acc >= target) break;
    }
  const orc_drop *dr = &s->drops[pick];
  const double seed = fmax(dr->seed, 0.5);
  const int ring = (int)(((long)s->emit_cursor + g) % s->np);
  double *P = s->parcels + 8L * ring;
  if (P[3] > 1e-6) { /* the overwritten parcel's mass goes to the lysed counter */
    double px = P[0], py = P[1], pz = P[2];
    if (py < ytop - 0.32) py = y0 + 0.012;
    double u[3] = {(px + hx) / (2.0 * hx), (py - y0) / (ytop - y0), (pz + hx) / (2.0 * hx)};
    if (u[0] > 0.002 && u[1] > 0.002 && u[2] > 0.002 && u[0] < 0.998 && u[1] < 0.998 &&
        u[2] < 0.998) {
      int ci = clampi((int)floor(u[0] * d.nx), 0, d.nx - 1);
      int cj = clampi((int)floor(u[1] * d.ny), 0, d.ny - 1);
      int ck = clampi((int)floor(u[2] * d.nz), 0, d.nz - 1);
      uint32_t q = (uint32_t)nearbyint(P[3] * 65536.0);
      if (q > 0) s->splat[1][IDX(d, ci, cj, ck)] += q;
    }
  }
  const float seedf = (float)seed;
  const float seq = (float)ring + seedf * 137.0f;
  float h[3];
  hash33_f((float)g, (float)s->wp.sim_time * 60.0f, seedf, h);
  const float r3x = fr32(seq * 0.8191725134f + h[0] * 0.032f);
  const float r3y = fr32(seq * 0.6710436067f + h[1] * 0.032f);
  const float r3z = fr32(seq * 0.5497004779f + h[2] * 0.032f);
  const double th = 6.28318530718 * r3x;
  const double az = 1.0 + 0.036 * cos(4.0 * th + seed * 1.73) + 0.026 * cos(6.0 * th - seed * 2.31) +
                    0.018 * sin(8.0 * th + seed * 0.89);
  const double rdrop = fmax(dr->radius, 0.010);
  const double emerg = (dr->y - ytop) / fmax(rdrop, 0.018);
  const double nasc = sstep(-0.32, 0.18, emerg);
  const double slug = sstep(0.26, 0.85, dr->drain);
  const double wake = mixd(0.46, 1.0, sstep(0.018, 0.26, dr->drain));
  const double rneck = mixd(0.013 * wake, clampd(rdrop * 0.84 + 0.014, 0.020, 0.056), slug) * az;
  const double inner = 0.37 * slug;
  const int cap = slug > 0.08 && r3z < 0.16f;
  const double rann = mixd(inner, 1.0, pow((double)r3y, mixd(0.50, 0.44, slug)));
  const double rcap = mixd(0.14 * slug, fmax(inner * 1.32, 0.08), sqrt((double)r3y));
  const double rfrac = mixd(cap ? rcap : rann, sqrt((double)r3y), nasc * 0.82);
  const double r = rneck * rfrac;
  const double wa = s->wp.sim_time * 1.85 + seed * 3.17;
  const double wd = 0.0038 * (1.0 - slug) * (1.0 - wake * 0.45);
  const double px = dr->x + cos(wa) * wd + cos(th) * r;
  const double pz = dr->z + sin(wa * 0.83 + seed) * wd + sin(th) * r;
  const double E = s->wp.half_extent;
  const double eta = clampd(wv_sample(s->wave, s->wp.n, clampd((px + E) / (2.0 * E), 0.002, 0.998),
                                      clampd((pz + E) / (2.0 * E), 0.002, 0.998), 0),
                            -0.085, 0.015);
  const double neck_top = ytop + fmin(0.0, eta) - 0.0038;
  const double sag = (1.0 - rfrac * rfrac) * (0.0085 * slug + 0.0045 * nasc);
  const double zs = cap ? r3z * 3.6 : (r3z - 0.16 * slug) / fmax(1.0 - 0.16 * slug, 0.5);
  const double py = neck_top - sag - mixd(0.0014, 0.0135, clampd(zs, 0.0, 1.0));
  const double frag = 1.0 - sstep(0.022, 0.056, rdrop);
  const double rim = sstep(0.58, 1.0, rfrac);
  const double lys = clampd(0.14 * rim + 0.28 * frag * (0.45 + 0.55 * rfrac), 0.0, 0.44);
  double t[3] = {-sin(th) * fmax(rfrac, 0.25), -0.80, cos(th) * fmax(rfrac, 0.25)};
  const double tl = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
  P[0] = px; P[1] = py; P[2] = pz; P[3] = 0.13400;
  for (int a = 0; a < 3; ++a) P[4 + a] = t[a] / tl * 0.0036;
  P[7] = lys;
}

/* the parcel stage: both counters cleared (BloodDropEngine.ts:2073-2074),
 * the pending emission (:2082-2087), then every parcel stepped and splatted
 * (:2089-2093) */
static void parcels_stage(orc_state *s, double dt) {
  if (!s->wave_on || !s->parcels) return;
  parcel_scene S = {s->wp.half_xz, s->wp.surface_y - s->wp.depth, s->wp.surface_y, s->wp.sim_time,
                    s->wp.hemolysis, clampd(dt, 0.0, 0.025)};
  memset(s->splat[0], 0, sizeof(uint32_t) * splat_cells(s));
  memset(s->splat[1], 0, sizeof(uint32_t) * splat_cells(s));
  if (s->emit_count > 0 && s->np > 0)
    for (int g = 0; g < s->emit_count; ++g) parcel_emit(s, g, s->emit_count);
  s->emit_count = 0;
  for (int g = 0; g < s->np; ++g) parcel_step(s, g, &S);
}

/* ---- CONC_PREDICT (advect.ts:28-72) of the lysed channel (.g) on the
 * concentration grid: uvw = (cell + 1/2) / N; the lateral walls and the top
 * row 0 (:42-45); v0 the centre velocity (the MAC grid's) sampled at uvw;
 * below speed^2 2.5e-5 the cell's own value (:54-58); below 0.012 a one-
 * stage trace uvw - v0 dt / domainSize (:60-62); else the RK2 trace through
 * the midpoint velocity (:63-67); uvw clamped to [0.001, 0.999]; dt clamped
 * to [0, 0.025]; domainSize the world box (common.ts:52-54). */
static void conc_predict(orc_state *s, double dt) {
  dims3 d = s->d[0], dc = s->dc;
  const double dom[3] = {2.0 * s->wp.half_xz, s->wp.depth, 2.0 * s->wp.half_xz};
  const double *g = s->conc + s->nc;
  dt = clampd(dt, 0.0, 0.025);
#pragma omp parallel for
  for (int k = 0; k < dc.nz; ++k)
    for (int j = 0; j < dc.ny; ++j)
      for (int i = 0; i < dc.nx; ++i) {
        long c = IDX(dc, i, j, k);
        if (conc_wall(dc, i, j, k)) { s->cpred[c] = 0.0; continue; }
        double uvw[3] = {(i + 0.5) / dc.nx, (j + 0.5) / dc.ny, (k + 0.5) / dc.nz};
        double v0[3] = {sample_cell(s->cu, d, uvw[0], uvw[1], uvw[2]),
                        sample_cell(s->cv, d, uvw[0], uvw[1], uvw[2]),
                        sample_cell(s->cw, d, uvw[0], uvw[1], uvw[2])};
        double sp2 = v0[0] * v0[0] + v0[1] * v0[1] + v0[2] * v0[2], back[3];
        if (sp2 < 2.5e-5) { s->cpred[c] = g[c]; continue; }
        if (sp2 < 0.012) {
          for (int a = 0; a < 3; ++a) back[a] = clampd(uvw[a] - (v0[a] * dt) / dom[a], 0.001, 0.999);
        } else {
          double mid[3];
          for (int a = 0; a < 3; ++a) mid[a] = clampd(uvw[a] - 0.5 * (v0[a] * dt) / dom[a], 0.001, 0.999);
          double vm[3] = {sample_cell(s->cu, d, mid[0], mid[1], mid[2]),
                          sample_cell(s->cv, d, mid[0], mid[1], mid[2]),
                          sample_cell(s->cw, d, mid[0], mid[1], mid[2])};
          for (int a = 0; a < 3; ++a) back[a] = clampd(uvw[a] - (vm[a] * dt) / dom[a], 0.001, 0.999);
        }
        s->cpred[c] = sample_cell(g, dc, back[0], back[1], back[2]);
      }
}

/* ---- CONC_RESOLVE (advect.ts:81-298): the 4-channel composite.  Per cell
 * (lateral walls and the top row 0):
 *  - nothing to do where neither counter holds mass, the predictor's two
 *    octant samples are below 1e-7 and (the 8 floor rows) the bed below
 *    1e-6 (advect.ts:151-164), or (above the floor rows) the predicted lysed
 *    value and its 6-neighbour mean are below 1.2e-6 (:176-179);
 *  - lysed: MacCormack of the predictor (forward estimate at uvw + v0 dt) with
 *    the 8-texel limiter of the source, a 6-neighbour diffusion blend
 *    clamp(5.6 dt, 0, 0.105), an ambient dilution exp(-dt mix(0.016, 0.034,
 *    height)), the lysed counter's mass (72% of it captured by the bed near
 *    the floor), cut below 2.5e-6 (:195-237);
 *  - bed (floor rows): self-drain, settling from above, radial spreading and
 *    vertical diffusion, decay with radius, plus the captured mass, a
 *    row-height fade, cut below 1e-6 (:241-272);
 *  - intact: the intact counter blended with its 6-neighbour mean
 *    (18% / 36% where it holds mass) plus the bed, under the radial wall
 *    sponge; the bed also capped by a row profile (:276-285);
 *  - oxygenation/shear state from the flow speed, the trace's convective
 *    shear, the sheet thickness and the bed fraction (:290-297). */
static double conc_resolve_cell(const orc_state *s, const double *src, double dt, int ix, int iy,
                                int iz, double out[4]) {
  /* d: the concentration grid (its cells, counters and predictor); the
   * centre velocity is sampled on the MAC grid dv */
  dims3 d = s->dc, dv = s->d[0];
  const long n = s->nc;
  const double *og = src + n, *oa = src + 3 * n, *pg = s->cpred;
  const uint32_t *si = s->splat[0], *sl = s->splat[1];
  out[0] = out[1] = out[2] = out[3] = 0.0;
  if (ix == 0 || ix >= d.nx - 1 || iz == 0 || iz >= d.nz - 1 || iy >= d.ny - 1) return 0.0;
  const double N[3] = {(double)d.nx, (double)d.ny, (double)d.nz};
  const long c = IDX(d, ix, iy, iz);
  const uint32_t ri = si[c], rl = sl[c];
  const int floor_zone = iy <= 7;
  if (ri == 0 && rl == 0) {
    double o0 = sample_cell(pg, d, ix / N[0], iy / N[1], iz / N[2]);
    double o1 = sample_cell(pg, d, (ix + 1.0) / N[0], (iy + 1.0) / N[1], (iz + 1.0) / N[2]);
    double bed_on = floor_zone ? oa[c] : 0.0;
    if (o0 + o1 <= 1.0e-7 && bed_on <= 1.0e-6) return 0.0;
  }
#define PG(i, j, k) pg[IDX(d, clampi(i, 0, d.nx - 1), clampi(j, 0, d.ny - 1), clampi(k, 0, d.nz - 1))]
#define OA(i, j, k) oa[IDX(d, clampi(i, 0, d.nx - 1), clampi(j, 0, d.ny - 1), clampi(k, 0, d.nz - 1))]
#define SI(i, j, k) (si[IDX(d, clampi(i, 0, d.nx - 1), clampi(j, 0, d.ny - 1), clampi(k, 0, d.nz - 1))] / 65536.0)
  const double gpred = pg[c];
  const double lapg = (PG(ix + 1, iy, iz) + PG(ix - 1, iy, iz) + PG(ix, iy + 1, iz) +
                       PG(ix, iy - 1, iz) + PG(ix, iy, iz + 1) + PG(ix, iy, iz - 1)) * (1.0 / 6.0);
  if (!floor_zone && ri == 0 && rl == 0 && gpred <= 1.2e-6 && lapg <= 1.2e-6) return 0.0;
  double lapi = 0.0;
  if (ri > 0 || gpred > 8.0e-5 || lapg > 8.0e-5)
    lapi = (SI(ix + 1, iy, iz) + SI(ix - 1, iy, iz) + SI(ix, iy + 1, iz) + SI(ix, iy - 1, iz) +
            SI(ix, iy, iz + 1) + SI(ix, iy, iz - 1)) * (1.0 / 6.0);
  const double hx = s->wp.half_xz, ytop = s->wp.surface_y, y0 = ytop - s->wp.depth;
  dt = clampd(dt, 0.0, 0.025);
  const double uvw[3] = {(ix + 0.5) / N[0], (iy + 0.5) / N[1], (iz + 0.5) / N[2]};
  const double dom[3] = {2.0 * hx, ytop - y0, 2.0 * hx};
  const double p[3] = {-hx + uvw[0] * dom[0], y0 + uvw[1] * dom[1], -hx + uvw[2] * dom[2]};
  double v0[3] = {sample_cell(s->cu, dv, uvw[0], uvw[1], uvw[2]), sample_cell(s->cv, dv, uvw[0], uvw[1], uvw[2]),
                  sample_cell(s->cw, dv, uvw[0], uvw[1], uvw[2])};
  double mid[3], back[3], fwd[3];
  for (int a = 0; a < 3; ++a) mid[a] = clampd(uvw[a] - 0.5 * (v0[a] * dt) / dom[a], 0.001, 0.999);
  double vm[3] = {sample_cell(s->cu, dv, mid[0], mid[1], mid[2]), sample_cell(s->cv, dv, mid[0], mid[1], mid[2]),
                  sample_cell(s->cw, dv, mid[0], mid[1], mid[2])};
  for (int a = 0; a < 3; ++a) {
    back[a] = clampd(uvw[a] - (vm[a] * dt) / dom[a], 0.001, 0.999);
    fwd[a] = clampd(uvw[a] + (v0[a] * dt) / dom[a], 0.001, 0.999);
  }
  const double est = sample_cell(pg, d, fwd[0], fwd[1], fwd[2]);
  const double corr = gpred + s->p.maccormack * (og[c] - est);
  int bi = (int)floor(back[0] * N[0] - 0.5), bj = (int)floor(back[1] * N[1] - 0.5),
      bk = (int)floor(back[2] * N[2] - 0.5);
  double lo = 1e300, hi = -1e300;
  for (int dz = 0; dz < 2; ++dz)
    for (int dy = 0; dy < 2; ++dy)
      for (int dx = 0; dx < 2; ++dx) {
        double v = og[IDX(d, clampi(bi + dx, 0, d.nx - 1), clampi(bj + dy, 0, d.ny - 1),
                          clampi(bk + dz, 0, d.nz - 1))];
        lo = fmin(lo, v);
        hi = fmax(hi, v);
      }
  double cl = clampd(corr, lo, hi);
  cl = mixd(cl, lapg, clampd(dt * 5.60, 0.0, 0.105));
  cl *= exp(-dt * mixd(0.016, 0.034, sstep(y0 + 0.045, y0 + 0.26, p[1])));
  const double dm = rl / 65536.0;
  const double capture = 1.0 - sstep(y0 + 0.018, y0 + 0.058, p[1]);
  cl += dm * (1.0 - capture * 0.72);
  if (cl < 2.5e-6) cl = 0.0;
  const double rh = sqrt(p[0] * p[0] + p[2] * p[2]);
  double bed = 0.0;
  if (floor_zone) {
    const double bc = oa[c];
    const double bh = 0.25 * (OA(ix + 1, iy, iz) + OA(ix - 1, iy, iz) + OA(ix, iy, iz + 1) + OA(ix, iy, iz - 1));
    const double ba = OA(ix, iy + 1, iz), bb = OA(ix, iy > 0 ? iy - 1 : 0, iz);
    const double spread = clampd(dt * (3.20 + 6.50 * fmin(bc, 0.65)), 0.0, 0.165);
    const double vert = clampd(dt * 0.75, 0.0, 0.026);
    const double settle = clampd(dt * 2.80, 0.0, 0.090);
    const double decay = exp(-dt * mixd(0.095, 0.28, sstep(hx * 0.35, hx * 0.78, rh)));
    const double drained = (iy == 0 ? bc : bc * (1.0 - settle)) * decay;
    const double above = iy >= 7 ? 0.0 : ba * settle;
    const double sm = mixd(mixd(drained + above, bh, spread), 0.5 * (ba + bb), vert);
    const double fade = 1.0 - 0.38 * sstep(3.5, 7.0, (double)iy);
    bed = (sm + dm * capture * 0.68) * fade;
    if (bed < 1.0e-6) bed = 0.0;
  }
  const double wall = 1.0 - sstep(hx * 0.68, hx * 0.94, rh);
  const double taper = 1.0 - sstep(hx * 0.52, hx * 0.88, rh);
  bed = fmin(bed * wall, 0.54 * exp(-0.56 * iy) * taper);
  const double craw = ri / 65536.0;
  const double intact = mixd(craw, lapi, ri > 0 ? 0.36 : 0.18) * wall + bed;
  cl *= wall;
  const double tot = intact + cl;
  double ox = 0.0;
  if (tot > 1.0e-5) {
    const double fl = clampd(cl / fmax(tot, 1.0e-5), 0.0, 1.0), fi = 1.0 - fl;
    const double fb = sstep(0.0, 0.85, bed / fmax(tot, 1.0e-5));
    const double speed = sqrt(v0[0] * v0[0] + v0[1] * v0[1] + v0[2] * v0[2]);
    const double dvx = v0[0] - vm[0], dvy = v0[1] - vm[1], dvz = v0[2] - vm[2];
    const double conv = sqrt(dvx * dvx + dvy * dvy + dvz * dvz) / fmax(0.5 * dt, 1.0e-3);
    const double sh = sstep(0.014, 0.135, speed + 0.38 * conv);
    const double thin = (1.0 - sstep(0.024, 0.24, intact)) * sstep(0.0006, 0.0065, intact);
    const double act = thin * (0.38 + 0.62 * sh) * fi;
    const double raw = clampd(0.08 + 0.62 * act + 0.36 * sh * fi, 0.0, 1.0);
    ox = raw * (1.0 - 0.82 * fb);
  }
#undef PG
#undef OA
#undef SI
  out[0] = intact; out[1] = cl; out[2] = ox; out[3] = bed;
  return 0.0;
}
static void conc_resolve(orc_state *s, double dt) {
  dims3 d = s->dc;
  const long n = s->nc;
#pragma omp parallel for
  for (int k = 0; k < d.nz; ++k)
    for (int j = 0; j < d.ny; ++j)
      for (int i = 0; i < d.nx; ++i) {
        double o[4];
        conc_resolve_cell(s, s->conc, dt, i, j, k, o);
        long c = IDX(d, i, j, k);
        for (int a = 0; a < 4; ++a) s->conc2[a * n + c] = o[a];
      }
  double *t = s->conc; s->conc = s->conc2; s->conc2 = t;
}

/* the two splat counters hold at least the splat grid's cells (zeroed when
 * they grow) */
static void ensure_splat(orc_state *s) {
  long need = splat_cells(s);
  if (s->splat_cap >= need) return;
  free(s->splat[0]);
  free(s->splat[1]);
  s->splat[0] = (uint32_t *)calloc(need, sizeof(uint32_t));
  s->splat[1] = (uint32_t *)calloc(need, sizeof(uint32_t));
  s->splat_cap = need;
}

/* on: 0 off; 1 the concentration on the MAC grid; 2 on a grid twice as fine
 * per axis (the reference's CONC_GRID over its VEL_GRID).  A change of grid
 * starts from zero. */
void orc_conc_enable(orc_state *s, int on) {
  if (on != 1 && on != 2) {
    s->conc_on = 0;
    return;
  }
  if (s->conc && s->conc_r != on) {
    free(s->conc); free(s->conc2); free(s->cpred);
    s->conc = s->conc2 = s->cpred = NULL;
  }
  s->conc_r = on;
  s->dc.nx = on * s->d[0].nx;
  s->dc.ny = on * s->d[0].ny;
  s->dc.nz = on * s->d[0].nz;
  s->nc = (long)s->dc.nx * s->dc.ny * s->dc.nz;
  if (!s->conc) {
    s->conc = (double *)calloc(4 * (size_t)s->nc, sizeof(double));
    s->conc2 = (double *)calloc(4 * (size_t)s->nc, sizeof(double));
    s->cpred = (double *)calloc((size_t)s->nc, sizeof(double));
  }
  s->conc_on = 1;
  ensure_splat(s);
}

void orc_parcels_emit(orc_state *s, int cursor, int count) {
  s->emit_cursor = cursor;
  s->emit_count = count;
}

void orc_parcels_set(orc_state *s, const double *p, int n) {
  if (n > s->np_cap) {
    free(s->parcels);
    s->parcels = (double *)malloc(sizeof(double) * 8 * (size_t)n);
    s->np_cap = n;
  }
  ensure_splat(s);
  memcpy(s->parcels, p, sizeof(double) * 8 * (size_t)n);
  s->np = n;
}

double *orc_parcels(orc_state *s) { return s->parcels; }
uint32_t *orc_splat(orc_state *s, int which) {
  ensure_splat(s);
  return (which == 0 || which == 1) ? s->splat[which] : NULL;
}
/* the splat grid (the concentration grid while it is on, else the MAC
 * grid): out = nx, ny, nz */
void orc_splat_dims(orc_state *s, int out[3]) {
  dims3 d = conc_dims(s);
  out[0] = d.nx;
  out[1] = d.ny;
  out[2] = d.nz;
}

void orc_wave_enable(orc_state *s, const orc_wave_params *p) {
  if (!p) {
    s->wave_on = 0;
    return;
  }
  if (!s->wave || s->wp.n != p->n) {
    free(s->wave); free(s->wz); free(s->wz2); free(s->wfilm);
    long nn = (long)p->n * p->n;
    s->wave = (double *)calloc(4 * nn, sizeof(double));
    s->wz = (double *)calloc(2 * nn, sizeof(double));
    s->wz2 = (double *)calloc(2 * nn, sizeof(double));
    s->wfilm = (double *)calloc(nn, sizeof(double));
  }
  if (!s->ptop) s->ptop = (double *)calloc((long)s->d[0].nx * s->d[0].nz, sizeof(double));
  if (p->scene && !s->curl) s->curl = (double *)calloc(4 * s->n[0], sizeof(double));
  s->wp = *p;
  if (!(s->wp.half_xz > 0.0)) s->wp.half_xz = 0.5 * s->p.nx * s->p.dx;
  if (!(s->wp.depth > 0.0)) s->wp.depth = s->p.ny * s->p.dx;
  memset(s->wave, 0, sizeof(double) * 4 * (long)p->n * p->n);
  memset(s->ptop, 0, sizeof(double) * (long)s->d[0].nx * s->d[0].nz);
  memset(s->drops, 0, sizeof(s->drops));
  s->wave_on = 1;
}

void orc_set_drops(orc_state *s, const orc_drop *dr, int n) {
  memset(s->drops, 0, sizeof(s->drops));
  for (int q = 0; q < n && q < MAX_DROPS; ++q) s->drops[q] = dr[q];
}

/* ---- state ---- */
orc_state *orc_create(const orc_params *p) {
  if (p->levels < 1 || p->levels > MAXLEV) return NULL;
  orc_state *s = (orc_state *)calloc(1, sizeof(orc_state));
  s->p = *p;
  for (int l = 0; l <= p->levels; ++l) {
    s->d[l].nx = p->nx >> l;
    s->d[l].ny = p->ny >> l;
    s->d[l].nz = p->nz >> l;
    s->n[l] = (long)s->d[l].nx * s->d[l].ny * s->d[l].nz;
  }
  long n = s->n[0];
  double **f0[] = {&s->u, &s->v, &s->w, &s->pr, &s->dye, &s->cu, &s->cv, &s->cw,
                   &s->rhs, &s->dpred, &s->u2, &s->v2, &s->w2};
  for (unsigned a = 0; a < sizeof(f0) / sizeof(f0[0]); ++a) *f0[a] = (double *)calloc(n, sizeof(double));
  s->mgx[0] = s->pr;
  s->mgb[0] = s->rhs;
  s->mgr[0] = (double *)calloc(n, sizeof(double));
  for (int l = 1; l <= p->levels; ++l) {
    s->mgx[l] = (double *)calloc(s->n[l], sizeof(double));
    s->mgb[l] = (double *)calloc(s->n[l], sizeof(double));
    s->mgr[l] = (double *)calloc(s->n[l], sizeof(double));
  }
  for (int a = 0; a < 4; ++a) s->cg[a] = (double *)calloc(s->n[p->levels], sizeof(double));
  return s;
}

void orc_destroy(orc_state *s) {
  if (!s) return;
  double *f0[] = {s->u, s->v, s->w, s->pr, s->dye, s->cu, s->cv, s->cw, s->rhs, s->dpred,
                  s->u2, s->v2, s->w2, s->mgr[0]};
  for (unsigned a = 0; a < sizeof(f0) / sizeof(f0[0]); ++a) free(f0[a]);
  for (int l = 1; l <= s->p.levels; ++l) { free(s->mgx[l]); free(s->mgb[l]); free(s->mgr[l]); }
  for (int a = 0; a < 4; ++a) free(s->cg[a]);
  free(s->wave); free(s->wz); free(s->wz2); free(s->wfilm); free(s->ptop); free(s->curl);
  free(s->parcels); free(s->splat[0]); free(s->splat[1]); free(s->conc); free(s->conc2);
  free(s->cpred);
  free(s);
}

double *orc_field(orc_state *s, int field, int level) {
  if (level < 0 || level > s->p.levels) return NULL;
  switch (field) {
    case F_U: return s->u;
    case F_V: return s->v;
    case F_W: return s->w;
    case F_P: return s->pr;
    case F_DYE: return s->dye;
    case F_CU: return s->cu;
    case F_CV: return s->cv;
    case F_CW: return s->cw;
    case F_RHS: return s->rhs;
    case F_MGX: return s->mgx[level];
    case F_MGB: return s->mgb[level];
    case F_MGR: return s->mgr[level];
    case F_DYE_PRED: return s->dpred;
    case F_WAVE: return s->wave_on && level == 0 ? s->wave : NULL;
    case F_PTOP: return s->wave_on && level == 0 ? s->ptop : NULL;
    case F_CURL: return level == 0 ? s->curl : NULL;
    case F_CONC: return level == 0 && s->conc_on ? s->conc : NULL;
    case F_CONC_PRED: return level == 0 && s->conc_on ? s->cpred : NULL;
  }
  return NULL;
}

long orc_field_len(orc_state *s, int field, int level) {
  if (field == F_MGX || field == F_MGB || field == F_MGR) return s->n[level];
  if (field == F_WAVE) return 4L * s->wp.n * s->wp.n;
  if (field == F_PTOP) return (long)s->d[0].nx * s->d[0].nz;
  if (field == F_CURL) return 4L * s->n[0];
  if (field == F_CONC) return s->conc_on ? 4L * s->nc : 0;
  if (field == F_CONC_PRED) return s->conc_on ? s->nc : 0;
  return s->n[0];
}

/* Seeded blob initial condition (no reference counterpart: the workload
 * replaces FLUID_INIT_VEL, fluid.ts:80-121).  dye = sum over blobs of
 * 1 - smoothstep(r - 1/2, r + 1/2, |x - c|) at cell centres; V on y-faces
 * (x = i+1/2, y = j+1, z = k+1/2) = -sum w0 * (same profile at the face);
 * U = W = P = 0. */
void orc_init_blobs(orc_state *s, const orc_blob *b, int nb) {
  dims3 d = s->d[0];
  long n = s->n[0];
  memset(s->u, 0, sizeof(double) * n);
  memset(s->w, 0, sizeof(double) * n);
  memset(s->pr, 0, sizeof(double) * n);
#pragma omp parallel for
  for (int k = 0; k < d.nz; ++k)
    for (int j = 0; j < d.ny; ++j)
      for (int i = 0; i < d.nx; ++i) {
        double dy = 0.0, vv = 0.0;
        for (int q = 0; q < nb; ++q) {
          double ex = i + 0.5 - b[q].cx, ey = j + 0.5 - b[q].cy, ez = k + 0.5 - b[q].cz;
          double r = sqrt(ex * ex + ey * ey + ez * ez);
          dy += 1.0 - orc_smoothstep(b[q].radius - 0.5, b[q].radius + 0.5, r);
          double fy = j + 1.0 - b[q].cy;
          double rf = sqrt(ex * ex + fy * fy + ez * ez);
          vv -= b[q].w0 * (1.0 - orc_smoothstep(b[q].radius - 0.5, b[q].radius + 0.5, rf));
        }
        long c = IDX(d, i, j, k);
        s->dye[c] = dy;
        s->v[c] = vv;
      }
}

void orc_stage(orc_state *s, int stage, double dt) {
  switch (stage) {
    case ST_F2C: face_to_centre(s); break;
    case ST_ADVECT: advect_mac(s, dt); break;
    case ST_CASSON: casson_stage(s, dt); break;
    case ST_DIV: divergence(s, s->u, s->v, s->w, s->rhs); break;
    case ST_VCYCLE: vcycle(s); break;
    case ST_PROJECT: project(s); break;
    case ST_DYE_PRED: dye_predict(s, dt); break;
    case ST_DYE_RES: dye_resolve(s, dt); break;
    case ST_WAVE: if (s->wave_on) wave_step(s, dt); break;
    case ST_TOP_PRESSURE: if (s->wave_on) top_pressure(s, dt); break;
    case ST_PARCELS: parcels_stage(s, dt); break;
    case ST_CONC_PRED: if (s->conc_on) conc_predict(s, dt); break;
    case ST_CONC_RES: if (s->conc_on) conc_resolve(s, dt); break;
  }
}

void orc_mg_op(orc_state *s, int op, int level, int count) {
  int L = s->p.levels;
  if (level < 0 || level > L) return;
  dims3 d = s->d[level];
  switch (op) {
    case MG_SMOOTH: smooth(s->mgx[level], s->mgb[level], d, level, count, smoother_omega(s)); break;
    case MG_RESIDUAL: residual(s->mgx[level], s->mgb[level], s->mgr[level], d, level); break;
    case MG_RESTRICT:
      if (level < L) restrict_op(s->mgr[level], d, s->mgb[level + 1], s->d[level + 1]);
      break;
    case MG_PROLONG:
      if (level < L) prolong_add(s->mgx[level], d, s->mgx[level + 1], s->d[level + 1]);
      break;
    case MG_COARSE:
      if (level == L) coarse_solve(s, L);
      break;
    case MG_ZERO: memset(s->mgx[level], 0, sizeof(double) * s->n[level]); break;
  }
}

/* The step order of encodePhysicsStep (BloodDropEngine.ts:1886-2069) and
 * pass B's dye transport (:2096-2108): with the wave coupling, the wave step
 * and the lid pressure first (:1886-1914); then advect, Casson + buoyancy,
 * RHS, V-cycles, project, face->centre, dye.  The advection reads the
 * centre velocity of the previous step's projected faces (:2067-2069), formed
 * here at the top of each step; the post-advection reconstruction
 * (:1922-1925) feeds only the curl, inside casson_stage with the scene. */
int orc_step(orc_state *s, double dt, int nsteps, double *rel_res) {
  int cycles = 0;
  double rr = 0.0;
  for (int n = 0; n < nsteps; ++n) {
    if (s->wave_on) {
      wave_step(s, dt);
      top_pressure(s, dt);
    }
    face_to_centre(s);
    advect_mac(s, dt);
    casson_stage(s, dt);
    divergence(s, s->u, s->v, s->w, s->rhs);
    if (s->p.solve_to_tol) {
      rr = rel_residual(s);
      int it = 0;
      while (rr > s->p.tol && it < s->p.max_cycles) {
        vcycle(s);
        ++it;
        rr = rel_residual(s);
      }
      cycles += it;
    } else {
      for (int c = 0; c < s->p.cycles; ++c) vcycle(s);
      cycles += s->p.cycles;
      if (rel_res) rr = rel_residual(s);
    }
    project(s);
    face_to_centre(s);
    if (s->conc_on) {  /* pass B (:2073-2108): parcels, then the concentration */
      parcels_stage(s, dt);
      conc_predict(s, dt);
      conc_resolve(s, dt);
    } else {
      dye_predict(s, dt);
      dye_resolve(s, dt);
    }
  }
  if (rel_res) *rel_res = rr;
  return cycles;
}

void orc_div_norms(orc_state *s, double *max_div, double *l2_div) {
  dims3 d = s->d[0];
  double *D = s->mgr[0];
  divergence(s, s->u, s->v, s->w, D);
  double mx = 0.0, ss = 0.0;
  for (long c = 0; c < s->n[0]; ++c) {
    mx = fmax(mx, fabs(D[c]));
    ss += D[c] * D[c];
  }
  (void)d;
  *max_div = mx;
  *l2_div = sqrt(ss / (double)s->n[0]);
}
