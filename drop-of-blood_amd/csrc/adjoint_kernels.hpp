// adjoint_kernels.hpp — reverse-mode (vector-Jacobian) kernels of one step.
//
// Each kernel recomputes the forward quantities of its stage from the stored
// stage inputs and pushes the adjoint of the stage's outputs back onto its
// inputs.  Gathers where the transpose is a fixed stencil (face <-> centre,
// divergence, projection); scatters with atomics where it is data dependent
// (semi-Lagrangian samples, the Casson neighbour stencil), so float results
// depend on the atomic order at the last bit.
#pragma once

#include "common.hpp"
#include "step_kernels.hpp"

namespace mac3d {

// trilinear sample set-up shared by evaluation, gradient and scatter
template <class T> struct Tri {
  size_t r00, r10, r01, r11;  // row bases (k0,j0) (k0,j1) (k1,j0) (k1,j1)
  int i0, i1;
  T fx, fy, fz;
};

template <class T>
__device__ inline Tri<T> tri_setup(const Dims &d, T ux, T uy, T uz) {
  Tri<T> t;
  T tx = ux * (T)d.nx - T(0.5), ty = uy * (T)d.ny - T(0.5), tz = uz * (T)d.nz - T(0.5);
  T bx = floor(tx), by = floor(ty), bz = floor(tz);
  t.fx = tx - bx;
  t.fy = ty - by;
  t.fz = tz - bz;
  int ib = (int)bx, jb = (int)by, kb = (int)bz;
  t.i0 = clampi(ib, 0, d.nx - 1);
  t.i1 = clampi(ib + 1, 0, d.nx - 1);
  int j0 = clampi(jb, 0, d.ny - 1), j1 = clampi(jb + 1, 0, d.ny - 1);
  int k0 = clampi(kb, 0, d.nz - 1), k1 = clampi(kb + 1, 0, d.nz - 1);
  t.r00 = ((size_t)k0 * d.ny + j0) * d.nx;
  t.r10 = ((size_t)k0 * d.ny + j1) * d.nx;
  t.r01 = ((size_t)k1 * d.ny + j0) * d.nx;
  t.r11 = ((size_t)k1 * d.ny + j1) * d.nx;
  return t;
}

// d sample / d (normalised position), per axis
template <class T>
__device__ inline void tri_grad(const T *__restrict__ f, const Dims &d, const Tri<T> &t, T g[3]) {
  T c000 = f[t.r00 + t.i0], c100 = f[t.r00 + t.i1], c010 = f[t.r10 + t.i0], c110 = f[t.r10 + t.i1];
  T c001 = f[t.r01 + t.i0], c101 = f[t.r01 + t.i1], c011 = f[t.r11 + t.i0], c111 = f[t.r11 + t.i1];
  T ey = T(1) - t.fy, ez = T(1) - t.fz;
  T dx = ey * ez * (c100 - c000) + t.fy * ez * (c110 - c010) + ey * t.fz * (c101 - c001) +
         t.fy * t.fz * (c111 - c011);
  T x00 = mixv(c000, c100, t.fx), x10 = mixv(c010, c110, t.fx);
  T x01 = mixv(c001, c101, t.fx), x11 = mixv(c011, c111, t.fx);
  T dy = ez * (x10 - x00) + t.fz * (x11 - x01);
  T dz = mixv(x00, x10, t.fy);
  dz = mixv(x01, x11, t.fy) - dz;
  // a clamped index pair (i0 == i1) has no slope: c1 - c0 == 0 already
  g[0] = dx * (T)d.nx;
  g[1] = dy * (T)d.ny;
  g[2] = dz * (T)d.nz;
}

// d sample / d position of the three packed centre-velocity components:
// g[a][b] = d v_a / d position_b
template <class T>
__device__ inline void tri_grad4(const typename cvec_of<T>::type *__restrict__ f, const Dims &d,
                                 const Tri<T> &t, T g[3][3]) {
  auto c000 = f[t.r00 + t.i0], c100 = f[t.r00 + t.i1], c010 = f[t.r10 + t.i0], c110 = f[t.r10 + t.i1];
  auto c001 = f[t.r01 + t.i0], c101 = f[t.r01 + t.i1], c011 = f[t.r11 + t.i0], c111 = f[t.r11 + t.i1];
  T ey = T(1) - t.fy, ez = T(1) - t.fz;
#define MAC3D_G(C, a)                                                                    \
  {                                                                                      \
    T dx = ey * ez * (c100.C - c000.C) + t.fy * ez * (c110.C - c010.C) +                 \
           ey * t.fz * (c101.C - c001.C) + t.fy * t.fz * (c111.C - c011.C);              \
    T x00 = mixv(c000.C, c100.C, t.fx), x10 = mixv(c010.C, c110.C, t.fx);                \
    T x01 = mixv(c001.C, c101.C, t.fx), x11 = mixv(c011.C, c111.C, t.fx);                \
    T dy = ez * (x10 - x00) + t.fz * (x11 - x01);                                        \
    T dz = mixv(x01, x11, t.fy) - mixv(x00, x10, t.fy);                                  \
    g[a][0] = dx * (T)d.nx;                                                              \
    g[a][1] = dy * (T)d.ny;                                                              \
    g[a][2] = dz * (T)d.nz;                                                              \
  }
  MAC3D_G(x, 0)
  MAC3D_G(y, 1)
  MAC3D_G(z, 2)
#undef MAC3D_G
}

template <class T>
__device__ inline void tri_scatter(T *__restrict__ lam, const Tri<T> &t, T val) {
  T ex = T(1) - t.fx, ey = T(1) - t.fy, ez = T(1) - t.fz;
  atomicAdd(&lam[t.r00 + t.i0], val * ex * ey * ez);
  atomicAdd(&lam[t.r00 + t.i1], val * t.fx * ey * ez);
  atomicAdd(&lam[t.r10 + t.i0], val * ex * t.fy * ez);
  atomicAdd(&lam[t.r10 + t.i1], val * t.fx * t.fy * ez);
  atomicAdd(&lam[t.r01 + t.i0], val * ex * ey * t.fz);
  atomicAdd(&lam[t.r01 + t.i1], val * t.fx * ey * t.fz);
  atomicAdd(&lam[t.r11 + t.i0], val * ex * t.fy * t.fz);
  atomicAdd(&lam[t.r11 + t.i1], val * t.fx * t.fy * t.fz);
}

// derivative of clamp(x, lo, hi) with respect to x
template <class T> __device__ inline T dclamp(T x, T lo, T hi) {
  return (x > lo && x < hi) ? T(1) : T(0);
}

// ---------------------------------------------------------------------------
// face -> centre transpose (FLUID_MAC_TO_CENTER, fluid.ts:917-943):
// c_u(i) = (u(i-1) + u(i))/2 with wall faces 0, so lam_u(i) += (lam_cu(i) +
// lam_cu(i+1))/2 for i < nx-1; likewise v (top face free) and w.
template <class T>
__global__ void __launch_bounds__(256)
k_adj_face_to_centre(Dims d, const T *__restrict__ lcu, const T *__restrict__ lcv,
                     const T *__restrict__ lcw, T *__restrict__ lu, T *__restrict__ lv,
                     T *__restrict__ lw) {
  MAC3D_CELL_PROLOGUE
  if (i < d.nx - 1) lu[c] += T(0.5) * (lcu[c] + lcu[c + sx]);
  lv[c] += T(0.5) * (lcv[c] + ((j < d.ny - 1) ? lcv[c + sy] : T(0)));
  if (k < d.nz - 1) lw[c] += T(0.5) * (lcw[c] + lcw[c + sz]);
}

// Projection and divergence transposes in one gather
// (FLUID_MAC_PROJECT fluid.ts:863-911, FLUID_MAC_RHS fluid.ts:517-555):
// u' = u - (p(i+1) - p(i)) on interior faces, so lam_p = D(lam_u') with the
// divergence's wall rule; rhs = D(u), so lam_u += lam_rhs(i) - lam_rhs(i+1).
// k_adj_project_p: lam_p = D(lam')     (lam' masked at the solid faces)
template <class T, class R>
__global__ void __launch_bounds__(256)
k_adj_project_p(Dims d, const T *__restrict__ lu, const T *__restrict__ lv,
                const T *__restrict__ lw, R *__restrict__ lp) {
  MAC3D_CELL_PROLOGUE
  typedef typename compute_of<R>::type C;
  C uR = (i >= d.nx - 1) ? C(0) : (C)lu[c];
  C uL = (i == 0) ? C(0) : (C)lu[c - sx];
  C vT = (C)lv[c];
  C vB = (j == 0) ? C(0) : (C)lv[c - sy];
  C wF = (k >= d.nz - 1) ? C(0) : (C)lw[c];
  C wB = (k == 0) ? C(0) : (C)lw[c - sz];
  lp[c] = (R)((uR - uL) + (vT - vB) + (wF - wB));
}

// lam_velc = mask(lam') + G^T-part of the divergence: face (i) between cells
// i and i+1 gets lam_rhs(i) - lam_rhs(i+1) (top face: lam_rhs(top) only)
template <class T, class R>
__global__ void __launch_bounds__(256)
k_adj_div(Dims d, const T *__restrict__ lu1, const T *__restrict__ lv1, const T *__restrict__ lw1,
          const R *__restrict__ lrhs, T *__restrict__ lu, T *__restrict__ lv, T *__restrict__ lw) {
  MAC3D_CELL_PROLOGUE
  typedef typename compute_of<R>::type C;
  C r = (C)lrhs[c];
  lu[c] = (i < d.nx - 1) ? (T)((C)lu1[c] + r - (C)lrhs[c + sx]) : T(0);
  lv[c] = (T)((C)lv1[c] + r - ((j < d.ny - 1) ? (C)lrhs[c + sy] : C(0)));
  lw[c] = (k < d.nz - 1) ? (T)((C)lw1[c] + r - (C)lrhs[c + sz]) : T(0);
}

// ---------------------------------------------------------------------------
// Dye transport transposes (CONC_PREDICT advect.ts:28-72, MacCormack core
// advect.ts:196-222).  Resolve first: the clamp passes the adjoint to the
// corrected value when it is inside the 8-texel bounds and to the texel that
// set the bound otherwise; the corrected value pred + s (orig - est) then
// feeds pred, orig and the forward estimate, whose sample point moves with
// the centred velocity at the cell.
template <class T>
__global__ void __launch_bounds__(256)
k_adj_dye_resolve(Dims d, T sc0, T sc1, T sc2, T strength,
                  const typename cvec_of<T>::type *__restrict__ cvel, const T *__restrict__ dye,
                  const T *__restrict__ pred, const T *__restrict__ ldye_out,
                  T *__restrict__ ldye, T *__restrict__ lpred, T *__restrict__ lcu,
                  T *__restrict__ lcv, T *__restrict__ lcw) {
  MAC3D_CELL_PROLOGUE
  const T g = ldye_out[c];
  if (g == T(0)) return;
  const T sc[3] = {sc0, sc1, sc2};
  const T lo = T(0.001), hi = T(0.999);
  T uvw[3] = {((T)i + T(0.5)) / (T)d.nx, ((T)j + T(0.5)) / (T)d.ny, ((T)k + T(0.5)) / (T)d.nz};
  T v0[3], back[3];
  dye_back(d, cvel, sc, uvw, v0, back);
  T fr[3], f[3];
  for (int a = 0; a < 3; ++a) {
    fr[a] = uvw[a] + v0[a] * sc[a];
    f[a] = clampv(fr[a], lo, hi);
  }
  Tri<T> tf = tri_setup(d, f[0], f[1], f[2]);
  T est = mixv(mixv(mixv(pred[tf.r00 + tf.i0], pred[tf.r00 + tf.i1], tf.fx),
                    mixv(pred[tf.r10 + tf.i0], pred[tf.r10 + tf.i1], tf.fx), tf.fy),
               mixv(mixv(pred[tf.r01 + tf.i0], pred[tf.r01 + tf.i1], tf.fx),
                    mixv(pred[tf.r11 + tf.i0], pred[tf.r11 + tf.i1], tf.fx), tf.fy),
               tf.fz);
  T corr = pred[c] + strength * (dye[c] - est);
  int bi = (int)floor(back[0] * (T)d.nx - T(0.5));
  int bj = (int)floor(back[1] * (T)d.ny - T(0.5));
  int bk = (int)floor(back[2] * (T)d.nz - T(0.5));
  size_t imin = 0, imax = 0;
  T mn = T(0), mx = T(0);
  for (int q = 0; q < 8; ++q) {
    int ii = clampi(bi + (q & 1), 0, d.nx - 1), jj = clampi(bj + ((q >> 1) & 1), 0, d.ny - 1),
        kk = clampi(bk + (q >> 2), 0, d.nz - 1);
    size_t id = idx3(d, ii, jj, kk);
    T s = dye[id];
    if (q == 0 || s < mn) { mn = s; imin = id; }
    if (q == 0 || s > mx) { mx = s; imax = id; }
  }
  if (corr < mn) { atomicAdd(&ldye[imin], g); return; }
  if (corr > mx) { atomicAdd(&ldye[imax], g); return; }
  atomicAdd(&lpred[c], g);
  atomicAdd(&ldye[c], strength * g);
  T ge = -strength * g;  // adjoint of est
  tri_scatter(lpred, tf, ge);
  T gr[3];
  tri_grad(pred, d, tf, gr);
  T gv0[3];
  for (int a = 0; a < 3; ++a) gv0[a] = ge * gr[a] * dclamp(fr[a], lo, hi) * sc[a];
  Tri<T> t0 = tri_setup(d, uvw[0], uvw[1], uvw[2]);
  tri_scatter(lcu, t0, gv0[0]);
  tri_scatter(lcv, t0, gv0[1]);
  tri_scatter(lcw, t0, gv0[2]);
}

// Transpose of the RK2 backtrace-and-sample g = S(F, back(x)), back from the
// centred velocity (sample at p0, midpoint, sample at mid): the adjoint gv of
// the traced position back is pushed through both velocity samples.
// lo/hi: the clamp of the trace points; shift: the sample-point shift per
// axis (MAC faces) with its own clamp sl/sh.
template <class T>
__device__ inline void adj_trace(const Dims &d, const typename cvec_of<T>::type *__restrict__ cvel,
                                 const T sc[3], const T f[3], bool clamp_p0, T lo, T hi,
                                 const T gback[3], T *__restrict__ lcu, T *__restrict__ lcv,
                                 T *__restrict__ lcw) {
  T p0[3], v0[3], midr[3], mid[3], vm[3], backr[3];
  for (int a = 0; a < 3; ++a) p0[a] = clamp_p0 ? clampv(f[a], lo, hi) : f[a];
  Tri<T> t0 = tri_setup(d, p0[0], p0[1], p0[2]);
  sample_vec<T>(cvel, d, p0[0], p0[1], p0[2], v0);
  for (int a = 0; a < 3; ++a) {
    midr[a] = f[a] - T(0.5) * (v0[a] * sc[a]);
    mid[a] = clampv(midr[a], lo, hi);
  }
  Tri<T> tm = tri_setup(d, mid[0], mid[1], mid[2]);
  sample_vec<T>(cvel, d, mid[0], mid[1], mid[2], vm);
  T gvm[3];
  for (int a = 0; a < 3; ++a) {
    backr[a] = f[a] - vm[a] * sc[a];
    gvm[a] = -gback[a] * dclamp(backr[a], lo, hi) * sc[a];
  }
  tri_scatter(lcu, tm, gvm[0]);
  tri_scatter(lcv, tm, gvm[1]);
  tri_scatter(lcw, tm, gvm[2]);
  // d vm_a / d mid_b
  T ga[3][3];
  tri_grad4<T>(cvel, d, tm, ga);
  T gmid[3];
  for (int b = 0; b < 3; ++b) gmid[b] = gvm[0] * ga[0][b] + gvm[1] * ga[1][b] + gvm[2] * ga[2][b];
  T gv0[3];
  for (int a = 0; a < 3; ++a) gv0[a] = -T(0.5) * gmid[a] * dclamp(midr[a], lo, hi) * sc[a];
  tri_scatter(lcu, t0, gv0[0]);
  tri_scatter(lcv, t0, gv0[1]);
  tri_scatter(lcw, t0, gv0[2]);
}

template <class T>
__global__ void __launch_bounds__(256)
k_adj_dye_predict(Dims d, T sc0, T sc1, T sc2, const typename cvec_of<T>::type *__restrict__ cvel,
                  const T *__restrict__ dye, const T *__restrict__ lpred, T *__restrict__ ldye,
                  T *__restrict__ lcu, T *__restrict__ lcv, T *__restrict__ lcw) {
  MAC3D_CELL_PROLOGUE
  const T g = lpred[c];
  if (g == T(0)) return;
  const T sc[3] = {sc0, sc1, sc2};
  T uvw[3] = {((T)i + T(0.5)) / (T)d.nx, ((T)j + T(0.5)) / (T)d.ny, ((T)k + T(0.5)) / (T)d.nz};
  T v0[3], back[3];
  dye_back(d, cvel, sc, uvw, v0, back);
  Tri<T> tb = tri_setup(d, back[0], back[1], back[2]);
  tri_scatter(ldye, tb, g);
  T gr[3], gb[3];
  tri_grad(dye, d, tb, gr);
  for (int a = 0; a < 3; ++a) gb[a] = g * gr[a];
  adj_trace(d, cvel, sc, uvw, false, T(0.001), T(0.999), gb, lcu, lcv, lcw);
}

// Transpose of FLUID_ADVECT_MAC (fluid.ts:144-192): per face, the sampled
// staggered component scatters into its source field and the traced point
// moves with the centred velocity at p0 and the midpoint.
template <class T>
__global__ void __launch_bounds__(256)
k_adj_advect(Dims d, T sc0, T sc1, T sc2, T damp, const T *__restrict__ u,
             const T *__restrict__ v, const T *__restrict__ w,
             const typename cvec_of<T>::type *__restrict__ cvel, const T *__restrict__ lout_u,
             const T *__restrict__ lout_v, const T *__restrict__ lout_w, T *__restrict__ lu,
             T *__restrict__ lv, T *__restrict__ lw, T *__restrict__ lcu, T *__restrict__ lcv,
             T *__restrict__ lcw) {
  MAC3D_CELL_PROLOGUE
  const T sc[3] = {sc0, sc1, sc2};
  const T lo = T(0.001), hi = T(0.999);
  const T *fld[3] = {u, v, w};
  T *lfl[3] = {lu, lv, lw};
  const T gouts[3] = {(i >= d.nx - 1) ? T(0) : lout_u[c], lout_v[c],
                      (k >= d.nz - 1) ? T(0) : lout_w[c]};
  for (int a = 0; a < 3; ++a) {
    T g = gouts[a] * damp;
    if (g == T(0)) continue;
    T f[3] = {((T)i + T(0.5)) / (T)d.nx, ((T)j + T(0.5)) / (T)d.ny, ((T)k + T(0.5)) / (T)d.nz};
    f[a] = ((a == 0 ? (T)i : (a == 1 ? (T)j : (T)k)) + T(1)) /
           (T)(a == 0 ? d.nx : (a == 1 ? d.ny : d.nz));
    T back[3];
    trace_back(d, cvel, sc, f, back);
    T qr[3], q[3];
    for (int b = 0; b < 3; ++b) {
      T shift = (b == a) ? T(0.5) / (T)(b == 0 ? d.nx : (b == 1 ? d.ny : d.nz)) : T(0);
      qr[b] = back[b] - shift;
      q[b] = clampv(qr[b], lo, hi);
    }
    Tri<T> tq = tri_setup(d, q[0], q[1], q[2]);
    tri_scatter(lfl[a], tq, g);
    T gr[3], gb[3];
    tri_grad(fld[a], d, tq, gr);
    for (int b = 0; b < 3; ++b) gb[b] = g * gr[b] * dclamp(qr[b], lo, hi);
    adj_trace(d, cvel, sc, f, true, T(0.002), T(0.998), gb, lcu, lcv, lcw);
  }
}

// scatter of a neighbour-face adjoint onto its source with the wall rules of
// face_u / face_v / face_w
template <class T>
__device__ inline void scat_u(const Dims &d, T *__restrict__ l, int i, int j, int k, T val) {
  if (i < 0 || i >= d.nx) return;
  if (j < 0) { atomicAdd(&l[idx3(d, i, 0, k)], -val); return; }
  if (j >= d.ny) { atomicAdd(&l[idx3(d, i, d.ny - 1, k)], val); return; }
  if (k < 0) { atomicAdd(&l[idx3(d, i, j, 0)], -val); return; }
  if (k >= d.nz) { atomicAdd(&l[idx3(d, i, j, d.nz - 1)], -val); return; }
  atomicAdd(&l[idx3(d, i, j, k)], val);
}
template <class T>
__device__ inline void scat_v(const Dims &d, T *__restrict__ l, int i, int j, int k, T val) {
  if (j < 0) return;
  if (j >= d.ny) { atomicAdd(&l[idx3(d, i, d.ny - 1, k)], val); return; }
  if (i < 0) { atomicAdd(&l[idx3(d, 0, j, k)], -val); return; }
  if (i >= d.nx) { atomicAdd(&l[idx3(d, d.nx - 1, j, k)], -val); return; }
  if (k < 0) { atomicAdd(&l[idx3(d, i, j, 0)], -val); return; }
  if (k >= d.nz) { atomicAdd(&l[idx3(d, i, j, d.nz - 1)], -val); return; }
  atomicAdd(&l[idx3(d, i, j, k)], val);
}
template <class T>
__device__ inline void scat_w(const Dims &d, T *__restrict__ l, int i, int j, int k, T val) {
  if (k < 0 || k >= d.nz) return;
  if (i < 0) { atomicAdd(&l[idx3(d, 0, j, k)], -val); return; }
  if (i >= d.nx) { atomicAdd(&l[idx3(d, d.nx - 1, j, k)], -val); return; }
  if (j < 0) { atomicAdd(&l[idx3(d, i, 0, k)], -val); return; }
  if (j >= d.ny) { atomicAdd(&l[idx3(d, i, d.ny - 1, k)], val); return; }
  atomicAdd(&l[idx3(d, i, j, k)], val);
}

// Transpose of the Casson viscous update and buoyancy (FLUID_CASSON_FORCES,
// fluid.ts:377-511): out = m + alpha(gamma(m), c) lap(m) (+ buoyancy on v).
// Linearisation: d out/d nb = alpha + (sum_b G_b lap_b) d alpha/d gamma
// d gamma/d nb per neighbour value, through the Casson law
// mu = (sqrt(mu_inf(c)) + sqrt(tau_y(c)/(gamma + reg)))^2 (the max(., reg)
// floor never binds since gamma > 0), alpha clamped to [0, alpha_max] (zero
// slope when clamped).  Also d/d dye and d/d rho_b (per-block partials).
template <class T>
__global__ void __launch_bounds__(256)
k_adj_casson(Dims d, StepParams P, const T *__restrict__ u, const T *__restrict__ v,
             const T *__restrict__ w, const T *__restrict__ dye, const T *__restrict__ lou,
             const T *__restrict__ lov, const T *__restrict__ low, T *__restrict__ lu,
             T *__restrict__ lv, T *__restrict__ lw, T *__restrict__ ldye,
             double *__restrict__ part) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int k = blockIdx.z;
  double drho = 0.0;
  if (i < d.nx && j < d.ny) {
    size_t c = idx3(d, i, j, k);
    const size_t sy = (size_t)d.nx;
    const T inv2h = T(1.0 / (2.0 * P.dx));
    const T dtoh2 = T(P.dt / (P.dx * P.dx));
    const T buoy = T(-P.g * (P.rho_b - P.rho_w) / P.rho_w * P.dt);
    T G[3] = {(i >= d.nx - 1) ? T(0) : lou[c], lov[c], (k >= d.nz - 1) ? T(0) : low[c]};
    T m[3] = {u[c], v[c], w[c]};
    // neighbour values: [dir][comp], dir = x+, x-, y+, y-, z+, z-
    const int di[6] = {1, -1, 0, 0, 0, 0}, dj[6] = {0, 0, 1, -1, 0, 0}, dk[6] = {0, 0, 0, 0, 1, -1};
    T nb[6][3];
    for (int q = 0; q < 6; ++q) {
      nb[q][0] = face_u(d, u, i + di[q], j + dj[q], k + dk[q]);
      nb[q][1] = face_v(d, v, i + di[q], j + dj[q], k + dk[q]);
      nb[q][2] = face_w(d, w, i + di[q], j + dj[q], k + dk[q]);
    }
    T ss = T(0);
    for (int a = 0; a < 3; ++a)
      for (int ax = 0; ax < 3; ++ax) {
        T df = nb[2 * ax][a] - nb[2 * ax + 1][a];
        ss += df * df;
      }
    T sroot = sqrt(ss + T(P.shear_eps2));
    T gam = inv2h * sroot;
    T cr = dye[c];
    T cc = clampv(cr, T(0), T(1));
    T muw = T(P.mu_w), dmu = T(P.mu_inf - P.mu_w);
    T sq = sqrt(muw + cc * dmu);
    T reg = T(P.gamma_reg);
    T gr = gam + reg;
    T sy_ = sqrt(T(P.tau_y) / gr);  // sqrt(tau_y / (gamma + reg))
    T yterm = cc * sy_;
    T mu = (sq + yterm) * (sq + yterm);
    T rho = T(P.rho_w) + cc * T(P.rho_b - P.rho_w);
    T araw = mu / rho * dtoh2;
    T alpha = clampv(araw, T(0), T(P.alpha_max));
    T aslope = dclamp(araw, T(0), T(P.alpha_max));
    T lap[3];
    for (int a = 0; a < 3; ++a)
      lap[a] = (nb[0][a] + nb[1][a] + nb[2][a] + nb[3][a] + nb[4][a] + nb[5][a]) - T(6) * m[a];
    T Gl = G[0] * lap[0] + G[1] * lap[1] + G[2] * lap[2];
    // d alpha / d gamma: d yterm / d gamma = -yterm / (2 (gamma + reg))
    T dmu_dg = T(2) * (sq + yterm) * (-yterm / (T(2) * gr));
    T H = Gl * aslope * dmu_dg / rho * dtoh2 * inv2h / sroot;
    // own cell
    atomicAdd(&lu[c], G[0] * (T(1) - T(6) * alpha));
    atomicAdd(&lv[c], G[1] * (T(1) - T(6) * alpha));
    atomicAdd(&lw[c], G[2] * (T(1) - T(6) * alpha));
    for (int q = 0; q < 6; ++q) {
      int ax = q >> 1;
      T sgn = (q & 1) ? T(-1) : T(1);
      T val[3];
      for (int a = 0; a < 3; ++a)
        val[a] = G[a] * alpha + sgn * H * (nb[2 * ax][a] - nb[2 * ax + 1][a]);
      scat_u(d, lu, i + di[q], j + dj[q], k + dk[q], val[0]);
      scat_v(d, lv, i + di[q], j + dj[q], k + dk[q], val[1]);
      scat_w(d, lw, i + di[q], j + dj[q], k + dk[q], val[2]);
    }
    // dye: alpha(c) and buoyancy on the face above the cell
    T cslope = dclamp(cr, T(0), T(1));
    T dsq = dmu / (T(2) * sq);
    T dmu_dc = T(2) * (sq + yterm) * (dsq + sy_);
    T drho_dc = T(P.rho_b - P.rho_w);
    T dalpha_dc = aslope * (dmu_dc / rho - mu * drho_dc / (rho * rho)) * dtoh2;
    T gdye = Gl * dalpha_dc * cslope;
    atomicAdd(&ldye[c], gdye + T(0.5) * buoy * G[1]);
    atomicAdd(&ldye[(j + 1 < d.ny) ? c + sy : c], T(0.5) * buoy * G[1]);
    // d / d rho_b (rho_w fixed): through rho(c) in alpha and the buoyancy
    T dalpha_drho = -aslope * mu * cc / (rho * rho) * dtoh2;
    T dyef = T(0.5) * (dye[c] + dye[(j + 1 < d.ny) ? c + sy : c]);
    drho = (double)(Gl * dalpha_drho) + (double)G[1] * (-P.g / P.rho_w * P.dt) * (double)dyef;
  }
  drho = wave_sum(drho);
  __shared__ double s[4];
  int t = threadIdx.y * blockDim.x + threadIdx.x;
  if ((t & 63) == 0) s[t >> 6] = drho;
  __syncthreads();
  if (t == 0) {
    size_t blk = ((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
    part[blk] = s[0] + s[1] + s[2] + s[3];
  }
}

// sum of n doubles into out[0] (+=), one block of 256
__global__ void __launch_bounds__(256)
k_sum_into(const double *__restrict__ part, size_t n, double *__restrict__ out) {
  double a = 0.0;
  for (size_t q = threadIdx.x; q < n; q += 256) a += part[q];
  a = wave_sum(a);
  __shared__ double s[4];
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) out[0] += s[0] + s[1] + s[2] + s[3];
}

// Transpose of the blob initial condition: per blob
// d/d(cx, cy, cz, radius, w0) of dye (cell centres) and V (y-faces),
// 1 - smoothstep(r - 1/2, r + 1/2, dist) with smoothstep' = 6 t (1 - t).
// Grid-stride; per-block partials part[blk * 5 * nb + 5 q + p].
template <class T>
__global__ void __launch_bounds__(256)
k_adj_init(Dims d, const BlobDev *__restrict__ blobs, int nb, const T *__restrict__ ldye,
           const T *__restrict__ lv, double *__restrict__ part) {
  extern __shared__ double sred[];  // 4 waves x 5 nb
  const int nparm = 5 * nb;
  double acc[5 * 16];
  for (int q = 0; q < nparm; ++q) acc[q] = 0.0;
  size_t n = d.cells();
  for (size_t c = (size_t)blockIdx.x * 256 + threadIdx.x; c < n; c += (size_t)gridDim.x * 256) {
    int i = (int)(c % d.nx), j = (int)((c / d.nx) % d.ny), k = (int)(c / ((size_t)d.nx * d.ny));
    double gd = (double)ldye[c], gv = (double)lv[c];
    for (int q = 0; q < nb; ++q) {
      BlobDev b = blobs[q];
      // cell centre (dye)
      double ex = i + 0.5 - b.cx, ey = j + 0.5 - b.cy, ez = k + 0.5 - b.cz;
      double r = sqrt(ex * ex + ey * ey + ez * ez);
      double t = (r - (b.radius - 0.5));
      if (gd != 0.0 && t > 0.0 && t < 1.0 && r > 0.0) {
        double sp = 6.0 * t * (1.0 - t);  // d smoothstep / d r
        // dye = 1 - ss: d/d centre = sp * (x - c)/r, d/d radius = sp
        acc[5 * q + 0] += gd * sp * ex / r;
        acc[5 * q + 1] += gd * sp * ey / r;
        acc[5 * q + 2] += gd * sp * ez / r;
        acc[5 * q + 3] += gd * sp;
      }
      // y-face (v = -w0 (1 - ss))
      double fy = j + 1.0 - b.cy;
      double rf = sqrt(ex * ex + fy * fy + ez * ez);
      if (gv != 0.0) {
        double tf = rf - (b.radius - 0.5);
        double prof = 1.0 - smoothstep_v(b.radius - 0.5, b.radius + 0.5, rf);
        acc[5 * q + 4] += -gv * prof;
        if (tf > 0.0 && tf < 1.0 && rf > 0.0) {
          double sp = 6.0 * tf * (1.0 - tf);
          acc[5 * q + 0] += -gv * b.w0 * sp * ex / rf;
          acc[5 * q + 1] += -gv * b.w0 * sp * fy / rf;
          acc[5 * q + 2] += -gv * b.w0 * sp * ez / rf;
          acc[5 * q + 3] += -gv * b.w0 * sp;
        }
      }
    }
  }
  for (int q = 0; q < nparm; ++q) {
    double a = wave_sum(acc[q]);
    if ((threadIdx.x & 63) == 0) sred[(threadIdx.x >> 6) * nparm + q] = a;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < nparm; q += 256)
    part[(size_t)blockIdx.x * nparm + q] =
        sred[q] + sred[nparm + q] + sred[2 * nparm + q] + sred[3 * nparm + q];
}

template <class T>
__global__ void __launch_bounds__(256) k_axpy(size_t n, T a, const T *__restrict__ x, T *__restrict__ y) {
  size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t < n) y[t] += a * x[t];
}

template <class T>
__global__ void __launch_bounds__(256)
k_sub_sq(size_t n, const T *__restrict__ a, const T *__restrict__ b, T *__restrict__ out,
         double *__restrict__ part) {
  size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  double s = 0.0;
  if (t < n) {
    T dlt = a[t] - b[t];
    out[t] = dlt;
    s = (double)dlt * (double)dlt;
  }
  s = wave_sum(s);
  __shared__ double sm[4];
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0] + sm[1] + sm[2] + sm[3];
}

}  // namespace mac3d
