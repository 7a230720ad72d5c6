// step_kernels.hpp — the non-solver kernels of one MAC step.
//
// One thread per cell, launched as (64, 4, 1) blocks over (x, y) with z on
// grid.z, so a wavefront covers 64 consecutive x cells of one row: every
// plain load/store is a coalesced 256 B (FP32) row segment.
#pragma once

#include "common.hpp"
#include "wave_kernels.hpp"

namespace mac3d {

struct StepParams {
  double dt, dx;
  double rho_w, mu_w, rho_b, tau_y, mu_inf, g;
  double gamma_reg, shear_eps2, alpha_max, vel_damp, maccormack;
  // the coupling's reference physics: Casson scale (lensAndRheology.z),
  // buoyancy strength (fluidParams.y), the world height of the lid (the
  // still surface) and the tank depth
  double casson_scale, buoyancy, surface_y, depth;
};

template <class T>
__device__ inline T smoothstep_v(T e0, T e1, T x) {
  T t = clampv((x - e0) / (e1 - e0), T(0), T(1));
  return t * t * (T(3) - T(2) * t);
}

#define MAC3D_CELL_PROLOGUE                                  \
  int i = blockIdx.x * blockDim.x + threadIdx.x;             \
  int j = blockIdx.y * blockDim.y + threadIdx.y;             \
  int k = blockIdx.z;                                        \
  if (i >= d.nx || j >= d.ny) return;                        \
  size_t c = idx3(d, i, j, k);                               \
  const size_t sx = 1, sy = (size_t)d.nx, sz = (size_t)d.nx * d.ny; \
  (void)sx; (void)sy; (void)sz;

// FLUID_MAC_TO_CENTER (src/shaders/fluid.ts:917-943): centre velocity is the
// mean of the two faces of each axis, solid wall faces read as 0.  SCENE: the
// top-row stagnation gate (fluid.ts:944-959) from the drop slots.
template <class T, bool SCENE>
__global__ void __launch_bounds__(256)
k_face_to_centre(Dims d, const T *__restrict__ u, const T *__restrict__ v,
                 const T *__restrict__ w, typename cvec_of<T>::type *__restrict__ cvel,
                 WaveParams wp, const WaveDrop *__restrict__ drops) {
  MAC3D_CELL_PROLOGUE
  T uR = (i >= d.nx - 1) ? T(0) : u[c];
  T uL = (i == 0) ? T(0) : u[c - sx];
  T vU = v[c];
  T vD = (j == 0) ? T(0) : v[c - sy];
  T wF = (k >= d.nz - 1) ? T(0) : w[c];
  T wB = (k == 0) ? T(0) : w[c - sz];
  typename cvec_of<T>::type o;
  o.x = T(0.5) * (uL + uR);
  o.y = T(0.5) * (vD + vU);
  o.z = T(0.5) * (wB + wF);
  if constexpr (SCENE)
    if (j >= d.ny - 2 && o.y < T(0)) o.y *= top_stagnation_gate<T>(d, i, j, k, wp, drops);
  cvel[c] = o;
}

// traceBackUVW (fluid.ts:169-174): RK2 midpoint backtrace on the centred
// velocity in normalised coordinates, clamped to [0.002, 0.998].
template <class T>
__device__ inline void trace_back(const Dims &d,
                                  const typename cvec_of<T>::type *__restrict__ cvel,
                                  const T sc[3], const T f[3], T out[3]) {
  const T lo = T(0.002), hi = T(0.998);
  T v0[3], vm[3];
  sample_vec<T>(cvel, d, clampv(f[0], lo, hi), clampv(f[1], lo, hi), clampv(f[2], lo, hi), v0);
  T mx = clampv(f[0] - T(0.5) * (v0[0] * sc[0]), lo, hi);
  T my = clampv(f[1] - T(0.5) * (v0[1] * sc[1]), lo, hi);
  T mz = clampv(f[2] - T(0.5) * (v0[2] * sc[2]), lo, hi);
  sample_vec<T>(cvel, d, mx, my, mz, vm);
  out[0] = clampv(f[0] - vm[0] * sc[0], lo, hi);
  out[1] = clampv(f[1] - vm[1] * sc[1], lo, hi);
  out[2] = clampv(f[2] - vm[2] * sc[2], lo, hi);
}

// traceBackUVW in texel space: the face point f (the cell's texel index +
// offset), RK2 displacement in texels v dt/h, clamps at [0.002 N - 1/2,
// 0.998 N - 1/2] (bw, relative to the cell).  f is the +1/2 face of cell c
// along axis a (c and the axis stride sa in the reader's offset unit): unless
// the clamp moves it, the first sample's weights are 1/2 along a and 0 on the
// other axes, so it is mixv(cvel[c], cvel[c + sa], 1/2) -- the value of the
// 8-texel sample (mixv(x, y, 0) = x exactly) from 2 loads.
template <class T, class CR>
__device__ inline void trace_back_tp(const Dims &d, const CR &cvel, T dth, const TPos<T> &f,
                                     const TBounds<T> &bw, typename CR::index c,
                                     typename CR::index sa, TPos<T> &out, T &nanc) {
  TPos<T> p0 = f;
  tpos_clamp(p0, bw);
  T v0[3], vm[3];
  bool moved = false;
#pragma unroll
  for (int q = 0; q < 3; ++q) moved |= p0.b[q] != f.b[q] || p0.o[q] != f.o[q];
  if (!moved) {
    auto a0 = cvel.ld(c), a1 = cvel.ld(c + sa);
    v0[0] = mixv(a0.x, a1.x, T(0.5));
    v0[1] = mixv(a0.y, a1.y, T(0.5));
    v0[2] = mixv(a0.z, a1.z, T(0.5));
  } else {
    sample_vec_tp_r<T>(cvel, d, p0, v0);
  }
  TPos<T> m = f;
  for (int a = 0; a < 3; ++a) m.o[a] -= T(0.5) * (v0[a] * dth);
  tpos_clamp(m, bw);
  sample_vec_tp_r<T>(cvel, d, m, vm);
  out = f;
  for (int a = 0; a < 3; ++a) out.o[a] -= vm[a] * dth;
  tpos_clamp(out, bw);
  nanc = trace_nan(v0, vm);
}

// FLUID_ADVECT_MAC (fluid.ts:144-192, 323-325): each staggered component is
// traced back from its own face and sampled with the -1/2 texel shift of
// sampleMacU/V/W (fluid.ts:154-167), then multiplied by the damping factor;
// the two solid wall faces stay 0.  sc = dt / domain size per axis.
// (FP32: at least 5 waves per SIMD, more gathers in flight; 6 spill since
// the traces carry their non-finite check, and 8 forced too few registers.
// FP64 needs its registers.  The plain FP32 step takes k_advect_mac_lds.)
// BUF: fields read through buffer
// resources with 32-bit byte offsets (fields under 2^31 bytes, planes under
// 2^24).  SCENE: the coupling's scene terms (advect_scene_terms) from the
// drop slots and the wave state.
template <class T, bool BUF, bool SCENE>
__global__ void __launch_bounds__(256, sizeof(T) == 4 ? 5 : 1)
k_advect_mac(Dims d, T dth, T damp, const T *__restrict__ u, const T *__restrict__ v,
             const T *__restrict__ w, const typename cvec_of<T>::type *__restrict__ cvel,
             T *__restrict__ uo, T *__restrict__ vo, T *__restrict__ wo, WaveParams wp,
             const WaveDrop *__restrict__ drops, const DropPhase *__restrict__ phs,
             const float4 *__restrict__ wave) {
  MAC3D_CELL_PROLOGUE
  const size_t n = d.cells();
  const auto cv = make_reader<BUF>(cvel, n);
  typedef typename decltype(cv)::index I;
  const decltype(make_reader<BUF>(u, n)) rf[3] = {make_reader<BUF>(u, n), make_reader<BUF>(v, n),
                                                  make_reader<BUF>(w, n)};
  const I cu = cv.cell(d, i, j, k);  // cell c in the reader's offset unit
  const T N[3] = {(T)d.nx, (T)d.ny, (T)d.nz};
  T lo2[3], hi2[3], lo1[3], hi1[3];
  for (int a = 0; a < 3; ++a) {
    lo2[a] = T(0.002) * N[a] - T(0.5);
    hi2[a] = T(0.998) * N[a] - T(0.5);
    lo1[a] = T(0.001) * N[a] - T(0.5);
    hi1[a] = T(0.999) * N[a] - T(0.5);
  }
  const TBounds<T> bw2 = tpos_bounds(lo2, hi2, i, j, k), bw1 = tpos_bounds(lo1, hi1, i, j, k);
  T res[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    // face +1/2 along axis a: texel (cell + 1/2) on that axis, the cell centre on the others
    TPos<T> f;
    f.b[0] = i; f.b[1] = j; f.b[2] = k;
    f.o[0] = f.o[1] = f.o[2] = T(0);
    f.o[a] = T(0.5);
    TPos<T> b;
    const I sa = (a == 0 ? 1 : (a == 1 ? (I)d.nx : (I)d.nx * d.ny)) * decltype(cv)::unit;
    T nanc;
    trace_back_tp<T>(d, cv, dth, f, bw2, cu, sa, b, nanc);
    b.o[a] -= T(0.5);  // the -1/2 texel shift of sampleMacU/V/W
    tpos_clamp(b, bw1);
    res[a] = sample_tp_r<T>(rf[a], d, b) * damp + nanc;
  }
  if constexpr (SCENE)
    advect_scene_terms<T>(d, i, j, k, wp, drops, phs, wave, res[0], res[1], res[2]);
  uo[c] = (i >= d.nx - 1) ? T(0) : res[0];
  vo[c] = res[1];
  wo[c] = (k >= d.nz - 1) ? T(0) : res[2];
}

// ---------------------------------------------------------------------------
// k_advect_mac with its samples taken from LDS (FP32).  The trilinear
// gathers of the backtraces (30 of 12 B and 24 of 4 B per cell) saturate the
// texture addresser; at the step's Courant numbers (|v| dt/h under 1) every
// sample of a cell has its corners within one cell of it.  A block stages
// its tile (AL_TX x AL_TY x AL_TZ cells) with a one-cell ring of the centre
// velocity and of u, v, w in LDS (structure of arrays, coalesced loads), the
// ring's cells outside the grid replicating the edge cells, which is the
// sampler's clamp-to-edge rule, and samples from there: the corner offsets
// are immediates and need no edge logic.  A sample whose corners leave the
// staged box falls back to the buffer gathers of k_advect_mac.  The sample
// arithmetic (weights, lerp order) is k_advect_mac's, so the results are
// identical.
#define AL_TX 32
#define AL_TY 8
#define AL_TZ 4
#define AL_EX (AL_TX + 2)
#define AL_EY (AL_TY + 2)
#define AL_EZ (AL_TZ + 2)
#define AL_N (AL_EX * AL_EY * AL_EZ)

// lower corner (staged index o) and weights of p, a sample point of the cell
// at staged index (lx, ly, lz); false when a corner lies outside the box
__device__ inline bool al_corner(const TPos<float> &p, int lx, int ly, int lz, int &o,
                                 float &fx, float &fy, float &fz) {
  const float f0 = floor(p.o[0]), f1 = floor(p.o[1]), f2 = floor(p.o[2]);
  fx = p.o[0] - f0;
  fy = p.o[1] - f1;
  fz = p.o[2] - f2;
  const int cx = lx + (int)f0, cy = ly + (int)f1, cz = lz + (int)f2;
  o = (cz * AL_EY + cy) * AL_EX + cx;
  return (unsigned)cx < AL_EX - 1 && (unsigned)cy < AL_EY - 1 && (unsigned)cz < AL_EZ - 1;
}
// The staged centre velocity interleaves its components by row
// ([plane][row][component][x]): a plane's 24 corner values of a sample then
// lie within the 8-bit dword offsets of paired LDS reads from one base.
#define AL_VR (3 * AL_EX)       // floats per staged row of the velocity
#define AL_VP (AL_EY * AL_VR)   // floats per staged plane
// lower corner in the staged velocity's layout (index v of component 0)
// and weights of p; false when a corner lies outside the box
__device__ inline bool al_corner3(const TPos<float> &p, int lx, int ly, int lz, int &v,
                                  float &fx, float &fy, float &fz) {
  const float f0 = floor(p.o[0]), f1 = floor(p.o[1]), f2 = floor(p.o[2]);
  fx = p.o[0] - f0;
  fy = p.o[1] - f1;
  fz = p.o[2] - f2;
  const int cx = lx + (int)f0, cy = ly + (int)f1, cz = lz + (int)f2;
  v = cz * AL_VP + cy * AL_VR + cx;
  return (unsigned)cx < AL_EX - 1 && (unsigned)cy < AL_EY - 1 && (unsigned)cz < AL_EZ - 1;
}
// trilinear sample of the staged velocity at corner v, the order of
// sample_vec_tp_r
__device__ inline void al_tri3(const float *f, int v, float fx, float fy, float fz, float out[3]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float *g = f + v + c * AL_EX;
    const float x00 = mixv(g[0], g[1], fx);
    const float x10 = mixv(g[AL_VR], g[AL_VR + 1], fx);
    const float x01 = mixv(g[AL_VP], g[AL_VP + 1], fx);
    const float x11 = mixv(g[AL_VP + AL_VR], g[AL_VP + AL_VR + 1], fx);
    out[c] = mixv(mixv(x00, x10, fy), mixv(x01, x11, fy), fz);
  }
}
// trilinear sample of a staged scalar at corner o (the order of sample_tp_r)
__device__ inline float al_tri(const float *f, int o, float fx, float fy, float fz) {
  constexpr int dy = AL_EX, dz = AL_EX * AL_EY;
  const float x00 = mixv(f[o], f[o + 1], fx);
  const float x10 = mixv(f[o + dy], f[o + dy + 1], fx);
  const float x01 = mixv(f[o + dz], f[o + dz + 1], fx);
  const float x11 = mixv(f[o + dz + dy], f[o + dz + dy + 1], fx);
  return mixv(mixv(x00, x10, fy), mixv(x01, x11, fy), fz);
}

template <bool SCENE>
__global__ void __launch_bounds__(AL_TX * AL_TY)
k_advect_mac_lds(Dims d, float dth, float damp, const float *__restrict__ u,
                 const float *__restrict__ v, const float *__restrict__ w,
                 const float3 *__restrict__ cvel, float *__restrict__ uo, float *__restrict__ vo,
                 float *__restrict__ wo, WaveParams wp, const WaveDrop *__restrict__ drops,
                 const DropPhase *__restrict__ phs, const float4 *__restrict__ wave) {
  __shared__ float sc[3 * AL_N];  // centre velocity, components interleaved by row
  __shared__ float sf[3][AL_N];   // u, v, w
  const int x0 = blockIdx.x * AL_TX, y0 = blockIdx.y * AL_TY, z0 = blockIdx.z * AL_TZ;
  const int tid = threadIdx.x;
  // staging by (x, y) columns through the box's planes: the column's
  // position is formed once, a plane adds the (uniform) plane offset.  A
  // thread's two columns load together (one round trip); the second is
  // clamped to the box's last column, which threads past it restage with
  // that column's own values.
  static_assert(AL_EX * AL_EY <= 2 * AL_TX * AL_TY, "two columns per thread");
  const size_t plane = (size_t)d.nx * d.ny;
  {
    int cl[2];
    float3 cvv[2][AL_EZ];
    float fu[2][AL_EZ], fv[2][AL_EZ], fw[2][AL_EZ];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int col = min(tid + h * AL_TX * AL_TY, AL_EX * AL_EY - 1);
      cl[h] = col;
      const size_t cxy = (size_t)clampi(y0 - 1 + col / AL_EX, 0, d.ny - 1) * d.nx +
                         clampi(x0 - 1 + col % AL_EX, 0, d.nx - 1);
#pragma unroll
      for (int lz = 0; lz < AL_EZ; ++lz) {
        const size_t c = cxy + (size_t)clampi(z0 - 1 + lz, 0, d.nz - 1) * plane;
        cvv[h][lz] = cvel[c];
        fu[h][lz] = u[c];
        fv[h][lz] = v[c];
        fw[h][lz] = w[c];
      }
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int col = cl[h], lx = col % AL_EX, ly = col / AL_EX;
#pragma unroll
      for (int lz = 0; lz < AL_EZ; ++lz) {
        const int t = lz * (AL_EX * AL_EY) + col, tv = lz * AL_VP + ly * AL_VR + lx;
        sc[tv] = cvv[h][lz].x;
        sc[tv + AL_EX] = cvv[h][lz].y;
        sc[tv + 2 * AL_EX] = cvv[h][lz].z;
        sf[0][t] = fu[h][lz];
        sf[1][t] = fv[h][lz];
        sf[2][t] = fw[h][lz];
      }
    }
  }
  __syncthreads();
  const int i = x0 + tid % AL_TX, j = y0 + tid / AL_TX;
  if (i >= d.nx || j >= d.ny) return;
  const size_t n = d.cells();
  const auto cvr = make_reader<true>(cvel, n);
  const BufRead<float> rf[3] = {make_reader<true>(u, n), make_reader<true>(v, n),
                                make_reader<true>(w, n)};
  const float N[3] = {(float)d.nx, (float)d.ny, (float)d.nz};
  float lo2[3], hi2[3], lo1[3], hi1[3];
  for (int a = 0; a < 3; ++a) {
    lo2[a] = 0.002f * N[a] - 0.5f;
    hi2[a] = 0.998f * N[a] - 0.5f;
    lo1[a] = 0.001f * N[a] - 0.5f;
    hi1[a] = 0.999f * N[a] - 0.5f;
  }
  const int lx = tid % AL_TX + 1, ly = tid / AL_TX + 1;
  // centre velocity at a point: from LDS, or gathered where it leaves the box
  auto vel_at = [&](const TPos<float> &p, int lz, float out[3]) {
    int o;
    float fx, fy, fz;
    if (al_corner3(p, lx, ly, lz, o, fx, fy, fz)) {
      al_tri3(sc, o, fx, fy, fz, out);
    } else {
      sample_vec_tp_r<float>(cvr, d, p, out);
    }
  };
#pragma unroll 1
  for (int kz = 0; kz < AL_TZ; ++kz) {
    const int k = z0 + kz;
    if (k >= d.nz) break;
    const int lz = kz + 1;
    const int l0 = lz * AL_VP + ly * AL_VR + lx;  // the cell's staged velocity index
    const TBounds<float> bw2 = tpos_bounds(lo2, hi2, i, j, k), bw1 = tpos_bounds(lo1, hi1, i, j, k);
    float res[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      // face +1/2 along axis a: the RK2 trace of trace_back_tp
      TPos<float> f;
      f.b[0] = i; f.b[1] = j; f.b[2] = k;
      f.o[0] = f.o[1] = f.o[2] = 0.f;
      f.o[a] = 0.5f;
      TPos<float> p0 = f;
      tpos_clamp(p0, bw2);
      bool moved = false;
#pragma unroll
      for (int q = 0; q < 3; ++q) moved |= p0.o[q] != f.o[q];
      float v0[3], vm[3];
      if (!moved) {
        const int sa = a == 0 ? 1 : (a == 1 ? AL_VR : AL_VP);
#pragma unroll
        for (int q = 0; q < 3; ++q)
          v0[q] = mixv(sc[l0 + q * AL_EX], sc[l0 + q * AL_EX + sa], 0.5f);
      } else {
        vel_at(p0, lz, v0);
      }
      TPos<float> m = f;
      for (int q = 0; q < 3; ++q) m.o[q] -= 0.5f * (v0[q] * dth);
      tpos_clamp(m, bw2);
      vel_at(m, lz, vm);
      TPos<float> b = f;
      for (int q = 0; q < 3; ++q) b.o[q] -= vm[q] * dth;
      tpos_clamp(b, bw2);
      b.o[a] -= 0.5f;  // the -1/2 texel shift of sampleMacU/V/W
      tpos_clamp(b, bw1);
      int o;
      float fx, fy, fz;
      float sv;
      if (al_corner(b, lx, ly, lz, o, fx, fy, fz)) sv = al_tri(sf[a], o, fx, fy, fz);
      else sv = sample_tp_r<float>(rf[a], d, b);
      res[a] = sv * damp + trace_nan(v0, vm);
    }
    if constexpr (SCENE)
      advect_scene_terms<float>(d, i, j, k, wp, drops, phs, wave, res[0], res[1], res[2]);
    const size_t c = idx3(d, i, j, k);
    uo[c] = (i >= d.nx - 1) ? 0.f : res[0];
    vo[c] = res[1];
    wo[c] = (k >= d.nz - 1) ? 0.f : res[2];
  }
}

// Neighbour face values with the workload's wall rules: no-slip solid walls
// (tangential ghost = -value, normal wall face = 0) and a free-slip lid
// (ghost = value).  At most one index is out of range per call.
template <class T>
__device__ inline T face_u(const Dims &d, const T *__restrict__ u, int i, int j, int k) {
  if (i < 0 || i >= d.nx) return T(0);
  if (j < 0) return -u[idx3(d, i, 0, k)];
  if (j >= d.ny) return u[idx3(d, i, d.ny - 1, k)];
  if (k < 0) return -u[idx3(d, i, j, 0)];
  if (k >= d.nz) return -u[idx3(d, i, j, d.nz - 1)];
  return u[idx3(d, i, j, k)];
}
template <class T>
__device__ inline T face_v(const Dims &d, const T *__restrict__ v, int i, int j, int k) {
  if (j < 0) return T(0);
  if (j >= d.ny) return v[idx3(d, i, d.ny - 1, k)];
  if (i < 0) return -v[idx3(d, 0, j, k)];
  if (i >= d.nx) return -v[idx3(d, d.nx - 1, j, k)];
  if (k < 0) return -v[idx3(d, i, j, 0)];
  if (k >= d.nz) return -v[idx3(d, i, j, d.nz - 1)];
  return v[idx3(d, i, j, k)];
}
template <class T>
__device__ inline T face_w(const Dims &d, const T *__restrict__ w, int i, int j, int k) {
  if (k < 0 || k >= d.nz) return T(0);
  if (i < 0) return -w[idx3(d, 0, j, k)];
  if (i >= d.nx) return -w[idx3(d, d.nx - 1, j, k)];
  if (j < 0) return -w[idx3(d, i, 0, k)];
  if (j >= d.ny) return w[idx3(d, i, d.ny - 1, k)];
  return w[idx3(d, i, j, k)];
}

// One cell of the Casson update and buoyancy (fluid.ts:451-486 with the
// workload's SI Casson law): m the cell's three faces, x/y/z p/n the face
// values one cell up/down each axis (wall ghosts applied), dc the cell's
// dye, df the dye at its top y-face.
template <class T>
__device__ inline void casson_cell(const StepParams &P, T m0, T m1, T m2, T xp0, T xp1, T xp2,
                                   T xn0, T xn1, T xn2, T yp0, T yp1, T yp2, T yn0, T yn1, T yn2,
                                   T zp0, T zp1, T zp2, T zn0, T zn1, T zn2, T dc, T df, T &o0,
                                   T &o1, T &o2) {
  const T inv2h = T(1.0 / (2.0 * P.dx));
  const T dtoh2 = T(P.dt / (P.dx * P.dx));
  const T buoy = T(-P.g * (P.rho_b - P.rho_w) / P.rho_w * P.dt);
  T ss = T(0);
  {
    T a = xp0 - xn0, b = yp0 - yn0, e = zp0 - zn0;
    ss += a * a + b * b + e * e;
    a = xp1 - xn1; b = yp1 - yn1; e = zp1 - zn1;
    ss += a * a + b * b + e * e;
    a = xp2 - xn2; b = yp2 - yn2; e = zp2 - zn2;
    ss += a * a + b * b + e * e;
  }
  T gam = inv2h * sqrt(ss + T(P.shear_eps2));
  T cc = clampv(dc, T(0), T(1));
  T sq_muinf = sqrt(T(P.mu_w) + cc * T(P.mu_inf - P.mu_w));
  T tau = T(P.tau_y) * cc * cc;
  T reg = T(P.gamma_reg);
  T gr = gam + reg;
  T yterm = sqrt(tau / (gr > reg ? gr : reg));
  T mu_eff = (sq_muinf + yterm) * (sq_muinf + yterm);
  T rho = T(P.rho_w) + cc * T(P.rho_b - P.rho_w);
  T alpha = clampv(mu_eff / rho * dtoh2, T(0), T(P.alpha_max));
  o0 = m0 + alpha * ((xp0 + xn0 + yp0 + yn0 + zp0 + zn0) - T(6) * m0);
  o1 = m1 + alpha * ((xp1 + xn1 + yp1 + yn1 + zp1 + zn1) - T(6) * m1);
  o2 = m2 + alpha * ((xp2 + xn2 + yp2 + yn2 + zp2 + zn2) - T(6) * m2);
  o1 += buoy * df;
}

// One cell of FLUID_CASSON_FORCES parts (1) and (3) with the reference's own
// physics (the coupling's physics flag): the lattice Casson law
//   gamma = 50 sqrt(|du|^2 + |dv|^2 + |dw|^2 + 1e-6)        (fluid.ts:451-454)
//   alpha = clamp((sqrt(1e-4 (1 + 1.35 c)) + sqrt(0.0036 scale c^2 /
//           max(gamma + 0.4, 0.4)))^2, 1e-4, 0.0044)      (fluid.ts:458-462)
// with c = clamp(dc, 0, 1.5) and no dt (a per-step rate), then the
// tanh-saturated buoyancy of the face concentration rho = 1.24 fr + 0.26 fg
// (fr, fg: the intact and lysed concentration at the cell's top y-face, fg 0
// without the 4-channel concentration), gated by the face's height under the
// surface (fluid.ts:477-486); the neighbours come clamped to the edge
// (fluid.ts:387-392).  j: the cell's row.
template <class T>
__device__ inline void casson_cell_ref(const StepParams &P, const Dims &d, int j, T m0, T m1, T m2,
                                       T xp0, T xp1, T xp2, T xn0, T xn1, T xn2, T yp0, T yp1,
                                       T yp2, T yn0, T yn1, T yn2, T zp0, T zp1, T zp2, T zn0,
                                       T zn1, T zn2, T dc, T fr, T fg, T &o0, T &o1, T &o2) {
  T ss = T(0);
  {
    T a = xp0 - xn0, b = yp0 - yn0, e = zp0 - zn0;
    ss += a * a + b * b + e * e;
    a = xp1 - xn1; b = yp1 - yn1; e = zp1 - zn1;
    ss += a * a + b * b + e * e;
    a = xp2 - xn2; b = yp2 - yn2; e = zp2 - zn2;
    ss += a * a + b * b + e * e;
  }
  const T gam = T(50) * sqrt(ss + T(1e-6));
  const T cr = clampv(dc, T(0), T(1.5));
  const T sq_muinf = sqrt(T(0.00010) * (T(1) + T(1.35) * cr));
  const T tau = T(0.0036) * T(P.casson_scale) * cr * cr;
  const T gr = gam + T(0.40);
  const T yterm = sqrt(tau / (gr > T(0.40) ? gr : T(0.40)));
  const T alpha = clampv((sq_muinf + yterm) * (sq_muinf + yterm), T(0.00010), T(0.0044));
  o0 = m0 + alpha * ((xp0 + xn0 + yp0 + yn0 + zp0 + zn0) - T(6) * m0);
  o1 = m1 + alpha * ((xp1 + xn1 + yp1 + yn1 + zp1 + zn1) - T(6) * m1);
  o2 = m2 + alpha * ((xp2 + xn2 + yp2 + yn2 + zp2 + zn2) - T(6) * m2);
  const T rho = T(1.24) * fr + T(0.26) * fg;
  if (rho > T(0.0008)) {
    const T yf = T(P.surface_y - P.depth + (j + 1.0) / d.ny * P.depth);
    const T below = T(P.surface_y) - yf;
    const T gate = smoothstep_v<T>(T(0.002), T(0.024), below > T(0) ? below : T(0));
    const T dtc = T(P.dt < 0.0 ? 0.0 : (P.dt > 0.025 ? 0.025 : P.dt));
    o1 += -T(P.buoyancy) * T(0.345) * (T(1.24) * tanh(rho / T(1.24))) * gate * dtc;
  }
}

// FLUID_CASSON_FORCES (fluid.ts:377-511) parts (1) and (3): Casson
// shear-thinning explicit viscous update and Boussinesq buoyancy.
//   gamma  = |+-1 neighbour differences| / (2h)          (fluid.ts:451-454)
//   mu_eff = (sqrt(mu_inf(c)) + sqrt(tau_y(c) / max(gamma + reg, reg)))^2
//                                                        (fluid.ts:458-462)
//   u     += alpha (sum of 6 neighbours - 6 u), alpha = mu_eff/rho dt/h^2
//                                                        (fluid.ts:464-465)
//   v     += -g (rho_b - rho_w)/rho_w * dye_face * dt    (fluid.ts:477-486)
// with SI Casson parameters blended by the dye fraction c.
//
// Marching in z: a 512-thread block owns a 64 x 8
// column tile and a z-chunk [k0, k1).  Planes of u, v, w with a one-cell
// ring are staged in a four-slot LDS ring (one barrier per plane) with the
// wall ghosts of face_u/v/w already applied: a ring cell outside the grid
// holds its mirror cell's value times the ghost sign (0 for the wall-normal
// face, -1 for the no-slip tangential faces, +1 at the free-slip lid), and
// the planes below 0 and above nz - 1 hold plane 0 / nz - 1 times (-1, -1,
// 0).  So every load is in the grid and the stencil needs no branches; each
// plane is read about 1.3 times instead of about 3 (by one thread per cell).
#define CS_TX 64
#define CS_TY 8
#define CS_T (CS_TX * CS_TY)
#define CS_EX (CS_TX + 2)
#define CS_EY (CS_TY + 2)
#define CS_PL (CS_EX * CS_EY)
#define CS_LPT ((CS_PL + CS_T - 1) / CS_T)

// (FP32: at least 6 waves per SIMD, three blocks per CU, measured 65-69
// against 84 us at 192x256x192; 8 waves forces too few registers and was
// slower.  FP64 needs its registers.)
// REF: the reference's own physics (casson_cell_ref): every ghost is the
// clamped edge value (sign +1), as loadMac clamps its indices.
template <class T, bool REF>
__global__ void __launch_bounds__(CS_T, sizeof(T) == 4 ? 6 : 1)
k_casson_zm(Dims d, StepParams P, const T *__restrict__ u, const T *__restrict__ v,
            const T *__restrict__ w, const T *__restrict__ dye, const T *__restrict__ lys,
            const T *__restrict__ dface, const T *__restrict__ lface, T *__restrict__ uo,
            T *__restrict__ vo, T *__restrict__ wo, int cz, int ncx, int ncy) {
  __shared__ T su[4][CS_EY][CS_EX], sv[4][CS_EY][CS_EX], sw[4][CS_EY][CS_EX];
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  const int lb = xcd_remap(blockIdx.x, gridDim.x);  // one z-chunk of tiles per XCD
  const int x0 = (lb % ncx) * CS_TX, y0 = ((lb / ncx) % ncy) * CS_TY;
  const int k0 = (lb / (ncx * ncy)) * cz, k1 = min(k0 + cz, nz);
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * CS_TX + tx;
  const size_t plane = (size_t)nx * ny;
  // staged positions: mirrored in-plane index and ghost signs per field
  int so[CS_LPT], sr[CS_LPT], sc[CS_LPT];
  T gu[CS_LPT], gv[CS_LPT], gw[CS_LPT];
#pragma unroll
  for (int q = 0; q < CS_LPT; ++q) {
    int vq = tid + q * CS_T;
    bool act = vq < CS_PL;
    if (!act) vq = 0;
    sr[q] = vq / CS_EX;
    sc[q] = vq % CS_EX;
    int gx = x0 - 1 + sc[q], gy = y0 - 1 + sr[q];
    T a = T(1), b = T(1), e = T(1);
    if (!REF) {
      if (gx < 0 || gx >= nx) { a = T(0); b = T(-1); e = T(-1); }
      else if (gy < 0) { a = T(-1); b = T(0); e = T(-1); }
    }
    int mx = gx < 0 ? 0 : (gx >= nx ? nx - 1 : gx), my = gy < 0 ? 0 : (gy >= ny ? ny - 1 : gy);
    so[q] = my * nx + mx;
    gu[q] = a;
    gv[q] = b;
    gw[q] = e;
    if (!act) so[q] = -1;
  }
  T pu[CS_LPT], pv[CS_LPT], pw[CS_LPT];
  auto fetch = [&](int k) {
    int kc = k < 0 ? 0 : (k >= nz ? nz - 1 : k);
    const size_t base = (size_t)kc * plane;
#pragma unroll
    for (int q = 0; q < CS_LPT; ++q) {
      size_t o = base + (so[q] < 0 ? 0 : so[q]);
      pu[q] = u[o];
      pv[q] = v[o];
      pw[q] = w[o];
    }
  };
  auto stage = [&](int k) {
    // ghost planes below and above the grid: (-u, -v, 0) of the edge plane
    const bool zg = !REF && (k < 0 || k >= nz);
    const T zu = zg ? T(-1) : T(1), zw = zg ? T(0) : T(1);
    const int sl = k & 3;
#pragma unroll
    for (int q = 0; q < CS_LPT; ++q)
      if (so[q] >= 0) {
        su[sl][sr[q]][sc[q]] = gu[q] * zu * pu[q];
        sv[sl][sr[q]][sc[q]] = gv[q] * zu * pv[q];
        sw[sl][sr[q]][sc[q]] = gw[q] * zw * pw[q];
      }
  };
  const int gx = x0 + tx, gy = y0 + ty;
  const bool inside = gx < nx && gy < ny;
  const size_t cc0 = inside ? (size_t)gy * nx + gx : 0;
  const size_t up = (gy + 1 < ny) ? (size_t)nx : 0;
  fetch(k0 - 1);
  stage(k0 - 1);
  fetch(k0);
  stage(k0);
  fetch(k0 + 1);
  // the concentration at the cell (dye; the lysed one, lys, with REF and the
  // 4-channel concentration, else 0) and at its top y-face: the face planes
  // when given (samples of a finer concentration grid), else the mean of the
  // cell and the cell above
  const T *fd = dface ? dface : dye, *fl = dface ? lface : lys;
  const size_t fo = dface ? 0 : up;
  T dn = dye[(size_t)k0 * plane + cc0], dnu = fd[(size_t)k0 * plane + cc0 + fo];
  T ln = T(0), lnu = T(0);
  if (REF && lys) {
    ln = lys[(size_t)k0 * plane + cc0];
    lnu = fl[(size_t)k0 * plane + cc0 + fo];
  }
  const int r = ty + 1, c = tx + 1;
  for (int k = k0; k < k1; ++k) {
    stage(k + 1);
    const T dc = dn;
    const T fr = dface ? dnu : T(0.5) * (dn + dnu), fg = dface ? lnu : T(0.5) * (ln + lnu);
    __syncthreads();
    fetch(k + 2);
    if (k + 1 < k1) {
      dn = dye[(size_t)(k + 1) * plane + cc0];
      dnu = fd[(size_t)(k + 1) * plane + cc0 + fo];
      if (REF && lys) {
        ln = lys[(size_t)(k + 1) * plane + cc0];
        lnu = fl[(size_t)(k + 1) * plane + cc0 + fo];
      }
    }
    if (inside) {
      const int s0 = k & 3, sm = (k - 1) & 3, sp = (k + 1) & 3;
      T o0, o1, o2;
      if (REF)
        casson_cell_ref<T>(P, d, gy, su[s0][r][c], sv[s0][r][c], sw[s0][r][c],
                           su[s0][r][c + 1], sv[s0][r][c + 1], sw[s0][r][c + 1],
                           su[s0][r][c - 1], sv[s0][r][c - 1], sw[s0][r][c - 1],
                           su[s0][r + 1][c], sv[s0][r + 1][c], sw[s0][r + 1][c],
                           su[s0][r - 1][c], sv[s0][r - 1][c], sw[s0][r - 1][c],
                           su[sp][r][c], sv[sp][r][c], sw[sp][r][c],
                           su[sm][r][c], sv[sm][r][c], sw[sm][r][c], dc, fr, fg, o0, o1, o2);
      else
      casson_cell<T>(P, su[s0][r][c], sv[s0][r][c], sw[s0][r][c],
                     su[s0][r][c + 1], sv[s0][r][c + 1], sw[s0][r][c + 1],
                     su[s0][r][c - 1], sv[s0][r][c - 1], sw[s0][r][c - 1],
                     su[s0][r + 1][c], sv[s0][r + 1][c], sw[s0][r + 1][c],
                     su[s0][r - 1][c], sv[s0][r - 1][c], sw[s0][r - 1][c],
                     su[sp][r][c], sv[sp][r][c], sw[sp][r][c],
                     su[sm][r][c], sv[sm][r][c], sw[sm][r][c], dc, fr, o0, o1, o2);
      const size_t ci = (size_t)k * plane + cc0;
      uo[ci] = (gx >= nx - 1) ? T(0) : o0;
      vo[ci] = o1;
      wo[ci] = (k >= nz - 1) ? T(0) : o2;
    }
  }
}

// FLUID_MAC_RHS (fluid.ts:517-555): D = (uR - uL) + (vT - vB) + (wF - wB),
// solid wall faces 0, top face free; on the top row the lid pressure moves
// to the right-hand side, D -= p_top (fluid.ts:548-551).  ptop (nx * nz,
// nullable: the flat lid, p_top = 0) holds evalTopDirichletPressure.
template <class T, class R>
__global__ void __launch_bounds__(256)
k_divergence(Dims d, const T *__restrict__ u, const T *__restrict__ v,
             const T *__restrict__ w, R *__restrict__ out, const float *__restrict__ ptop) {
  MAC3D_CELL_PROLOGUE
  typedef typename compute_of<R>::type C;
  C uR = (i >= d.nx - 1) ? C(0) : (C)u[c];
  C uL = (i == 0) ? C(0) : (C)u[c - sx];
  C vT = (C)v[c];
  C vB = (j == 0) ? C(0) : (C)v[c - sy];
  C wF = (k >= d.nz - 1) ? C(0) : (C)w[c];
  C wB = (k == 0) ? C(0) : (C)w[c - sz];
  C rhs = (uR - uL) + (vT - vB) + (wF - wB);
  if (ptop && j == d.ny - 1) rhs = rhs - (C)ptop[(size_t)k * d.nx + i];
  out[c] = (R)rhs;
}

// FLUID_MAC_PROJECT (fluid.ts:863-911): face -= p(+1) - p, solid wall faces
// 0, the top face against the lid's ghost pressure p_top (ptop, nullable:
// the flat lid's 0; fluid.ts:893-898).
// CLAMP2: the reference physics' +-2 clamp of the projected faces (fluid.ts:908).
template <class T, class Pt, bool CLAMP2>
__global__ void __launch_bounds__(256)
k_project(Dims d, T *__restrict__ u, T *__restrict__ v, T *__restrict__ w,
          const Pt *__restrict__ p, const float *__restrict__ ptop) {
  MAC3D_CELL_PROLOGUE
  typedef typename compute_of<Pt>::type C;
  C pc = (C)p[c];
  C uu = (i < d.nx - 1) ? (C)u[c] - ((C)p[c + sx] - pc) : C(0);
  C pu = (j < d.ny - 1) ? (C)p[c + sy] : (ptop ? (C)ptop[(size_t)k * d.nx + i] : C(0));
  C vv = (C)v[c] - (pu - pc);
  C ww = (k < d.nz - 1) ? (C)w[c] - ((C)p[c + sz] - pc) : C(0);
  if (CLAMP2) {
    uu = clampv(uu, C(-2), C(2));
    vv = clampv(vv, C(-2), C(2));
    ww = clampv(ww, C(-2), C(2));
  }
  u[c] = (T)uu;
  v[c] = (T)vv;
  w[c] = (T)ww;
}

// CONC_PREDICT (advect.ts:28-72): RK2 semi-Lagrangian predictor of the dye
// on the centred velocity, normalised coordinates clamped to [0.001, 0.999].
template <class T>
__device__ inline void dye_back(const Dims &d, const typename cvec_of<T>::type *__restrict__ cvel,
                                const T sc[3], const T uvw[3], T v0[3], T back[3]) {
  const T lo = T(0.001), hi = T(0.999);
  sample_vec<T>(cvel, d, uvw[0], uvw[1], uvw[2], v0);
  T mx = clampv(uvw[0] - T(0.5) * (v0[0] * sc[0]), lo, hi);
  T my = clampv(uvw[1] - T(0.5) * (v0[1] * sc[1]), lo, hi);
  T mz = clampv(uvw[2] - T(0.5) * (v0[2] * sc[2]), lo, hi);
  T vm[3];
  sample_vec<T>(cvel, d, mx, my, mz, vm);
  back[0] = clampv(uvw[0] - vm[0] * sc[0], lo, hi);
  back[1] = clampv(uvw[1] - vm[1] * sc[1], lo, hi);
  back[2] = clampv(uvw[2] - vm[2] * sc[2], lo, hi);
}

// dye_back in texel space from the cell centre (no clamp on the first
// sample), clamps at [0.001 N - 1/2, 0.999 N - 1/2]
template <class T, class CR>
__device__ inline void dye_back_tp(const Dims &d, const CR &cvel, T dth, int i, int j, int k,
                                   const T lo[3], const T hi[3], T v0[3], TPos<T> &back) {
  TPos<T> c;
  c.b[0] = i; c.b[1] = j; c.b[2] = k;
  c.o[0] = c.o[1] = c.o[2] = T(0);
  // the first sample sits on the texel centre (weights 0: mixv(x, y, 0) = x
  // exactly), so it is the texel's own value
  {
    auto a0 = cvel.ld(cvel.cell(d, i, j, k));
    v0[0] = a0.x;
    v0[1] = a0.y;
    v0[2] = a0.z;
  }
  const TBounds<T> bw = tpos_bounds(lo, hi, i, j, k);
  TPos<T> m = c;
  for (int a = 0; a < 3; ++a) m.o[a] -= T(0.5) * (v0[a] * dth);
  tpos_clamp(m, bw);
  T vm[3];
  sample_vec_tp_r<T>(cvel, d, m, vm);
  back = c;
  for (int a = 0; a < 3; ++a) back.o[a] -= vm[a] * dth;
  tpos_clamp(back, bw);
}

template <class T>
__device__ inline void dye_bounds(const Dims &d, T lo[3], T hi[3]) {
  const T N[3] = {(T)d.nx, (T)d.ny, (T)d.nz};
  for (int a = 0; a < 3; ++a) {
    lo[a] = T(0.001) * N[a] - T(0.5);
    hi[a] = T(0.999) * N[a] - T(0.5);
  }
}

// CONC_PREDICT (advect.ts:28-72): the dye sampled at the backtrace point.  mm
// (nullable) takes the least and greatest of the 8 dye texels of that sample
// per cell, which are the bounds CONC_RESOLVE's limiter clamps to, so that
// k_dye_resolve_mm need not trace back again.
// the cells CONC_PREDICT and CONC_RESOLVE hold at zero: the lateral walls and
// the top row (advect.ts:42-45, 100-103)
__device__ inline bool conc_wall(const Dims &d, int i, int j, int k) {
  return i == 0 || i >= d.nx - 1 || k == 0 || k >= d.nz - 1 || j >= d.ny - 1;
}

// REF: the reference's predictor branches (the coupling's physics flag): 0 at
// conc_wall cells; the cell's own dye below speed^2 2.5e-5 (advect.ts:54-58);
// a one-stage trace t - v0 dt/h below 0.012 (advect.ts:60-62); else RK2.
// (mm is null with REF: the resolve traces again.)
template <class T, bool BUF, bool REF>
__global__ void __launch_bounds__(256)
k_dye_predict(Dims d, T dth, const typename cvec_of<T>::type *__restrict__ cvel,
              const T *__restrict__ dye, T *__restrict__ pred,
              typename vec2_of<T>::type *__restrict__ mm) {
  MAC3D_CELL_PROLOGUE
  const size_t n = d.cells();
  const auto cv = make_reader<BUF>(cvel, n);
  const auto rd = make_reader<BUF>(dye, n);
  T lo[3], hi[3], v0[3];
  dye_bounds(d, lo, hi);
  if (REF) {
    if (conc_wall(d, i, j, k)) {
      pred[c] = T(0);
      return;
    }
    const auto a0 = cv.ld(cv.cell(d, i, j, k));
    const T sp2 = a0.x * a0.x + a0.y * a0.y + a0.z * a0.z;
    if (sp2 < T(2.5e-5)) {
      pred[c] = dye[c];
      return;
    }
    if (sp2 < T(0.012)) {
      TPos<T> b1;
      b1.b[0] = i; b1.b[1] = j; b1.b[2] = k;
      b1.o[0] = -a0.x * dth;
      b1.o[1] = -a0.y * dth;
      b1.o[2] = -a0.z * dth;
      tpos_clamp(b1, tpos_bounds(lo, hi, i, j, k));
      pred[c] = sample_tp_r<T>(rd, d, b1);
      return;
    }
  }
  TPos<T> back;
  dye_back_tp<T>(d, cv, dth, i, j, k, lo, hi, v0, back);
  if (mm) {
    typename vec2_of<T>::type b;
    pred[c] = sample_tp_mm_r<T>(rd, d, back, b.x, b.y);
    mm[c] = b;
  } else {
    pred[c] = sample_tp_r<T>(rd, d, back);
  }
}

// k_dye_predict for the plain FP32 step (no reference branches; the limiter
// bounds emitted) with its samples taken from LDS: a block stages its tile
// (AL_TX x AL_TY x AL_TZ cells) with a one-cell ring of the centre velocity
// (components interleaved by row) and of the dye, as k_advect_mac_lds does,
// the ring's cells outside the grid replicating the edge cells (the
// sampler's clamp-to-edge).  A sample whose corners leave the box gathers
// from memory.  Weights, lerp order and the min/max order are those of
// k_dye_predict, so the values are identical.
__device__ inline float al_tri_mm(const float *f, int o, float fx, float fy, float fz,
                                  float &mn, float &mx) {
  constexpr int dy = AL_EX, dz = AL_EX * AL_EY;
  const float s0 = f[o], s1 = f[o + 1], s2 = f[o + dy], s3 = f[o + dy + 1];
  const float s4 = f[o + dz], s5 = f[o + dz + 1], s6 = f[o + dz + dy], s7 = f[o + dz + dy + 1];
  mn = fmin(fmin(fmin(s0, s1), fmin(s2, s3)), fmin(fmin(s4, s5), fmin(s6, s7)));
  mx = fmax(fmax(fmax(s0, s1), fmax(s2, s3)), fmax(fmax(s4, s5), fmax(s6, s7)));
  const float x00 = mixv(s0, s1, fx);
  const float x10 = mixv(s2, s3, fx);
  const float x01 = mixv(s4, s5, fx);
  const float x11 = mixv(s6, s7, fx);
  return mixv(mixv(x00, x10, fy), mixv(x01, x11, fy), fz);
}

__global__ void __launch_bounds__(AL_TX * AL_TY)
k_dye_predict_lds(Dims d, float dth, const float3 *__restrict__ cvel,
                  const float *__restrict__ dye, float *__restrict__ pred,
                  float2 *__restrict__ mm) {
  __shared__ float sc[3 * AL_N];  // centre velocity, components interleaved by row
  __shared__ float sd[AL_N];      // dye
  const int x0 = blockIdx.x * AL_TX, y0 = blockIdx.y * AL_TY, z0 = blockIdx.z * AL_TZ;
  const int tid = threadIdx.x;
  // (staged as in k_advect_mac_lds: a thread's two columns load together)
  const size_t plane = (size_t)d.nx * d.ny;
  {
    int cl[2];
    float3 cvv[2][AL_EZ];
    float fd[2][AL_EZ];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int col = min(tid + h * AL_TX * AL_TY, AL_EX * AL_EY - 1);
      cl[h] = col;
      const size_t cxy = (size_t)clampi(y0 - 1 + col / AL_EX, 0, d.ny - 1) * d.nx +
                         clampi(x0 - 1 + col % AL_EX, 0, d.nx - 1);
#pragma unroll
      for (int lz = 0; lz < AL_EZ; ++lz) {
        const size_t c = cxy + (size_t)clampi(z0 - 1 + lz, 0, d.nz - 1) * plane;
        cvv[h][lz] = cvel[c];
        fd[h][lz] = dye[c];
      }
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int col = cl[h], lx = col % AL_EX, ly = col / AL_EX;
#pragma unroll
      for (int lz = 0; lz < AL_EZ; ++lz) {
        const int tv = lz * AL_VP + ly * AL_VR + lx;
        sc[tv] = cvv[h][lz].x;
        sc[tv + AL_EX] = cvv[h][lz].y;
        sc[tv + 2 * AL_EX] = cvv[h][lz].z;
        sd[lz * (AL_EX * AL_EY) + col] = fd[h][lz];
      }
    }
  }
  __syncthreads();
  const int i = x0 + tid % AL_TX, j = y0 + tid / AL_TX;
  if (i >= d.nx || j >= d.ny) return;
  const size_t n = d.cells();
  const auto cvr = make_reader<true>(cvel, n);
  const auto rd = make_reader<true>(dye, n);
  float lo[3], hi[3];
  dye_bounds(d, lo, hi);
  const int lx = tid % AL_TX + 1, ly = tid / AL_TX + 1;
#pragma unroll 1
  for (int kz = 0; kz < AL_TZ; ++kz) {
    const int k = z0 + kz;
    if (k >= d.nz) break;
    const int lz = kz + 1;
    const int l0 = lz * AL_VP + ly * AL_VR + lx;
    // dye_back_tp: the first sample is the cell's own centre velocity
    float v0[3], vm[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) v0[q] = sc[l0 + q * AL_EX];
    const TBounds<float> bw = tpos_bounds(lo, hi, i, j, k);
    TPos<float> m;
    m.b[0] = i; m.b[1] = j; m.b[2] = k;
    m.o[0] = m.o[1] = m.o[2] = 0.f;
    for (int a = 0; a < 3; ++a) m.o[a] -= 0.5f * (v0[a] * dth);
    tpos_clamp(m, bw);
    int o;
    float fx, fy, fz;
    if (al_corner3(m, lx, ly, lz, o, fx, fy, fz)) al_tri3(sc, o, fx, fy, fz, vm);
    else sample_vec_tp_r<float>(cvr, d, m, vm);
    TPos<float> back;
    back.b[0] = i; back.b[1] = j; back.b[2] = k;
    back.o[0] = back.o[1] = back.o[2] = 0.f;
    for (int a = 0; a < 3; ++a) back.o[a] -= vm[a] * dth;
    tpos_clamp(back, bw);
    float2 b;
    float s;
    if (al_corner(back, lx, ly, lz, o, fx, fy, fz)) s = al_tri_mm(sd, o, fx, fy, fz, b.x, b.y);
    else s = sample_tp_mm_r<float>(rd, d, back, b.x, b.y);
    const size_t c = idx3(d, i, j, k);
    pred[c] = s;
    mm[c] = b;
  }
}

// CONC_RESOLVE's MacCormack core (advect.ts:196-222): forward estimate from
// the predictor at uvw + v0 dt, corrected value pred + s (orig - estimate),
// clamped to the min/max of the 8 source texels around the backtrace point.
// REF: conc_wall cells resolve to 0 (advect.ts:100-103).
template <class T, bool BUF, bool REF>
__global__ void __launch_bounds__(256)
k_dye_resolve(Dims d, T dth, T strength, const typename cvec_of<T>::type *__restrict__ cvel,
              const T *__restrict__ dye, const T *__restrict__ pred, T *__restrict__ out) {
  MAC3D_CELL_PROLOGUE
  if (REF && conc_wall(d, i, j, k)) {
    out[c] = T(0);
    return;
  }
  const size_t n = d.cells();
  T lo[3], hi[3], v0[3];
  dye_bounds(d, lo, hi);
  TPos<T> back;
  dye_back_tp<T>(d, make_reader<BUF>(cvel, n), dth, i, j, k, lo, hi, v0, back);
  TPos<T> fw;
  fw.b[0] = i; fw.b[1] = j; fw.b[2] = k;
  for (int a = 0; a < 3; ++a) fw.o[a] = v0[a] * dth;
  tpos_clamp(fw, tpos_bounds(lo, hi, i, j, k));
  T est = sample_tp_r<T>(make_reader<BUF>(pred, n), d, fw);
  T corr = pred[c] + strength * (dye[c] - est);
  int bi = back.b[0] + (int)floor(back.o[0]);
  int bj = back.b[1] + (int)floor(back.o[1]);
  int bk = back.b[2] + (int)floor(back.o[2]);
  int i0 = clampi(bi, 0, d.nx - 1), i1 = clampi(bi + 1, 0, d.nx - 1);
  int j0 = clampi(bj, 0, d.ny - 1), j1 = clampi(bj + 1, 0, d.ny - 1);
  int k0 = clampi(bk, 0, d.nz - 1), k1 = clampi(bk + 1, 0, d.nz - 1);
  T s0 = dye[idx3(d, i0, j0, k0)], s1 = dye[idx3(d, i1, j0, k0)];
  T s2 = dye[idx3(d, i0, j1, k0)], s3 = dye[idx3(d, i1, j1, k0)];
  T s4 = dye[idx3(d, i0, j0, k1)], s5 = dye[idx3(d, i1, j0, k1)];
  T s6 = dye[idx3(d, i0, j1, k1)], s7 = dye[idx3(d, i1, j1, k1)];
  T mn = fmin(fmin(fmin(s0, s1), fmin(s2, s3)), fmin(fmin(s4, s5), fmin(s6, s7)));
  T mx = fmax(fmax(fmax(s0, s1), fmax(s2, s3)), fmax(fmax(s4, s5), fmax(s6, s7)));
  out[c] = clampv(corr, mn, mx);
}

// k_dye_resolve with the limiter's bounds from k_dye_predict (mm, formed from
// the same dye and centre velocity): the same value with no second backtrace.
template <class T, bool BUF>
__global__ void __launch_bounds__(256)
k_dye_resolve_mm(Dims d, T dth, T strength, const typename cvec_of<T>::type *__restrict__ cvel,
                 const T *__restrict__ dye, const T *__restrict__ pred,
                 const typename vec2_of<T>::type *__restrict__ mm, T *__restrict__ out) {
  MAC3D_CELL_PROLOGUE
  T lo[3], hi[3];
  dye_bounds(d, lo, hi);
  const auto a0 = cvel[c];
  TPos<T> fw;
  fw.b[0] = i; fw.b[1] = j; fw.b[2] = k;
  fw.o[0] = a0.x * dth;
  fw.o[1] = a0.y * dth;
  fw.o[2] = a0.z * dth;
  tpos_clamp(fw, tpos_bounds(lo, hi, i, j, k));
  T est = sample_tp_r<T>(make_reader<BUF>(pred, d.cells()), d, fw);
  T corr = pred[c] + strength * (dye[c] - est);
  const auto b = mm[c];
  out[c] = clampv(corr, b.x, b.y);
}

// Seeded blob initial condition: dye at cell centres and -w0 on V faces,
// each the sum over blobs of 1 - smoothstep(r - 1/2, r + 1/2, distance).
struct BlobDev {
  double cx, cy, cz, radius, w0;
};


template <class T>
__global__ void __launch_bounds__(256)
k_init_blobs(Dims d, const BlobDev *__restrict__ blobs, int nb, T *__restrict__ dye,
             T *__restrict__ v) {
  MAC3D_CELL_PROLOGUE
  double dy = 0.0, vv = 0.0;
  for (int q = 0; q < nb; ++q) {
    BlobDev b = blobs[q];
    double ex = i + 0.5 - b.cx, ey = j + 0.5 - b.cy, ez = k + 0.5 - b.cz;
    double r = sqrt(ex * ex + ey * ey + ez * ez);
    dy += 1.0 - smoothstep_v(b.radius - 0.5, b.radius + 0.5, r);
    double fy = j + 1.0 - b.cy;
    double rf = sqrt(ex * ex + fy * fy + ez * ez);
    vv -= b.w0 * (1.0 - smoothstep_v(b.radius - 0.5, b.radius + 0.5, rf));
  }
  dye[c] = (T)dy;
  v[c] = (T)vv;
}

// Divergence norms: per-block max |D| and sum D^2 in FP64, then one block
// folds the partials in block order (deterministic).
template <class T>
__global__ void __launch_bounds__(256)
k_div_partials(Dims d, const T *__restrict__ u, const T *__restrict__ v,
               const T *__restrict__ w, double *__restrict__ part) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int k = blockIdx.z;
  double mx = 0.0, ss = 0.0;
  if (i < d.nx && j < d.ny) {
    size_t c = idx3(d, i, j, k);
    const size_t sy = (size_t)d.nx, sz = (size_t)d.nx * d.ny;
    double uR = (i >= d.nx - 1) ? 0.0 : (double)u[c];
    double uL = (i == 0) ? 0.0 : (double)u[c - 1];
    double vT = (double)v[c];
    double vB = (j == 0) ? 0.0 : (double)v[c - sy];
    double wF = (k >= d.nz - 1) ? 0.0 : (double)w[c];
    double wB = (k == 0) ? 0.0 : (double)w[c - sz];
    double D = (uR - uL) + (vT - vB) + (wF - wB);
    mx = fabs(D);
    ss = D * D;
  }
  mx = wave_max(mx);
  ss = wave_sum(ss);
  __shared__ double smx[4], sss[4];
  int t = threadIdx.y * blockDim.x + threadIdx.x;
  if ((t & 63) == 0) { smx[t >> 6] = mx; sss[t >> 6] = ss; }
  __syncthreads();
  if (t == 0) {
    double a = smx[0], b = sss[0];
    for (int q = 1; q < 4; ++q) { a = smx[q] > a ? smx[q] : a; b += sss[q]; }
    size_t blk = ((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
    part[2 * blk] = a;
    part[2 * blk + 1] = b;
  }
}

// Fold n (max, sum) pairs into out[0] = max, out[1] = sum; one block of 256.
__global__ void __launch_bounds__(256)
k_fold_partials(const double *__restrict__ part, size_t n, double *__restrict__ out) {
  double mx = 0.0, ss = 0.0;
  for (size_t q = threadIdx.x; q < n; q += 256) {
    double a = part[2 * q];
    mx = a > mx ? a : mx;
    ss += part[2 * q + 1];
  }
  mx = wave_max(mx);
  ss = wave_sum(ss);
  __shared__ double smx[4], sss[4];
  int t = threadIdx.x;
  if ((t & 63) == 0) { smx[t >> 6] = mx; sss[t >> 6] = ss; }
  __syncthreads();
  if (t == 0) {
    double a = smx[0], b = sss[0];
    for (int q = 1; q < 4; ++q) { a = smx[q] > a ? smx[q] : a; b += sss[q]; }
    out[0] = a;
    out[1] = b;
  }
}

// k_fold_partials for the context-free divergence norms: out[0] = max,
// out[1] = sqrt(sum / ncells), the RMS (same summation order)
__global__ void __launch_bounds__(256)
k_fold_div_norms(const double *__restrict__ part, size_t n, double ncells,
                 double *__restrict__ out) {
  double mx = 0.0, ss = 0.0;
  for (size_t q = threadIdx.x; q < n; q += 256) {
    double a = part[2 * q];
    mx = a > mx ? a : mx;
    ss += part[2 * q + 1];
  }
  mx = wave_max(mx);
  ss = wave_sum(ss);
  __shared__ double smx[4], sss[4];
  int t = threadIdx.x;
  if ((t & 63) == 0) { smx[t >> 6] = mx; sss[t >> 6] = ss; }
  __syncthreads();
  if (t == 0) {
    double a = smx[0], b = sss[0];
    for (int q = 1; q < 4; ++q) { a = smx[q] > a ? smx[q] : a; b += sss[q]; }
    out[0] = a;
    out[1] = sqrt(b / ncells);
  }
}

// dtype conversion between a staging buffer and a field
template <class A, class B>
__global__ void __launch_bounds__(256) k_convert(const A *__restrict__ src, B *__restrict__ dst, size_t n) {
  size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t < n) dst[t] = (B)src[t];
}

template <class T>
__global__ void __launch_bounds__(256) k_fill(T *__restrict__ dst, size_t n, T val) {
  size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t < n) dst[t] = val;
}

}  // namespace mac3d

namespace mac3d {
// one component of the packed centre velocity <-> a plain array
template <class A, class B>
__global__ void __launch_bounds__(256)
k_convert_strided(const A *__restrict__ src, B *__restrict__ dst, size_t n, int sstride,
                  int soff, int dstride, int doff) {
  size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t < n) dst[t * dstride + doff] = (B)src[t * sstride + soff];
}
}  // namespace mac3d
