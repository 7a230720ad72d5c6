// Lagrangian blood parcels: advection through the centre velocity, stretch,
// settling into the floor bed, osmotic hemolysis, and the fixed-point trilinear
// splat of their mass into two 32-bit counters per cell (PARTICLE_ADVECT_SPLAT,
// src/shaders/particles.ts:189-463).  One thread per parcel; the parcels are
// FP32 as the reference stores them (RbcParcel, particles.ts:16-20).
//
// The splat adds integer weights with atomics, so the counters do not depend
// on the order in which parcels land.  Device-scope atomics bound the stage;
// two x-neighbour cells in one aligned 64-bit word take one 64-bit atomic,
// which equals two 32-bit adds while no counter reaches 2^32 (65,536 units of
// mass in one cell).  Past that the low counter would carry into its x + 1
// neighbour, where the reference's 32-bit atomics wrap; mac3d_parcels_upload
// keeps the parcels' total mass below 65,536 units, so no counter gets there
// (the reference's 393,216 parcels of 0.134 hold 52,691).  Each weight is
// round(mass 65536 w) with w a product of trilinear factors, round-half-even
// as WGSL's round.
#pragma once

#include "common.hpp"
#include "wave_kernels.hpp"

namespace mac3d {

// RbcParcel: pm = (world position, intact mass; 0 inactive),
// sl = (stretch vector, age ticks + lysis fraction)
struct Parcel {
  float4 pm, sl;
};

constexpr float kSplatScale = 65536.0f;      // SPLAT_FP_SCALE (common.ts:20)
constexpr float kParcelMass0 = 0.134f;       // INITIAL_PARCEL_VOXEL_MASS (particles.ts:14)

struct ParcelScene {
  float dt;        // clamp(dt, 0, 0.025)
  float half_xz;   // domainBounds.x
  float y_min;     // domainBounds.y = surface - depth
  float y_max;     // domainBounds.z = surface height
  float sim_time;  // cameraPosAndTime.w
  float k_lysis;   // lensAndRheology.w
};

__device__ inline float pc_ss(float e0, float e1, float x) {
  float t = clampv((x - e0) / (e1 - e0), 0.f, 1.f);
  return t * t * (3.f - 2.f * t);
}

// worldToUVW (common.ts:56-58) over the box [-hx, hx] x [y_min, y_max] x [-hx, hx]
__device__ inline float3 pc_uvw(const ParcelScene &S, float3 p) {
  const float w = 2.f * S.half_xz, h = S.y_max - S.y_min;
  return make_float3((p.x + S.half_xz) / w, (p.y - S.y_min) / h, (p.z + S.half_xz) / w);
}

__device__ inline float3 pc_clamp_uvw(float3 u) {
  return make_float3(clampv(u.x, 0.002f, 0.998f), clampv(u.y, 0.002f, 0.998f),
                     clampv(u.z, 0.002f, 0.998f));
}

// sampleVelWorld (particles.ts:201-236): the centre velocity at p plus Stokes
// settling (slower under the surface and for held meniscus tails) and the
// sub-grid solenoidal unfurling deeper than 3.5 cm
__device__ inline float3 pc_velocity(const ParcelScene &S, const Dims &d,
                                     const float3 *__restrict__ cvel, float3 p, float mature,
                                     float hold) {
  const float below = S.y_max - p.y;
  const float3 u = pc_clamp_uvw(pc_uvw(S, p));
  const float settle_gate = mixv(0.38f, 1.f, pc_ss(0.002f, 0.055f, below));
  const float mono = mixv(0.020f, 0.046f, pc_ss(0.04f, 0.42f, below)) * settle_gate;
  const float stack = mixv(0.032f, 0.062f, pc_ss(0.04f, 0.48f, below));
  const float tail = hold * (1.f - pc_ss(0.003f, 0.034f, fmaxf(0.f, below)));
  const float settle = mixv(mono, stack, mature) * mixv(1.f, 0.035f, tail);
  float3 sub = make_float3(0.f, 0.f, 0.f);
  const float deep = pc_ss(0.035f, 0.24f, below);
  if (deep > 0.001f) {
    const float tp = S.sim_time * 0.42f;
    const float a1 = 8.4f * p.x + 5.2f * p.y + 6.6f * p.z + tp;
    const float a2 = -7.2f * p.x + 6.0f * p.y + 9.4f * p.z - tp * 0.81f;
    const float a3 = 11.5f * p.x - 7.5f * p.y - 8.5f * p.z + tp * 1.13f;
    const float c1 = cosf(a1), c2 = cosf(a2), s3 = 0.45f * sinf(a3);
    const float k = 0.00028f * deep;
    sub.x = (-8.16f * c1 - 8.20f * c2 + 8.30f * s3) * k;
    sub.y = (3.84f * c1 - 1.38f * c2 - 1.10f * s3) * pc_ss(0.06f, 0.24f, below) * k;
    sub.z = (7.36f * c1 - 5.40f * c2 + 12.20f * s3) * k;
  }
  float v[3];
  sample_vec<float>(cvel, d, u.x, u.y, u.z, v);
  v[1] *= mixv(1.f, 0.045f, tail);
  return make_float3(v[0] + sub.x, v[1] - settle + sub.y, v[2] + sub.z);
}

// splatTrilinearIntact/Lysed (particles.ts:238-301): integer trilinear weights
// round(mass 65536 w) added to the eight cells around uvw (clamped to the grid)
__device__ inline void pc_splat(const Dims &d, unsigned *__restrict__ acc, float3 u, float mass) {
  const float cx = u.x * d.nx - 0.5f, cy = u.y * d.ny - 0.5f, cz = u.z * d.nz - 0.5f;
  const float bx = floorf(cx), by = floorf(cy), bz = floorf(cz);
  const float fx = clampv(cx - bx, 0.f, 1.f), fy = clampv(cy - by, 0.f, 1.f),
              fz = clampv(cz - bz, 0.f, 1.f);
  const float gx = 1.f - fx, gy = 1.f - fy, gz = 1.f - fz;
  const int ix = (int)bx, iy = (int)by, iz = (int)bz;
  const int x0 = clampi(ix, 0, d.nx - 1), x1 = clampi(ix + 1, 0, d.nx - 1);
  const int y0 = clampi(iy, 0, d.ny - 1), y1 = clampi(iy + 1, 0, d.ny - 1);
  const int z0 = clampi(iz, 0, d.nz - 1), z1 = clampi(iz + 1, 0, d.nz - 1);
  const float s = mass * kSplatScale;
  const float wy[2] = {gy, fy}, wz[2] = {gz, fz};
  const int ys[2] = {y0, y1}, zs[2] = {z0, z1};
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int b = c & 1, e = c >> 1;
    // ((s wx) wy) wz, each product rounded (no fused multiply-add)
    const unsigned q0 = (unsigned)rintf(__fmul_rn(__fmul_rn(__fmul_rn(s, gx), wy[b]), wz[e]));
    const unsigned q1 = (unsigned)rintf(__fmul_rn(__fmul_rn(__fmul_rn(s, fx), wy[b]), wz[e]));
    unsigned *row = acc + idx3(d, 0, ys[b], zs[e]);
    if (x1 == x0 + 1 && !(x0 & 1)) {
      // both cells in one aligned 64-bit word (rows start even: nx is even):
      // one atomic adds both counts (no carry while a counter stays below 2^32)
      if (q0 | q1)
        atomicAdd((unsigned long long *)(row + x0),
                  ((unsigned long long)q1 << 32) | (unsigned long long)q0);
    } else if (x1 == x0) {  // clamped at a wall: both weights on one cell
      if (q0 + q1) atomicAdd(row + x0, q0 + q1);
    } else {
      if (q0) atomicAdd(row + x0, q0);
      if (q1) atomicAdd(row + x1, q1);
    }
  }
}

// One parcel step (PARTICLE_ADVECT_SPLAT main, particles.ts:303-462).
// cvel: the centre velocity of the MAC grid d.  ds: the splat grid (the
// concentration's, CONC_GRID, while it is on; else d), which holds the two
// counters intact/lysed and the hemolysis exposure: dye, standing for the
// concentration's intact channel, plus lysed_conc (nullable).
__global__ void __launch_bounds__(256)
k_parcel_advect_splat(Dims d, Dims ds, ParcelScene S, int np, Parcel *__restrict__ parcels,
                      const float3 *__restrict__ cvel, const float *__restrict__ dye,
                      const float *__restrict__ lysed_conc, unsigned *__restrict__ intact,
                      unsigned *__restrict__ lysed) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= np) return;
  Parcel P = parcels[g];
  float mass = P.pm.w;
  if (mass <= 1e-6f) return;
  const float dt = S.dt, hx = S.half_xz, ymin = S.y_min;
  float3 pos = make_float3(P.pm.x, P.pm.y, P.pm.z);
  float3 st = make_float3(P.sl.x, P.sl.y, P.sl.z);
  const float raw = fmaxf(0.f, P.sl.w);
  float ticks = floorf(raw);
  float lysis = clampv(raw - ticks, 0.f, 0.98f);
  const float age = ticks * (1.f / 240.f);
  const float mature = pc_ss(8.5f, 18.f, age);
  float tail_rank;
  {  // each product and sum rounded on its own (no contraction into an FMA)
#pragma clang fp contract(off)
    const float tr = (float)g * 0.7548776662f;
    tail_rank = tr - floorf(tr);
  }
  const float hold = powf(pc_ss(0.36f, 0.995f, tail_rank), 1.65f) * (1.f - pc_ss(5.2f, 10.2f, age));

  const float3 v1 = pc_velocity(S, d, cvel, pos, mature, hold);
  const float sp2 = v1.x * v1.x + v1.y * v1.y + v1.z * v1.z;
  if (dt > 0.f) {
    ticks = fminf(8191.f, ticks + fmaxf(1.f, rintf(dt * 240.f)));
    const float3 v2 = pc_velocity(
        S, d, cvel, make_float3(pos.x + 0.5f * v1.x * dt, pos.y + 0.5f * v1.y * dt, pos.z + 0.5f * v1.z * dt),
        mature, hold);
    float3 va = v2;
    if (sp2 > 0.0025f) {  // Ralston RK3
      const float3 v3 = pc_velocity(
          S, d, cvel,
          make_float3(pos.x + 0.75f * v2.x * dt, pos.y + 0.75f * v2.y * dt, pos.z + 0.75f * v2.z * dt),
          mature, hold);
      va = make_float3((2.f / 9.f) * v1.x + (3.f / 9.f) * v2.x + (4.f / 9.f) * v3.x,
                       (2.f / 9.f) * v1.y + (3.f / 9.f) * v2.y + (4.f / 9.f) * v3.y,
                       (2.f / 9.f) * v1.z + (3.f / 9.f) * v2.z + (4.f / 9.f) * v3.z);
    }
    // floor contact: damping and an outward glide (particles.ts:341-349)
    const float slip = pc_ss(ymin + 0.005f, ymin + 0.042f, pos.y);
    const float rxz = sqrtf(pos.x * pos.x + pos.z * pos.z);
    float ox = 0.f, oz = 0.f;
    if (rxz > 1e-4f) {
      const float k = 0.016f * (1.f - slip) * pc_ss(0.01f, 0.22f, rxz) / fmaxf(rxz, 1e-4f);
      ox = pos.x * k;
      oz = pos.z * k;
    }
    va = make_float3(va.x * mixv(0.45f, 1.f, slip) + ox, va.y * mixv(0.15f, 1.f, slip),
                     va.z * mixv(0.45f, 1.f, slip) + oz);
    pos = make_float3(pos.x + va.x * dt, pos.y + va.y * dt, pos.z + va.z * dt);
    pos.y = clampv(pos.y, ymin + 0.0055f, S.y_max - 0.0018f);
    // material line element ds/dt = (s . grad) u (particles.ts:353-374)
    if (sp2 > 2.5e-7f && slip > 0.15f) {
      const float sl = fmaxf(sqrtf(st.x * st.x + st.y * st.y + st.z * st.z), 1e-5f);
      const float3 sd = make_float3(st.x / sl, st.y / sl, st.z / sl);
      const float pr = 0.0065f;
      const float3 vp = pc_velocity(S, d, cvel, make_float3(pos.x + sd.x * pr, pos.y + sd.y * pr, pos.z + sd.z * pr), mature, hold);
      const float3 vm = pc_velocity(S, d, cvel, make_float3(pos.x - sd.x * pr, pos.y - sd.y * pr, pos.z - sd.z * pr), mature, hold);
      const float k = sl / (2.f * pr);
      st = make_float3(st.x + (vp.x - vm.x) * k * dt, st.y + (vp.y - vm.y) * k * dt,
                       st.z + (vp.z - vm.z) * k * dt);
      const float vmag = sqrtf(sp2);
      if (vmag > 0.001f && sl < 0.006f) {
        const float a = 0.0022f * dt / vmag;
        st = make_float3(st.x + v1.x * a, st.y + v1.y * a, st.z + v1.z * a);
      }
      const float maxlen =
          mixv(0.0270f, 0.0420f, hold * (1.f - pc_ss(0.005f, 0.065f, fmaxf(0.f, S.y_max - pos.y))));
      const float len = sqrtf(st.x * st.x + st.y * st.y + st.z * st.z);
      const float k2 = clampv(len, 0.0026f, maxlen) / fmaxf(len, 1e-5f);
      st = make_float3(st.x * k2, st.y * k2, st.z * k2);
    } else if (slip <= 0.15f) {
      st.y *= expf(-dt * 5.f);
    }
  }

  const float3 u = pc_uvw(S, pos);
  const float rh = sqrtf(pos.x * pos.x + pos.z * pos.z);
  const float sponge = 1.f - pc_ss(hx * 0.68f, hx * 0.94f, rh);
  mass *= mixv(1.f, sponge, clampv(dt * 6.f, 0.f, 1.f));
  if (mass <= 1e-5f || rh >= hx * 0.95f || u.x <= 0.002f || u.x >= 0.998f || u.z <= 0.002f ||
      u.z >= 0.998f || u.y >= 0.998f) {
    P.pm.w = 0.f;  // retired where it stood at the start of the step
    parcels[g] = P;
    return;
  }
  const float3 uc = pc_clamp_uvw(u);
  // settling into the floor bed: lysed counter (particles.ts:386-398)
  if (dt > 0.f && pos.y < ymin + 0.026f) {
    const float w = 1.f - pc_ss(ymin + 0.005f, ymin + 0.026f, pos.y);
    const float dm = mass * clampv(w * dt * 1.65f, 0.f, 0.095f);
    if (dm > 1e-7f) {
      mass = fmaxf(0.f, mass - dm);
      if (dm * 0.68f > 1e-7f) pc_splat(ds, lysed, uc, dm * 0.68f);
    }
    if (mass <= 1e-5f) {
      P.pm.w = 0.f;
      parcels[g] = P;
      return;
    }
  }
  // hypotonic hemolysis (particles.ts:400-425)
  float cell_res;
  {
#pragma clang fp contract(off)
    const float cr = (float)g * 0.6180339887f + 0.17f;
    cell_res = cr - floorf(cr);
  }
  // cSample.r + cSample.g (particles.ts:403-404): the dye (the intact
  // concentration) plus the lysed one when the 4-channel concentration is on
  float cn = sample_cell<float>(dye, ds, uc.x, uc.y, uc.z);
  if (lysed_conc) cn += sample_cell<float>(lysed_conc, ds, uc.x, uc.y, uc.z);
  const float dilute = 1.f - pc_ss(0.012f, 0.22f, cn);
  const float disp = pc_ss(1.1f, 7.8f, age + dilute * 2.2f + lysis * 2.5f);
  const float late = pc_ss(7.2f, 16.f, age) * dilute;
  const float min_intact =
      kParcelMass0 * mixv(mixv(0.46f, 0.64f, cell_res),
                          mixv(mixv(0.07f, 0.20f, cell_res), mixv(0.014f, 0.052f, cell_res), late), disp);
  const float headroom = mass - min_intact;
  const float shelter = pc_ss(ymin + 0.024f, ymin + 0.062f, pos.y);
  if (dt > 0.f && headroom > 1e-7f && shelter > 0.02f) {
    const float shear = mixv(0.22f, 1.f, pc_ss(0.0005f, 0.0120f, sp2));
    const float water = (1.f - 0.72f * pc_ss(0.016f, 0.30f, cn)) * shelter * shear;
    lysis = clampv(lysis + S.k_lysis * (1.85f * water) * dt, 0.f, 0.98f);
    const float frag = pc_ss(0.04f, 0.48f, lysis) * water;
    const float dm = fminf(headroom, mass * clampv(frag * S.k_lysis * 1.95f * dt, 0.f, 0.16f));
    if (dm > 1e-7f) {
      mass = fmaxf(0.f, mass - dm);
      pc_splat(ds, lysed, uc, dm);
    }
  }
  P.pm = make_float4(pos.x, pos.y, pos.z, mass);
  P.sl = make_float4(st.x, st.y, st.z, ticks + clampv(lysis, 0.f, 0.98f));
  parcels[g] = P;

  // the intact mass: one trilinear splat for short parcels, else a 4-point
  // tapered ribbon along the stretch and a sheet tangent (particles.ts:430-461)
  const float sl = sqrtf(st.x * st.x + st.y * st.y + st.z * st.z);
  if (sl < 0.0038f) {
    if (mass > 1e-6f) pc_splat(ds, intact, uc, mass);
    return;
  }
  const float sxz = sqrtf(st.x * st.x + st.z * st.z);
  float3 ht = make_float3(1.f, 0.f, 0.f);
  if (sxz > 1e-4f) {
    const float m = fmaxf(sxz, 1e-5f);
    ht = make_float3(-st.z / m, 0.f, st.x / m);
  }
  const float3 sc = make_float3(st.y * ht.z - st.z * ht.y, st.z * ht.x - st.x * ht.z,
                                st.x * ht.y - st.y * ht.x);
  const float scl = sqrtf(sc.x * sc.x + sc.y * sc.y + sc.z * sc.z);
  float3 cc = ht;
  if (scl > 1e-4f) {
    const float m = fmaxf(scl, 1e-5f);
    cc = make_float3(sc.x / m, sc.y / m, sc.z / m);
  }
  const float hm = pc_ss(0.35f, 0.85f, fabsf(st.y) / fmaxf(sl, 1e-5f));
  float3 nn = make_float3(mixv(cc.x, ht.x, hm), mixv(cc.y, ht.y, hm), mixv(cc.z, ht.z, hm));
  const float nl = sqrtf(nn.x * nn.x + nn.y * nn.y + nn.z * nn.z);
  const float span = clampv(sl * 0.62f, 0.0032f, 0.0085f) / nl;
  const float3 sv = make_float3(nn.x * span, nn.y * span, nn.z * span);
  const float a[4] = {-0.42f, -0.14f, 0.14f, 0.42f}, b[4] = {0.f, 1.f, -1.f, 0.f},
              wt[4] = {0.21f, 0.29f, 0.29f, 0.21f};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const float3 pq = make_float3(pos.x + a[q] * st.x + b[q] * sv.x, pos.y + a[q] * st.y + b[q] * sv.y,
                                  pos.z + a[q] * st.z + b[q] * sv.z);
    const float mq = wt[q] * mass;
    if (mq > 1e-6f) pc_splat(ds, intact, pc_clamp_uvw(pc_uvw(S, pq)), mq);
  }
}

// fills n counters with 0 (a kernel, not a memset, so a captured graph may hold it)
// with 16-byte stores (p 16-byte aligned), the n mod 4 tail one by one
__global__ void __launch_bounds__(256) k_zero_u32(unsigned *__restrict__ p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x, n4 = n / 4;
  if (i < n4)
    reinterpret_cast<uint4 *>(p)[i] = make_uint4(0u, 0u, 0u, 0u);
  else if (i < n4 + (n & 3))
    p[4 * n4 + (i - n4)] = 0u;
}

}  // namespace mac3d
