// conc_kernels.hpp — the reference's pass B: parcel emission from the
// draining necks (PARTICLE_EMIT, src/shaders/particles.ts:53-186), the lysed
// channel's predictor (CONC_PREDICT, src/shaders/advect.ts:28-72) and the
// 4-channel concentration resolve (CONC_RESOLVE, advect.ts:81-298).  FP32, as
// the reference's parcels are (its concentration is rgba16float; FP32 here).
//
// The concentration is four planes of n floats on its own grid: intact RBC
// (.r), lysed Hb (.g), oxygenation/shear state (.b), floor bed (.a).  That
// grid is the MAC grid or, as the reference holds it (CONC_GRID 192x256x192
// over VEL_GRID 96x128x96, common.ts:4-15), twice as fine per axis; the
// kernels address it in uvw = (cell + 1/2) / N and sample the MAC grid's
// centre velocity there, as the reference's textureSampleLevel does.
#pragma once

#include "common.hpp"
#include "parcel_kernels.hpp"
#include "step_kernels.hpp"
#include "wave_kernels.hpp"

namespace mac3d {

// the reference caps a substep's emission at 8192 parcels
// (BloodDropEngine.ts:1835); the emission kernel runs that many threads
constexpr int kMaxEmit = 8192;

// the pending emission of the next parcel stage, in device memory so a
// captured step reads the current one
struct EmitCtl {
  int cursor, count;
};

struct ConcScene {
  float half_xz, y_min, y_max;  // world box (domainBounds)
  float ext;                    // wave half-extent (WAVE_HALF_EXTENT)
  int wave_n;
};

__device__ inline float fr32(float x) {
#pragma clang fp contract(off)
  return x - floorf(x);
}

// hash33 (common.ts:76-80) in single precision with every product and sum
// rounded on its own: contraction into fused multiply-adds is off in this
// block.  Plain operators, not __fmul_rn/__fadd_rn: those are inline header
// functions compiled with the file's contraction setting, so once inlined
// their product and sum still fuse.
__device__ inline void hash33_rn(float px, float py, float pz, float out[3]) {
#pragma clang fp contract(off)
  const float a0 = fr32(px * 0.1031f), b0 = fr32(py * 0.1030f), c0 = fr32(pz * 0.0973f);
  const float t = a0 * (b0 + 33.33f) + b0 * (a0 + 33.33f) + c0 * (c0 + 33.33f);
  const float a = a0 + t, b = b0 + t, c = c0 + t;
  out[0] = fr32((a + b) * c);
  out[1] = fr32((a + a) * b);
  out[2] = fr32((b + a) * a);
}

// PARTICLE_EMIT: thread g < ctl->count writes a new parcel at ring slot
// (cursor + g) mod np, in the drop slot its share of the summed drain rates
// picks, after moving a live parcel's mass there into the lysed counter (a
// cell of the splat grid d).
__global__ void __launch_bounds__(256)
k_parcel_emit(Dims d, ConcScene S, const EmitCtl *__restrict__ ctl,
              const WaveDrop *__restrict__ drops, const SceneClock *__restrict__ clk,
              const float4 *__restrict__ wave, int np, Parcel *__restrict__ parcels,
              unsigned *__restrict__ lysed) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int count = ctl->count;
  if (g >= count || np <= 0) return;
  float total = 0.f;
  for (int q = 0; q < WV_MAX_DROPS; ++q)
    if (drops[q].seed > 0.f && drops[q].drain > 0.0001f) total += drops[q].drain;
  if (total <= 1e-6f) return;
  const float target = (g + 0.5f) / (float)(count > 1 ? count : 1) * total;
  float acc = 0.f;
  int pick = 0;
  for (int q = 0; q < WV_MAX_DROPS; ++q)
    if (drops[q].seed > 0.f && drops[q].drain > 0.0001f) {
      acc += drops[q].drain;
      pick = q;
      if (acc >= target) break;
    }
  const WaveDrop dr = drops[pick];
  const float seed = fmaxf(dr.seed, 0.5f);
  const int ring = (int)(((long long)ctl->cursor + g) % np);
  Parcel P = parcels[ring];
  const float hx = S.half_xz, y0 = S.y_min, ytop = S.y_max;
  if (P.pm.w > 1e-6f) {  // the overwritten parcel's mass goes to the lysed counter
    float px = P.pm.x, py = P.pm.y, pz = P.pm.z;
    if (py < ytop - 0.32f) py = y0 + 0.012f;
    const float ux = (px + hx) / (2.f * hx), uy = (py - y0) / (ytop - y0),
                uz = (pz + hx) / (2.f * hx);
    if (ux > 0.002f && uy > 0.002f && uz > 0.002f && ux < 0.998f && uy < 0.998f && uz < 0.998f) {
      const int ci = clampi((int)floorf(ux * d.nx), 0, d.nx - 1);
      const int cj = clampi((int)floorf(uy * d.ny), 0, d.ny - 1);
      const int ck = clampi((int)floorf(uz * d.nz), 0, d.nz - 1);
      const unsigned q = (unsigned)rintf(P.pm.w * kSplatScale);
      if (q) atomicAdd(lysed + idx3(d, ci, cj, ck), q);
    }
  }
  // the R3 plastic sequence with the hash jitter (particles.ts:120-126),
  // every product and sum rounded on its own
  float r3x, r3y, r3z;
  {
#pragma clang fp contract(off)
    const float seq = (float)ring + seed * 137.0f;
    float h[3];
    hash33_rn((float)g, clk->sim_time * 60.0f, seed, h);
    r3x = fr32(seq * 0.8191725134f + h[0] * 0.032f);
    r3y = fr32(seq * 0.6710436067f + h[1] * 0.032f);
    r3z = fr32(seq * 0.5497004779f + h[2] * 0.032f);
  }
  const float th = 6.28318530718f * r3x;
  const float az = 1.f + 0.036f * cosf(4.f * th + seed * 1.73f) +
                   0.026f * cosf(6.f * th - seed * 2.31f) + 0.018f * sinf(8.f * th + seed * 0.89f);
  const float rdrop = fmaxf(dr.r, 0.010f);
  const float emerg = (dr.y - ytop) / fmaxf(rdrop, 0.018f);
  const float nasc = wv_smoothstep(-0.32f, 0.18f, emerg);
  const float slug = wv_smoothstep(0.26f, 0.85f, dr.drain);
  const float wake = mixv(0.46f, 1.f, wv_smoothstep(0.018f, 0.26f, dr.drain));
  const float rneck = mixv(0.013f * wake, clampv(rdrop * 0.84f + 0.014f, 0.020f, 0.056f), slug) * az;
  const float inner = 0.37f * slug;
  const bool cap = slug > 0.08f && r3z < 0.16f;
  const float rann = mixv(inner, 1.f, powf(r3y, mixv(0.50f, 0.44f, slug)));
  const float rcap = mixv(0.14f * slug, fmaxf(inner * 1.32f, 0.08f), sqrtf(r3y));
  const float rfrac = mixv(cap ? rcap : rann, sqrtf(r3y), nasc * 0.82f);
  const float r = rneck * rfrac;
  const float wa = clk->sim_time * 1.85f + seed * 3.17f;
  const float wd = 0.0038f * (1.f - slug) * (1.f - wake * 0.45f);
  const float px = dr.x + cosf(wa) * wd + cosf(th) * r;
  const float pz = dr.z + sinf(wa * 0.83f + seed) * wd + sinf(th) * r;
  // the live crater under the parcel (particles.ts:148-151)
  const float E = S.ext;
  const float eta = clampv(wv_sample(wave, S.wave_n, clampv((px + E) / (2.f * E), 0.002f, 0.998f),
                                     clampv((pz + E) / (2.f * E), 0.002f, 0.998f)).x,
                           -0.085f, 0.015f);
  const float top = ytop + fminf(0.f, eta) - 0.0038f;
  const float sag = (1.f - rfrac * rfrac) * (0.0085f * slug + 0.0045f * nasc);
  const float zs = cap ? r3z * 3.6f : (r3z - 0.16f * slug) / fmaxf(1.f - 0.16f * slug, 0.5f);
  const float py = top - sag - mixv(0.0014f, 0.0135f, clampv(zs, 0.f, 1.f));
  const float frag = 1.f - wv_smoothstep(0.022f, 0.056f, rdrop);
  const float rim = wv_smoothstep(0.58f, 1.f, rfrac);
  const float lys = clampv(0.14f * rim + 0.28f * frag * (0.45f + 0.55f * rfrac), 0.f, 0.44f);
  const float tx = -sinf(th) * fmaxf(rfrac, 0.25f), ty = -0.80f, tz = cosf(th) * fmaxf(rfrac, 0.25f);
  const float tl = 0.0036f / sqrtf(tx * tx + ty * ty + tz * tz);
  P.pm = make_float4(px, py, pz, kParcelMass0);
  P.sl = make_float4(tx * tl, ty * tl, tz * tl, lys);
  parcels[ring] = P;
}

// the emission is consumed by the stage that ran it
__global__ void k_emit_done(EmitCtl *__restrict__ ctl) {
  if (threadIdx.x == 0 && blockIdx.x == 0) ctl->count = 0;
}

// CONC_PREDICT (advect.ts:28-72) of the lysed channel g on the
// concentration grid d: the lateral walls and the top row 0 (:42-45); v0 the
// centre velocity (MAC grid dv) at uvw = (cell + 1/2) / N; below speed^2
// 2.5e-5 the cell's own value (:54-58); below 0.012 the one-stage trace
// uvw - v0 dt / domainSize (:60-62), else the RK2 trace through the midpoint
// velocity (:63-67); uvw clamped to [0.001, 0.999], dt to [0, 0.025].
__global__ void __launch_bounds__(256)
k_conc_predict(Dims d, Dims dv, ConcScene S, float dt, const float3 *__restrict__ cvel,
               const float *__restrict__ g, float *__restrict__ pred) {
  MAC3D_CELL_PROLOGUE
  if (conc_wall(d, i, j, k)) {
    pred[c] = 0.f;
    return;
  }
  const float dtc = clampv(dt, 0.f, 0.025f);
  const float dom0 = 2.f * S.half_xz, dom1 = S.y_max - S.y_min;
  const float u0 = (i + 0.5f) / (float)d.nx, u1 = (j + 0.5f) / (float)d.ny,
              u2 = (k + 0.5f) / (float)d.nz;
  float v0[3];
  sample_vec<float>(cvel, dv, u0, u1, u2, v0);
  const float sp2 = v0[0] * v0[0] + v0[1] * v0[1] + v0[2] * v0[2];
  if (sp2 < 2.5e-5f) {
    pred[c] = g[c];
    return;
  }
  float v[3] = {v0[0], v0[1], v0[2]};
  if (sp2 >= 0.012f) {
    const float m0 = clampv(u0 - 0.5f * (v0[0] * dtc) / dom0, 0.001f, 0.999f);
    const float m1 = clampv(u1 - 0.5f * (v0[1] * dtc) / dom1, 0.001f, 0.999f);
    const float m2 = clampv(u2 - 0.5f * (v0[2] * dtc) / dom0, 0.001f, 0.999f);
    sample_vec<float>(cvel, dv, m0, m1, m2, v);
  }
  const float b0 = clampv(u0 - (v[0] * dtc) / dom0, 0.001f, 0.999f);
  const float b1 = clampv(u1 - (v[1] * dtc) / dom1, 0.001f, 0.999f);
  const float b2 = clampv(u2 - (v[2] * dtc) / dom0, 0.001f, 0.999f);
  pred[c] = sample_cell<float>(g, d, b0, b1, b2);
}

// The intact (.r) and lysed (.g) channels of a concentration grid dc finer
// than the MAC grid d, as the Casson pass, its buoyancy and the unfurling
// gate read them per MAC cell (fluid.ts:446-447, 477-478, 491): samples at
// the cell centre and at the top y-face (uvw clamped to [0.001, 0.999]).
// out: four MAC-grid planes, .r at the centre, .g at the centre, .r at the
// face, .g at the face.
__global__ void __launch_bounds__(256)
k_conc_at_mac(Dims d, Dims dc, const float *__restrict__ conc, float *__restrict__ out) {
  MAC3D_CELL_PROLOGUE
  const size_t n = d.cells(), ncc = dc.cells();
  const float *r = conc, *g = conc + ncc;
  const float u0 = (i + 0.5f) / (float)d.nx, u1 = (j + 0.5f) / (float)d.ny,
              u2 = (k + 0.5f) / (float)d.nz;
  out[c] = sample_cell<float>(r, dc, u0, u1, u2);
  out[n + c] = sample_cell<float>(g, dc, u0, u1, u2);
  const float f0 = clampv(u0, 0.001f, 0.999f), f2 = clampv(u2, 0.001f, 0.999f);
  const float f1 = clampv((j + 1.f) / (float)d.ny, 0.001f, 0.999f);
  out[2 * n + c] = sample_cell<float>(r, dc, f0, f1, f2);
  out[3 * n + c] = sample_cell<float>(g, dc, f0, f1, f2);
}

// CONC_RESOLVE (advect.ts:81-298) on the concentration grid d: src/out the
// four planes (intact, lysed, oxygenation, bed); pg the lysed predictor;
// si/sl the splat counters; cvel the centre velocity of the MAC grid dv.
__global__ void __launch_bounds__(256)
k_conc_resolve(Dims d, Dims dv, ConcScene S, float dt, float strength,
               const float3 *__restrict__ cvel, const float *__restrict__ src,
               const float *__restrict__ pg, const unsigned *__restrict__ si,
               const unsigned *__restrict__ sl, float *__restrict__ out) {
  MAC3D_CELL_PROLOGUE
  const size_t n = d.cells();
  const float *og = src + n, *oa = src + 3 * n;
  float *o0 = out, *o1 = out + n, *o2 = out + 2 * n, *o3 = out + 3 * n;
  if (conc_wall(d, i, j, k)) {
    o0[c] = 0.f; o1[c] = 0.f; o2[c] = 0.f; o3[c] = 0.f;
    return;
  }
  const float Nx = (float)d.nx, Ny = (float)d.ny, Nz = (float)d.nz;
  const unsigned ri = si[c], rl = sl[c];
  const bool floor_zone = j <= 7;
  auto zero = [&]() { o0[c] = 0.f; o1[c] = 0.f; o2[c] = 0.f; o3[c] = 0.f; };
  if (ri == 0u && rl == 0u) {  // the octant positivity filter (advect.ts:151-164)
    const float a = sample_cell<float>(pg, d, i / Nx, j / Ny, k / Nz);
    const float b = sample_cell<float>(pg, d, (i + 1.f) / Nx, (j + 1.f) / Ny, (k + 1.f) / Nz);
    const float bed_on = floor_zone ? oa[c] : 0.f;
    if (a + b <= 1.0e-7f && bed_on <= 1.0e-6f) { zero(); return; }
  }
  auto at = [&](const float *f, int x, int y, int z) {
    return f[idx3(d, clampi(x, 0, d.nx - 1), clampi(y, 0, d.ny - 1), clampi(z, 0, d.nz - 1))];
  };
  auto cnt = [&](int x, int y, int z) {
    return (float)si[idx3(d, clampi(x, 0, d.nx - 1), clampi(y, 0, d.ny - 1),
                          clampi(z, 0, d.nz - 1))] * (1.f / kSplatScale);
  };
  const float gpred = pg[c];
  const float lapg = (at(pg, i + 1, j, k) + at(pg, i - 1, j, k) + at(pg, i, j + 1, k) +
                      at(pg, i, j - 1, k) + at(pg, i, j, k + 1) + at(pg, i, j, k - 1)) * (1.f / 6.f);
  if (!floor_zone && ri == 0u && rl == 0u && gpred <= 1.2e-6f && lapg <= 1.2e-6f) { zero(); return; }
  float lapi = 0.f;
  if (ri > 0u || gpred > 8.0e-5f || lapg > 8.0e-5f)
    lapi = (cnt(i + 1, j, k) + cnt(i - 1, j, k) + cnt(i, j + 1, k) + cnt(i, j - 1, k) +
            cnt(i, j, k + 1) + cnt(i, j, k - 1)) * (1.f / 6.f);
  const float hx = S.half_xz, y0 = S.y_min, ytop = S.y_max;
  const float dtc = clampv(dt, 0.f, 0.025f);
  const float u0 = (i + 0.5f) / Nx, u1 = (j + 0.5f) / Ny, u2 = (k + 0.5f) / Nz;
  const float dom0 = 2.f * hx, dom1 = ytop - y0, dom2 = 2.f * hx;
  const float px = -hx + u0 * dom0, py = y0 + u1 * dom1, pz = -hx + u2 * dom2;
  float v0[3], vm[3];
  sample_vec<float>(cvel, dv, u0, u1, u2, v0);
  const float m0 = clampv(u0 - 0.5f * (v0[0] * dtc) / dom0, 0.001f, 0.999f);
  const float m1 = clampv(u1 - 0.5f * (v0[1] * dtc) / dom1, 0.001f, 0.999f);
  const float m2 = clampv(u2 - 0.5f * (v0[2] * dtc) / dom2, 0.001f, 0.999f);
  sample_vec<float>(cvel, dv, m0, m1, m2, vm);
  const float b0 = clampv(u0 - (vm[0] * dtc) / dom0, 0.001f, 0.999f);
  const float b1 = clampv(u1 - (vm[1] * dtc) / dom1, 0.001f, 0.999f);
  const float b2 = clampv(u2 - (vm[2] * dtc) / dom2, 0.001f, 0.999f);
  const float f0 = clampv(u0 + (v0[0] * dtc) / dom0, 0.001f, 0.999f);
  const float f1 = clampv(u1 + (v0[1] * dtc) / dom1, 0.001f, 0.999f);
  const float f2 = clampv(u2 + (v0[2] * dtc) / dom2, 0.001f, 0.999f);
  // lysed: MacCormack with the 8-texel limiter (advect.ts:195-222)
  const float est = sample_cell<float>(pg, d, f0, f1, f2);
  const float corr = gpred + strength * (og[c] - est);
  const int bi = (int)floorf(b0 * Nx - 0.5f), bj = (int)floorf(b1 * Ny - 0.5f),
            bk = (int)floorf(b2 * Nz - 0.5f);
  float lo = 3.4e38f, hi = -3.4e38f;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const float v = at(og, bi + (q & 1), bj + ((q >> 1) & 1), bk + (q >> 2));
    lo = fminf(lo, v);
    hi = fmaxf(hi, v);
  }
  float cl = clampv(corr, lo, hi);
  cl = mixv(cl, lapg, clampv(dtc * 5.60f, 0.f, 0.105f));
  cl *= expf(-dtc * mixv(0.016f, 0.034f, wv_smoothstep(y0 + 0.045f, y0 + 0.26f, py)));
  const float dm = (float)rl * (1.f / kSplatScale);
  const float capture = 1.f - wv_smoothstep(y0 + 0.018f, y0 + 0.058f, py);
  cl += dm * (1.f - capture * 0.72f);
  if (cl < 2.5e-6f) cl = 0.f;
  // floor bed (advect.ts:241-272)
  const float rh = sqrtf(px * px + pz * pz);
  float bed = 0.f;
  if (floor_zone) {
    const float bc = oa[c];
    const float bh = 0.25f * (at(oa, i + 1, j, k) + at(oa, i - 1, j, k) + at(oa, i, j, k + 1) +
                              at(oa, i, j, k - 1));
    const float ba = at(oa, i, j + 1, k), bb = at(oa, i, j > 0 ? j - 1 : 0, k);
    const float spread = clampv(dtc * (3.20f + 6.50f * fminf(bc, 0.65f)), 0.f, 0.165f);
    const float vert = clampv(dtc * 0.75f, 0.f, 0.026f);
    const float settle = clampv(dtc * 2.80f, 0.f, 0.090f);
    const float decay = expf(-dtc * mixv(0.095f, 0.28f, wv_smoothstep(hx * 0.35f, hx * 0.78f, rh)));
    const float drained = (j == 0 ? bc : bc * (1.f - settle)) * decay;
    const float above = j >= 7 ? 0.f : ba * settle;
    const float sm = mixv(mixv(drained + above, bh, spread), 0.5f * (ba + bb), vert);
    const float fade = 1.f - 0.38f * wv_smoothstep(3.5f, 7.0f, (float)j);
    bed = (sm + dm * capture * 0.68f) * fade;
    if (bed < 1.0e-6f) bed = 0.f;
  }
  // intact and the wall sponge (advect.ts:276-285)
  const float wall = 1.f - wv_smoothstep(hx * 0.68f, hx * 0.94f, rh);
  const float taper = 1.f - wv_smoothstep(hx * 0.52f, hx * 0.88f, rh);
  bed = fminf(bed * wall, 0.54f * expf(-0.56f * (float)j) * taper);
  const float craw = (float)ri * (1.f / kSplatScale);
  const float intact = mixv(craw, lapi, ri > 0u ? 0.36f : 0.18f) * wall + bed;
  cl *= wall;
  // oxygenation / shear dispersion (advect.ts:290-297)
  const float tot = intact + cl;
  float ox = 0.f;
  if (tot > 1.0e-5f) {
    const float fl = clampv(cl / fmaxf(tot, 1.0e-5f), 0.f, 1.f), fi = 1.f - fl;
    const float fb = wv_smoothstep(0.f, 0.85f, bed / fmaxf(tot, 1.0e-5f));
    const float speed = sqrtf(v0[0] * v0[0] + v0[1] * v0[1] + v0[2] * v0[2]);
    const float dx = v0[0] - vm[0], dy = v0[1] - vm[1], dz = v0[2] - vm[2];
    const float conv = sqrtf(dx * dx + dy * dy + dz * dz) / fmaxf(0.5f * dtc, 1.0e-3f);
    const float sh = wv_smoothstep(0.014f, 0.135f, speed + 0.38f * conv);
    const float thin = (1.f - wv_smoothstep(0.024f, 0.24f, intact)) *
                       wv_smoothstep(0.0006f, 0.0065f, intact);
    const float act = thin * (0.38f + 0.62f * sh) * fi;
    const float raw = clampv(0.08f + 0.62f * act + 0.36f * sh * fi, 0.f, 1.f);
    ox = raw * (1.f - 0.82f * fb);
  }
  o0[c] = intact; o1[c] = cl; o2[c] = ox; o3[c] = bed;
}

// FIELD_CONC transfers: planes <-> 4 values per cell (r, g, b, a), the
// layout of the reference's rgba concentration texture
template <class A, class B>
__global__ void __launch_bounds__(256)
k_planes_to_rgba(const A *__restrict__ planes, B *__restrict__ rgba, size_t n) {
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
#pragma unroll
  for (int q = 0; q < 4; ++q) rgba[4 * t + q] = (B)planes[q * n + t];
}
template <class A, class B>
__global__ void __launch_bounds__(256)
k_rgba_to_planes(const A *__restrict__ rgba, B *__restrict__ planes, size_t n) {
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
#pragma unroll
  for (int q = 0; q < 4; ++q) planes[q * n + t] = (B)rgba[4 * t + q];
}

}  // namespace mac3d
