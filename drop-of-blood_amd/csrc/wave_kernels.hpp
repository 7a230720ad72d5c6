// Free-surface wave solver and the top Dirichlet pressure it feeds: the
// reference's WAVE_SPATIAL_FORCE / radix-2 FFT / WAVE_SPECTRAL_PROPAGATE /
// WAVE_FINALIZE passes (src/shaders/wave.ts:56-388) and
// evalTopDirichletPressure (src/shaders/fluid.ts:16-78), which the MAC
// divergence (fluid.ts:548-551) and projection (fluid.ts:896-898) read on
// the lid plane.
//
// The wave lives on an n x n texel grid over [-E, E]^2 (E = 1.60 m, n = 512 in
// the reference), texel (ix, iy) at uv = (i + 1/2) / n, x fastest.  State per
// texel: (eta, d eta/dt, 9-point Laplacian of eta, surfactant film), FP32
// (the reference stores it as rgba16float; the spectra are FP32 in both).
// One wave step: force pass -> forward 2-D FFT of eta + i deta/dt -> spectral
// propagation -> inverse 2-D FFT -> finalize, in four launches (force + row
// FFT; column FFT + propagation + inverse column FFT; inverse row FFT;
// finalize).  The FFTs are unitary (1/sqrt(n) per 1-D transform), as the
// reference's.
#pragma once
#include "common.hpp"

namespace mac3d {

#define WV_MAX_DROPS 8

// one drop slot of SimUniforms.drops (common.ts:26-43): posAndRadius and
// velOscAndDrain
struct WaveDrop {
  float x, y, z, r;
  float vy, osc, drain, seed;
};

// per-drop constants of the advection's neck terms, formed on the host when
// the drops are set: (cos, sin) of the azimuthal phases 1.73 seed,
// -2.31 seed and 0.89 seed, and the slug drive
// clamp(max(drain, 0)^0.66 * 0.84, 0, 1.24) (fluid.ts:211-240)
struct DropPhase {
  float c1, s1, c2, s2, c3, s3, slug_drive, pad;
};

struct WaveParams {
  int n;            // texels per side (power of two)
  float ext;        // wave domain half-width E (m)
  float half_xz;    // tank half-width (domainBounds.x)
  float depth;      // tank depth (domainBounds.z - domainBounds.y)
  float surface_y;  // still surface height (surfaceParams.x)
  float marangoni;  // surfaceParams.y
  float dt;         // substep dt clamped to [0, 0.025] (viewportAndDt.z)
};

__device__ inline float wv_smoothstep(float e0, float e1, float x) {
  float t = clampv((x - e0) / (e1 - e0), 0.f, 1.f);
  return t * t * (3.f - 2.f * t);
}

// texel of the wave state, indices clamped to the grid (loadWaveClamped)
__device__ inline float4 wv_load(const float4 *__restrict__ w, int n, int x, int y) {
  x = clampi(x, 0, n - 1);
  y = clampi(y, 0, n - 1);
  return w[(size_t)y * n + x];
}

// bilinear clamp-to-edge sample of the wave state at uv (textureSampleLevel
// with the linear sampler, BloodDropEngine.ts:497-505)
__device__ inline float4 wv_sample(const float4 *__restrict__ w, int n, float u, float v) {
  float tx = u * n - 0.5f, ty = v * n - 0.5f;
  float fx0 = floorf(tx), fy0 = floorf(ty);
  int x0 = (int)fx0, y0 = (int)fy0;
  float fx = tx - fx0, fy = ty - fy0;
  float4 a = wv_load(w, n, x0, y0), b = wv_load(w, n, x0 + 1, y0);
  float4 c = wv_load(w, n, x0, y0 + 1), e = wv_load(w, n, x0 + 1, y0 + 1);
  auto mx = [](float4 p, float4 q, float t) {
    return make_float4(mixv(p.x, q.x, t), mixv(p.y, q.y, t), mixv(p.z, q.z, t), mixv(p.w, q.w, t));
  };
  return mx(mx(a, b, fx), mx(c, e, fx), fy);
}

// trilinear clamp-to-edge sample of a MAC face field stored per cell at
// normalised uvw (the velMac texture read of the force pass)
template <class T>
__device__ inline float wv_sample_mac(const T *__restrict__ f, const Dims &d, float u, float v,
                                      float w) {
  float t[3] = {u * d.nx - 0.5f, v * d.ny - 0.5f, w * d.nz - 0.5f};
  const int N[3] = {d.nx, d.ny, d.nz};
  int i0[3], i1[3];
  float fr[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float fl = floorf(t[a]);
    fr[a] = t[a] - fl;
    i0[a] = clampi((int)fl, 0, N[a] - 1);
    i1[a] = clampi((int)fl + 1, 0, N[a] - 1);
  }
  auto at = [&](int i, int j, int k) { return (float)f[idx3(d, i, j, k)]; };
  float x00 = mixv(at(i0[0], i0[1], i0[2]), at(i1[0], i0[1], i0[2]), fr[0]);
  float x10 = mixv(at(i0[0], i1[1], i0[2]), at(i1[0], i1[1], i0[2]), fr[0]);
  float x01 = mixv(at(i0[0], i0[1], i1[2]), at(i1[0], i0[1], i1[2]), fr[0]);
  float x11 = mixv(at(i0[0], i1[1], i1[2]), at(i1[0], i1[1], i1[2]), fr[0]);
  return mixv(mixv(x00, x10, fr[1]), mixv(x01, x11, fr[1]), fr[2]);
}

// WAVE_SPATIAL_FORCE (wave.ts:56-178) at texel (ix, iy): drop impulses
// (zero-integral polar Ricker craters and capillary contact impulses), the
// surfactant film and its Marangoni pull, the MAC grid's upward top-layer
// velocity, the radial absorber.  Gives z = eta + i deta/dt for the FFT and
// the updated film.
template <class T>
__device__ inline void wave_force_texel(const WaveParams &P, const WaveDrop *__restrict__ drops,
                                        const float4 *__restrict__ wave, const Dims &d,
                                        const T *__restrict__ vmac, int ix, int iy, float2 &zv,
                                        float &film_out) {
  const int n = P.n;
  const float4 c = wv_load(wave, n, ix, iy);
  float eta = c.x, vel = c.y, film = c.w;
  const float dt = P.dt;
  const float time_scale = dt / 0.0166667f;
  const float lap_scale = ((float)n / 256.f) * ((float)n / 256.f);
  const float raw_lap = 0.20f * (wv_load(wave, n, ix, iy + 1).w + wv_load(wave, n, ix, iy - 1).w +
                                 wv_load(wave, n, ix + 1, iy).w + wv_load(wave, n, ix - 1, iy).w) +
                        0.05f * (wv_load(wave, n, ix + 1, iy + 1).w + wv_load(wave, n, ix - 1, iy + 1).w +
                                 wv_load(wave, n, ix + 1, iy - 1).w + wv_load(wave, n, ix - 1, iy - 1).w) -
                        film;
  const float lap_film = raw_lap * lap_scale;
  const float wx = ((ix + 0.5f) / n) * (2.f * P.ext) - P.ext;
  const float wz = ((iy + 0.5f) / n) * (2.f * P.ext) - P.ext;

  float force_drops = 0.f, active = 0.f, max_drain = 0.f;
  for (int q = 0; q < WV_MAX_DROPS; ++q) {
    const WaveDrop s = drops[q];
    if (!(s.seed > 0.f && s.drain > 0.0005f)) continue;
    max_drain = fmaxf(max_drain, s.drain);
    const float rd = fmaxf(s.r, 0.018f);
    const float speed = clampv(-s.vy, 0.06f, 0.65f);
    const float ex = wx - s.x, ez = wz - s.z;
    const float r = sqrtf(ex * ex + ez * ez);
    const float r_foot = clampv(rd * 1.05f, 0.024f, 0.082f);
    const float rn = r / r_foot;
    const float emerg = (s.y - P.surface_y) / fmaxf(rd, 0.018f);
    const float excav = wv_smoothstep(-0.78f, -0.04f, emerg);
    const float bell = sinf(clampv((1.05f - emerg) / 1.83f, 0.f, 1.f) * 3.14159265f) * excav;
    const float shape = expf(-rn * rn) * (1.f - rn * rn);
    const float dyn = (6.8f * bell * (rd / 0.056f)) * (0.72f + 2.65f * speed);
    const float r_cap = clampv(rd * 0.44f, 0.013f, 0.030f);
    const float rnc = r / r_cap;
    const float cap_win = wv_smoothstep(-0.35f, 0.28f, emerg) * (1.f - wv_smoothstep(0.68f, 1.06f, emerg));
    const float cap_shape = expf(-rnc * rnc) * (1.f - rnc * rnc);
    const float cap = (4.6f * cap_win * (0.60f + 0.40f * (rd / 0.056f))) * (0.65f + 2.20f * speed);
    force_drops = force_drops - dyn * shape - cap * cap_shape;
    active = fmaxf(active, excav * expf(-0.5f * rn * rn));
    const float r_stain = fmaxf(0.018f, rd * 0.60f);
    const float stain = expf(-(r * r) / (r_stain * r_stain)) * clampv(s.drain * 0.42f, 0.f, 1.f);
    film = fmaxf(film, stain * 0.65f);
  }

  const float m = P.marangoni;
  const float eff_lap = mixv(fmaxf(lap_film, -0.0015f), lap_film, active);
  const float force_mar = m * 0.28f * eff_lap;
  film = clampv(film + m * 0.065f * raw_lap * time_scale, 0.f, 1.f);
  film = film * powf(0.88f, time_scale);

  float force_sub = 0.f;
  const float hx = P.half_xz;
  const float gate = clampv(active * 1.4f + wv_smoothstep(0.f, 0.06f, max_drain), 0.f, 1.f);
  const float rw = sqrtf(wx * wx + wz * wz);
  if (gate > 0.001f && rw < hx * 0.94f) {
    const float tu = (wx + hx) / (2.f * hx), tw = (wz + hx) / (2.f * hx);
    const float vsub = wv_sample_mac<T>(vmac, d, tu, 0.988f, tw);
    const float taper = 1.f - wv_smoothstep(hx * 0.62f, hx * 0.92f, rw);
    force_sub = clampv(fmaxf(vsub, 0.f), 0.f, 0.48f) * 0.42f * taper * gate;
  }
  vel = vel + (force_drops + force_mar + force_sub) * dt;

  const float absorb = wv_smoothstep(0.90f, 1.54f, rw);
  const float rdamp = expf(-dt * (24.f * absorb * absorb));
  const float border = 1.f - wv_smoothstep(1.22f, 1.56f, rw);
  eta = clampv(eta * rdamp * border, -0.14f, 0.14f);
  vel = clampv(vel * rdamp * border, -1.8f, 1.8f);
  zv = make_float2(eta, vel);
  film_out = film * border;
}

// log2(n) radix-2 butterfly stages in LDS over a line of n complex values
// held in bit-reversed order (decimation in time, as the reference's
// createRadix2Fft1DShader, wave.ts:180-252), natural order out, unscaled;
// tw[m] = exp(-2 pi i m / n) (in LDS: a stage waits on no memory load),
// conjugated for the inverse.  Every thread of the (1-D) block calls it: it
// holds a barrier per stage.
// (sb, nullable: a second line transformed in the same stages, so the two
// share the stages' barriers)
__device__ inline void wv_butterfly(float2 *sm, int i0, int i1, float2 w) {
  const float2 a = sm[i0], q = sm[i1];
  const float2 t = make_float2(q.x * w.x - q.y * w.y, q.x * w.y + q.y * w.x);
  sm[i0] = make_float2(a.x + t.x, a.y + t.y);
  sm[i1] = make_float2(a.x - t.x, a.y - t.y);
}
__device__ inline void wv_fft_stages(float2 *sm, int n, int logn, const float2 *tw,
                                     bool inverse, float2 *sb = nullptr) {
  const int h = n >> 1;
  for (int s = 0; s < logn; ++s) {
    const int half = 1 << s;
    for (int b = threadIdx.x; b < h; b += blockDim.x) {
      const int k = b & (half - 1);
      const int i0 = ((b >> s) << (s + 1)) + k, i1 = i0 + half;
      float2 w = tw[k << (logn - 1 - s)];
      if (inverse) w.y = -w.y;
      wv_butterfly(sm, i0, i1, w);
      if (sb) wv_butterfly(sb, i0, i1, w);
    }
    __syncthreads();
  }
}

__device__ inline int wv_brev(int i, int logn) { return (int)(__brev((unsigned)i) >> (32 - logn)); }

// Row pitch of the spectrum in complex values: n + 32, so that successive
// rows start 256 B apart modulo 4 KB and a column's elements spread over the
// L2 channels instead of all falling in one (a 4 KB row stride does that).
#define WV_PAD 32
__host__ __device__ inline size_t wv_pitch(int n) { return (size_t)n + WV_PAD; }

// the n/2 twiddles into LDS (the caller's barrier publishes them)
__device__ inline void wv_load_twiddles(float2 *tws, int n, const float2 *__restrict__ tw) {
  for (int m = threadIdx.x; m < (n >> 1); m += blockDim.x) tws[m] = tw[m];
}

// The force pass fused with the forward FFT of its row: one block per row
// iy evaluates the row's texels into LDS (bit-reversed), transforms it and
// writes the row's spectrum (1/sqrt(n) scaled) and the film.
template <class T>
__global__ void __launch_bounds__(512)
k_wave_force_rows(WaveParams P, const WaveDrop *__restrict__ drops,
                  const float4 *__restrict__ wave, Dims d, const T *__restrict__ vmac,
                  float2 *__restrict__ z, float *__restrict__ film, int logn,
                  const float2 *__restrict__ tw) {
  extern __shared__ float2 sm[];  // the row, then the twiddles
  const int n = P.n, iy = blockIdx.x;
  float2 *tws = sm + n;
  wv_load_twiddles(tws, n, tw);
  for (int ix = threadIdx.x; ix < n; ix += blockDim.x) {
    float2 zv;
    float f;
    wave_force_texel<T>(P, drops, wave, d, vmac, ix, iy, zv, f);
    film[(size_t)iy * n + ix] = f;
    sm[wv_brev(ix, logn)] = zv;
  }
  __syncthreads();
  wv_fft_stages(sm, n, logn, tws, false);
  const float sc = rsqrtf((float)n);
  for (int i = threadIdx.x; i < n; i += blockDim.x)
    z[(size_t)iy * wv_pitch(n) + i] = make_float2(sm[i].x * sc, sm[i].y * sc);
}

// One 1-D unitary complex FFT of length n per line, radix-2 in LDS (n/2
// threads, one butterfly each per stage), input in bit-reversed order as the
// reference's createRadix2Fft1DShader (wave.ts:180-252).  rows: lines are
// rows (x fastest, contiguous); else columns.  tw[m] = exp(-2 pi i m / n),
// conjugated for the inverse.
__global__ void __launch_bounds__(512)
k_wave_fft(float2 *__restrict__ z, int n, int logn, int rows, int inverse,
           const float2 *__restrict__ tw) {
  extern __shared__ float2 sm[];  // the line, then the twiddles
  const int line = blockIdx.x, tid = threadIdx.x;
  const size_t base = rows ? (size_t)line * wv_pitch(n) : (size_t)line;
  const size_t step = rows ? 1 : wv_pitch(n);
  float2 *tws = sm + n;
  wv_load_twiddles(tws, n, tw);
  for (int i = tid; i < n; i += blockDim.x) sm[i] = z[base + (size_t)wv_brev(i, logn) * step];
  __syncthreads();
  wv_fft_stages(sm, n, logn, tws, inverse != 0);
  const float sc = rsqrtf((float)n);
  for (int i = tid; i < n; i += blockDim.x) {
    const float2 v = sm[i];
    z[base + (size_t)i * step] = make_float2(v.x * sc, v.y * sc);
  }
}

// WAVE_SPECTRAL_PROPAGATE (wave.ts:254-331) of spectral texel (ux, uy), fp
// its value and fm the value at the mirror texel (-ux, -uy): Hermitian
// unpacking of the two real fields packed as Z = eta + i w, the
// gravity-capillary dispersion omega(k) = sqrt((g k + s k^3) tanh(k H)) over
// one dt with viscous damping and a high-k roll-off, DC and the Nyquist row
// and column set to 0, repacked.
__device__ inline float2 wv_propagate(const WaveParams &P, int ux, int uy, float2 fp, float2 fm) {
  const int n = P.n, hn = n >> 1;
  if ((ux == 0 && uy == 0) || ux == hn || uy == hn) return make_float2(0.f, 0.f);
  const float2 eh = make_float2(0.5f * (fp.x + fm.x), 0.5f * (fp.y - fm.y));
  const float2 vh = make_float2(0.5f * (fp.y + fm.y), 0.5f * (-fp.x + fm.x));
  const float mx = (float)(ux > hn ? ux - n : ux), mz = (float)(uy > hn ? uy - n : uy);
  const float two_pi_l = 6.28318530718f / (2.f * P.ext);
  const float kx = mx * two_pi_l, kz = mz * two_pi_l;
  const float k2 = kx * kx + kz * kz;
  const float k = sqrtf(k2);
  const float dt = P.dt;
  const float th = tanhf(fminf(k * P.depth, 12.f));
  const float omega = sqrtf(fmaxf((0.84f * k + 0.0024f * k * k2) * th, 1e-8f));
  const float visc = 0.36f + 0.00018f * k2;
  const float roll = 1.f - wv_smoothstep(335.f, 465.f, k);
  const float damp = expf(-visc * dt) * roll;
  const float ph = omega * dt;
  const float cp = cosf(ph), sp = sinf(ph);
  const float2 en = make_float2(damp * (eh.x * cp + vh.x * (sp / omega)),
                                damp * (eh.y * cp + vh.y * (sp / omega)));
  const float2 vn = make_float2(damp * (-eh.x * (omega * sp) + vh.x * cp),
                                damp * (-eh.y * (omega * sp) + vh.y * cp));
  return make_float2(en.x - vn.y, en.y + vn.x);
}

// Forward column FFT, propagation and inverse column FFT in one pass: block
// a (0 <= a <= n/2) holds column a and its mirror column n - a in LDS (the
// propagation pairs texel (a, y) with (n - a, n - y)); columns 0 and n/2 are
// their own mirrors.  Each thread propagates one mirror pair in place (it
// reads both values before writing either), so no two threads touch one
// texel.
__global__ void __launch_bounds__(512)
k_wave_colpair(WaveParams P, float2 *__restrict__ z, int logn, const float2 *__restrict__ tw) {
  extern __shared__ float2 sm[];  // column a, column n - a, the twiddles
  const int n = P.n, hn = n >> 1;
  const int a = blockIdx.x, b = (n - a) & (n - 1);
  const bool self = a == b;
  float2 *ca = sm, *cb = self ? sm : sm + n, *tws = sm + 2 * n;
  wv_load_twiddles(tws, n, tw);
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const size_t r = (size_t)wv_brev(i, logn) * wv_pitch(n);
    ca[i] = z[r + a];
    if (!self) cb[i] = z[r + b];
  }
  __syncthreads();
  wv_fft_stages(ca, n, logn, tws, false, self ? nullptr : cb);
  const float sc = rsqrtf((float)n);
  const int npairs = self ? hn + 1 : n;
  for (int uy = threadIdx.x; uy < npairs; uy += blockDim.x) {
    const int my = (n - uy) & (n - 1);
    const float2 fp = make_float2(ca[uy].x * sc, ca[uy].y * sc);
    const float2 fm = make_float2(cb[my].x * sc, cb[my].y * sc);
    const float2 op = wv_propagate(P, a, uy, fp, fm);
    const float2 om = wv_propagate(P, b, my, fm, fp);
    ca[uy] = op;
    cb[my] = om;
  }
  __syncthreads();
  // natural -> bit-reversed order for the inverse transform (pairwise swaps)
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const int r = wv_brev(i, logn);
    if (i < r) {
      float2 t = ca[i];
      ca[i] = ca[r];
      ca[r] = t;
      if (!self) {
        t = cb[i];
        cb[i] = cb[r];
        cb[r] = t;
      }
    }
  }
  __syncthreads();
  wv_fft_stages(ca, n, logn, tws, true, self ? nullptr : cb);
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    z[(size_t)i * wv_pitch(n) + a] = make_float2(ca[i].x * sc, ca[i].y * sc);
    if (!self) z[(size_t)i * wv_pitch(n) + b] = make_float2(cb[i].x * sc, cb[i].y * sc);
  }
}

// WAVE_FINALIZE (wave.ts:333-388): propagated (eta, deta/dt), the film of the
// force pass, the resolution-scaled 9-point Laplacian of eta, the outer
// border fade and clamps.
__global__ void __launch_bounds__(256)
k_wave_finalize(WaveParams P, const float2 *__restrict__ z, const float *__restrict__ film,
                float4 *__restrict__ wave) {
  const int n = P.n;
  const int ix = blockIdx.x * 16 + threadIdx.x, iy = blockIdx.y * 16 + threadIdx.y;
  if (ix >= n || iy >= n) return;
  auto eta = [&](int x, int y) {
    x = clampi(x, 0, n - 1);
    y = clampi(y, 0, n - 1);
    return z[(size_t)y * wv_pitch(n) + x].x;
  };
  const size_t t = (size_t)iy * n + ix;
  const float2 pr = z[(size_t)iy * wv_pitch(n) + ix];
  const float lap_scale = ((float)n / 256.f) * ((float)n / 256.f);
  const float lap = (0.20f * (eta(ix, iy + 1) + eta(ix, iy - 1) + eta(ix + 1, iy) + eta(ix - 1, iy)) +
                     0.05f * (eta(ix + 1, iy + 1) + eta(ix - 1, iy + 1) + eta(ix + 1, iy - 1) +
                              eta(ix - 1, iy - 1)) -
                     pr.x) *
                    lap_scale;
  const float wx = ((ix + 0.5f) / n) * (2.f * P.ext) - P.ext;
  const float wz = ((iy + 0.5f) / n) * (2.f * P.ext) - P.ext;
  const float border = 1.f - wv_smoothstep(1.22f, 1.56f, sqrtf(wx * wx + wz * wz));
  wave[t] = make_float4(clampv(pr.x * border, -0.14f, 0.14f), clampv(pr.y * border, -1.8f, 1.8f),
                        lap * border, film[t] * border);
}

// evalTopDirichletPressure (fluid.ts:16-78) on the lid plane, one value per
// (ix, iz): hydrostatic and capillary pressure of the wave under the cell,
// the drops' wetting-neck overpressure (with the m = 4, 6, 8 azimuthal
// modulation), times dt / h and a wall taper.  ptop[iz * nx + ix].
__global__ void __launch_bounds__(256)
k_top_pressure(WaveParams P, const WaveDrop *__restrict__ drops, const float4 *__restrict__ wave,
               int nx, int nz, float *__restrict__ ptop) {
  const int ix = blockIdx.x * 16 + threadIdx.x, iz = blockIdx.y * 16 + threadIdx.y;
  if (ix >= nx || iz >= nz) return;
  const float hx = P.half_xz;
  const float fx = (ix + 0.5f) / nx, fz = (iz + 0.5f) / nz;
  const float wx = mixv(-hx, hx, fx), wz = mixv(-hx, hx, fz);
  const float dt = P.dt;
  const float cell_h = (2.f * hx) / nx;
  const float dt_h = dt / cell_h;
  const float su = clampv((wx + P.ext) / (2.f * P.ext), 0.002f, 0.998f);
  const float sv = clampv((wz + P.ext) / (2.f * P.ext), 0.002f, 0.998f);
  const float4 ws = wv_sample(wave, P.n, su, sv);
  float pk = 2.45f * (0.155f * ws.x + 0.00095f * (-18.4f * ws.z));
  for (int q = 0; q < WV_MAX_DROPS; ++q) {
    const WaveDrop s = drops[q];
    if (!(s.seed > 0.f && s.drain > 0.0005f)) continue;
    const float ex = wx - s.x, ez = wz - s.z;
    const float r = sqrtf(ex * ex + ez * ez);
    const float th = atan2f(ez, ex);
    const float azim = 1.f + 0.022f * cosf(4.f * th + s.seed * 1.73f) +
                       0.016f * cosf(6.f * th - s.seed * 2.31f) + 0.010f * sinf(8.f * th + s.seed * 0.89f);
    const float slug = wv_smoothstep(0.32f, 0.92f, s.drain);
    const float r_or = clampv(s.r * 1.12f + mixv(0.018f, 0.032f, slug), 0.030f, 0.090f) * azim;
    const float rn = r / fmaxf(r_or, 1e-4f);
    const float speed = clampv(-s.vy, 0.08f, 0.65f);
    const float neck = (0.28f + 1.05f * rn * rn) * expf(-0.92f * rn * rn) * (1.f - 0.38f * rn * rn);
    pk = pk + s.drain * (0.014f + 0.048f * speed) * neck;
  }
  const float taper = 1.f - wv_smoothstep(hx * 0.72f, hx * 0.96f, fmaxf(fabsf(wx), fabsf(wz)));
  ptop[(size_t)iz * nx + ix] = pk * dt_h * taper;
}


// The advection's scene terms (FLUID_ADVECT_MAC, fluid.ts:198-321), applied to
// cell (i, j, k)'s advected and damped face values before the wall faces are
// zeroed: the drops' wetting-neck momentum (the vertical slug/stem draft on
// the y-face and the radial flare on the x- and z-faces, each drop's
// m = 4, 6, 8 azimuthal modulation), the surfactant film's Marangoni pull and
// the kinematic top condition from the wave (top three rows), and the soft
// circular wall and floor sponges.  World positions: the tank [-hx, hx] in x
// and z, [ytop - depth, ytop] in y with ytop the surface height (uvwToWorld,
// common.ts:45-64, with domainBounds.z = surfaceParams.x as in the scene).
// The azimuthal terms cos(m theta + phase) are formed from the unit direction
// (cos theta, sin theta) by complex powers and the drop's phase factors, which
// equals the reference's atan2-and-cosine form without per-cell trigonometry.
template <class T>
__device__ inline void advect_scene_terms(const Dims &d, int i, int j, int k, const WaveParams &P,
                                          const WaveDrop *__restrict__ drops,
                                          const DropPhase *__restrict__ phs,
                                          const float4 *__restrict__ wave, T &vx, T &vy, T &vz) {
  const T hx = (T)P.half_xz, ytop = (T)P.surface_y, depth = (T)P.depth, y0 = ytop - depth;
  const T nx = (T)d.nx, ny = (T)d.ny, nz = (T)d.nz;
  auto wx_of = [&](T u) { return -hx + u * (T(2) * hx); };
  auto wy_of = [&](T v) { return y0 + v * depth; };
  const T dt = (T)P.dt;
  const T cx = wx_of((i + T(0.5)) / nx), cy = wy_of((j + T(0.5)) / ny), cz = wx_of((k + T(0.5)) / nz);
  const T fy_y = wy_of((j + T(1)) / ny);              // y-face (uvwY)
  const T xf_x = wx_of((i + T(1)) / nx), zf_z = wx_of((k + T(1)) / nz);  // x-, z-faces
  const T surf = (T)P.surface_y;
  auto ss = [](T e0, T e1, T x) {
    T t = clampv((x - e0) / (e1 - e0), T(0), T(1));
    return t * t * (T(3) - T(2) * t);
  };
  // 1 + 0.036 cos(4 th + 1.73 seed) + 0.026 cos(6 th - 2.31 seed) [the lobe],
  // plus 0.018 sin(8 th + 0.89 seed) [the full azimuthal factor], from the
  // unit direction (c, s) of th
  auto azim_of = [](T c, T s, const DropPhase &ph, T &lobe) {
    const T c2 = c * c - s * s, s2 = T(2) * c * s;
    const T c4 = c2 * c2 - s2 * s2, s4 = T(2) * c2 * s2;
    const T c6 = c4 * c2 - s4 * s2, s6 = s4 * c2 + c4 * s2;
    const T c8 = c4 * c4 - s4 * s4, s8 = T(2) * c4 * s4;
    lobe = T(1) + T(0.036) * (c4 * (T)ph.c1 - s4 * (T)ph.s1) +
           T(0.026) * (c6 * (T)ph.c2 - s6 * (T)ph.s2);
    return lobe + T(0.018) * (s8 * (T)ph.c3 + c8 * (T)ph.s3);
  };
  // (cos, sin) of atan2(b, a); atan2(0, 0) = 0
  auto unit_of = [](T a, T b, T r, T &c, T &s) {
    if (r > T(0)) {
      c = a / r;
      s = b / r;
    } else {
      c = T(1);
      s = T(0);
    }
  };
  auto radial_of = [&](T rr, T slug) {
    const T sl = (T(0.14) + T(1) * fmin(rr * rr, T(1.20))) * (T(1) - ss(T(0.78), T(1.15), rr));
    const T st = (T(0.68) + T(0.32) * fmin(rr * rr, T(1))) * (T(1) - ss(T(0.74), T(1.10), rr));
    return mixv(st, sl, slug);
  };
  T active = T(0);
  // below surf - 0.058 both neck masks are 0, and `active` is read only on the
  // top two rows: skip the drops there
  const bool necks = fy_y > surf - T(0.058) || j >= d.ny - 2;
  for (int q = 0; necks && q < WV_MAX_DROPS; ++q) {
    const WaveDrop s = drops[q];
    if (!(s.seed > 0.f && s.drain > 0.0005f)) continue;
    const DropPhase ph = phs[q];
    const T dx0 = (T)s.x, dz0 = (T)s.z, drain = (T)s.drain;
    const T ex = cx - dx0, ez = cz - dz0;  // (the y-face shares the cell's x and z)
    const T r = sqrt(ex * ex + ez * ez);
    T cth, sth, lobe;
    unit_of(ex, ez, r, cth, sth);
    const T azim = azim_of(cth, sth, ph, lobe);
    const T slug = ss(T(0.14), T(0.76), drain);
    const T rbase = clampv((T)s.r * mixv(T(0.80), T(1.08), slug) + mixv(T(0.012), T(0.030), slug),
                           T(0.018), T(0.088));
    const T rr = r / fmax(rbase * azim, T(1e-4));
    const T radial = radial_of(rr, slug);
    const T vmask = ss(surf - T(0.058), surf - T(0.014), fy_y);
    const T shear = mixv(T(1), T(0.36), ss(surf - T(0.024), surf - T(0.002), fy_y));
    const T mask = radial * vmask;
    active = fmax(active, radial);
    if (mask > T(0.001)) {
      const T impact = clampv(-(T)s.vy, T(0.10), T(0.65));
      const T slug_drive = (T)ph.slug_drive;
      const T stem_drive = clampv(T(0.045) + drain * T(0.22), T(0), T(0.20)) * ss(T(0), T(0.05), drain);
      const T flux = mixv(stem_drive, slug_drive, slug);
      const T target = -clampv(T(0.14) + T(0.64) * impact, T(0.15), T(0.39)) * flux * lobe * shear;
      const T blend = clampv(dt * mixv(T(14), T(22), slug) * mask, T(0), T(1));
      vy = mixv(vy, target, blend);
    }
    const T vmask_xz = ss(surf - T(0.058), surf - T(0.014), cy);
    if (vmask_xz > T(0.001)) {
      const T flare = mixv(T(0.104) * (T(0.35) + T(0.65) * slug), T(-0.020),
                           ss(surf - T(0.066), surf - T(0.012), cy)) * clampv(drain, T(0), T(1.3));
      // x-face at (xf_x, cz), z-face at (cx, zf_z)
      const T fxs[2] = {xf_x - dx0, cx - dx0}, fzs[2] = {ez, zf_z - dz0};
#pragma unroll
      for (int f = 0; f < 2; ++f) {
        const T ax = fxs[f], az = fzs[f];
        const T rf = sqrt(ax * ax + az * az);
        if (rf > T(1e-4) && rf < rbase * T(1.34)) {
          T cf, sf, lobe_f;
          unit_of(ax, az, rf, cf, sf);
          const T rrf = rf / fmax(rbase * azim_of(cf, sf, ph, lobe_f), T(1e-4));
          if (rrf < T(1.22)) {
            const T mf = radial_of(rrf, slug) * vmask_xz;
            const T bf = clampv(dt * mixv(T(14), T(22), slug) * mf, T(0), T(1));
            if (f == 0) vx = mixv(vx, (ax / rf) * flare, bf * T(0.65));
            else vz = mixv(vz, (az / rf) * flare, bf * T(0.65));
          }
        }
      }
    }
  }
  if (j >= d.ny - 3) {
    const T fade = T(j - (d.ny - 3) + 1) / T(3);
    const float eps = 0.015f, E = P.ext;
    auto uvw = [&](float x) { return clampv((x + E) / (2.f * E), 0.002f, 0.998f); };
    const float4 wc = wv_sample(wave, P.n, uvw((float)cx), uvw((float)cz));
    const float wxp = wv_sample(wave, P.n, uvw((float)cx + eps), uvw((float)cz)).w;
    const float wxn = wv_sample(wave, P.n, uvw((float)cx - eps), uvw((float)cz)).w;
    const float wzp = wv_sample(wave, P.n, uvw((float)cx), uvw((float)cz + eps)).w;
    const float wzn = wv_sample(wave, P.n, uvw((float)cx), uvw((float)cz - eps)).w;
    const T cmc = T(1) - ss(T(0.45), T(0.88), (T)wc.w);
    const T gx = (T)(wxp - wxn) / (T(2) * (T)eps), gz = (T)(wzp - wzn) / (T(2) * (T)eps);
    const T m = (T)P.marangoni;
    vx = vx + (-m * T(0.026) * cmc * gx * fade) * dt;
    vz = vz + (-m * T(0.026) * cmc * gz * fade) * dt;
    if (j >= d.ny - 2 && active < T(0.02)) {
      const T calm = T(1) - ss(T(0.020), T(0.14), fabs((T)wc.y));
      const T layer = (j == d.ny - 1) ? T(0.72) : T(0.26);
      vy = mixv(vy, clampv((T)wc.y * T(0.35), T(-0.045), T(0.045)), calm * layer);
    }
  }
  const T rc = sqrt(cx * cx + cz * cz);
  const T fade = (T(1) - ss(hx * T(0.68), hx * T(0.94), rc)) * ss(y0 + T(0.004), y0 + T(0.052), cy);
  const T sponge = mixv(T(0.84), T(1), fade);
  vx = vx * sponge;
  vy = vy * sponge;
  vz = vz * sponge;
}

// The top-row stagnation gate of FLUID_MAC_TO_CENTER (fluid.ts:944-959): the
// factor on a downward centre velocity in the top two rows,
// mix(mix(0.45, 1, smoothstep(0.002, 0.018, surface - y)), 1, activeNeck),
// activeNeck the largest of each draining drop's (seed > 0, drain > 0.008)
// horizontal footprint times smoothstep(0.008, 0.065, drain).
template <class T>
__device__ inline T top_stagnation_gate(const Dims &d, int i, int j, int k, const WaveParams &P,
                                        const WaveDrop *__restrict__ drops) {
  const T hx = (T)P.half_xz, depth = (T)P.depth, surf = (T)P.surface_y, y0 = surf - depth;
  const T px = -hx + (i + T(0.5)) / (T)d.nx * (T(2) * hx);
  const T py = y0 + (j + T(0.5)) / (T)d.ny * depth;
  const T pz = -hx + (k + T(0.5)) / (T)d.nz * (T(2) * hx);
  auto ss = [](T e0, T e1, T x) {
    T t = clampv((x - e0) / (e1 - e0), T(0), T(1));
    return t * t * (T(3) - T(2) * t);
  };
  T act = T(0);
  for (int q = 0; q < WV_MAX_DROPS; ++q) {
    const WaveDrop s = drops[q];
    if (!(s.seed > 0.f && s.drain > 0.008f)) continue;
    const T ra = clampv((T)s.r * T(1.15) + T(0.032), T(0.046), T(0.092));
    const T ex = px - (T)s.x, ez = pz - (T)s.z;
    const T rd = sqrt(ex * ex + ez * ez);
    act = fmax(act, (T(1) - ss(ra * T(0.78), ra * T(1.20), rd)) * ss(T(0.008), T(0.065), (T)s.drain));
  }
  const T stag = mixv(T(0.45), T(1), ss(T(0.002), T(0.018), fmax(T(0), surf - py)));
  return mixv(stag, T(1), act);
}

// the scene's per-frame uniforms that the Casson pass reads from device memory
// (so a captured step keeps them current): simTime (cameraPosAndTime.w),
// schlierenStr (particleParams.x), vorticityConf (fluidParams.x), and the
// parcels' hemolysisRate (lensAndRheology.w)
struct SceneClock {
  float sim_time, schlieren, vort_conf, hemolysis;
};

template <class T> struct q4_of;
template <> struct q4_of<float> { typedef float4 type; };
template <> struct q4_of<double> { typedef double4 type; };

// FLUID_CURL (src/shaders/fluid.ts:331-369): the curl of the cell-centred
// velocity by central differences of the neighbouring cells (grid units: no
// 1/h), neighbour indices clamped to the grid, and its length in .w.
template <class T>
__global__ void __launch_bounds__(256)
k_curl(Dims d, const typename cvec_of<T>::type *__restrict__ cvel,
       typename q4_of<T>::type *__restrict__ curl) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= d.nx || j >= d.ny) return;
  auto at = [&](int a, int b, int e) { return cvel[idx3(d, a, b, e)]; };
  const auto vR = at(min(i + 1, d.nx - 1), j, k), vL = at(max(i - 1, 0), j, k);
  const auto vU = at(i, min(j + 1, d.ny - 1), k), vD = at(i, max(j - 1, 0), k);
  const auto vF = at(i, j, min(k + 1, d.nz - 1)), vB = at(i, j, max(k - 1, 0));
  typename q4_of<T>::type o;
  o.x = T(0.5) * ((vU.z - vD.z) - (vF.y - vB.y));
  o.y = T(0.5) * ((vF.x - vB.x) - (vR.z - vL.z));
  o.z = T(0.5) * ((vR.y - vL.y) - (vU.x - vD.x));
  o.w = sqrt(o.x * o.x + o.y * o.y + o.z * o.z);
  curl[idx3(d, i, j, k)] = o;
}

// evalVortConf (fluid.ts:394-423): the confinement acceleration at cell
// (i, j, k): zero within two cells of the side walls and the floor and three
// of the lid, and where |omega| <= 0.18; else
// scale * gate(|omega|) * (N x omega), N the unit central-difference gradient
// of |omega| (zero when that gradient is below 1e-5).
template <class T>
__device__ inline void vort_conf_at(const Dims &d, const typename q4_of<T>::type *__restrict__ curl,
                                    int i, int j, int k, T scale, T f[3]) {
  f[0] = f[1] = f[2] = T(0);
  if (i <= 1 || i >= d.nx - 2 || j <= 1 || j >= d.ny - 3 || k <= 1 || k >= d.nz - 2) return;
  const auto om = curl[idx3(d, i, j, k)];
  if (om.w <= T(0.18)) return;
  const T t = clampv((om.w - T(0.18)) / T(0.95 - 0.18), T(0), T(1));
  const T gate = t * t * (T(3) - T(2) * t) / (T(1) + T(0.28) * om.w);
  const T ex = T(0.5) * (curl[idx3(d, i + 1, j, k)].w - curl[idx3(d, i - 1, j, k)].w);
  const T ey = T(0.5) * (curl[idx3(d, i, j + 1, k)].w - curl[idx3(d, i, j - 1, k)].w);
  const T ez = T(0.5) * (curl[idx3(d, i, j, k + 1)].w - curl[idx3(d, i, j, k - 1)].w);
  const T el = sqrt(ex * ex + ey * ey + ez * ez);
  if (el <= T(1e-5)) return;
  const T g = scale * gate / el;
  f[0] = g * (ey * om.z - ez * om.y);
  f[1] = g * (ez * om.x - ex * om.z);
  f[2] = g * (ex * om.y - ey * om.x);
}

// The Casson pass's scene terms (FLUID_CASSON_FORCES, fluid.ts:467-509),
// added in place to the faces of cell (i, j, k) after the viscous update and
// buoyancy: the vorticity confinement averaged onto each face with its
// +1 neighbour (:467-472), the deep-column solenoidal unfurling forcing
// (:488-503; gate from the dye, a single species here, so the lysed channel
// is 0), the +-1.8 clamp and the wall faces zeroed (:505-507).  World
// positions as in advect_scene_terms; dt clamped to [0, 0.025].
template <class T>
__global__ void __launch_bounds__(256)
k_casson_scene(Dims d, WaveParams P, const SceneClock *__restrict__ clk,
               const typename q4_of<T>::type *__restrict__ curl, const T *__restrict__ dye,
               const T *__restrict__ lys,
               T *__restrict__ u, T *__restrict__ v, T *__restrict__ w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= d.nx || j >= d.ny) return;
  const size_t c = idx3(d, i, j, k);
  const T dt = (T)P.dt, scale = (T)clk->vort_conf;
  T f0[3], fx[3], fy[3], fz[3];
  vort_conf_at<T>(d, curl, i, j, k, scale, f0);
  vort_conf_at<T>(d, curl, min(i + 1, d.nx - 1), j, k, scale, fx);
  vort_conf_at<T>(d, curl, i, min(j + 1, d.ny - 1), k, scale, fy);
  vort_conf_at<T>(d, curl, i, j, min(k + 1, d.nz - 1), scale, fz);
  T mx = u[c] + T(0.5) * (f0[0] + fx[0]) * dt;
  T my = v[c] + T(0.5) * (f0[1] + fy[1]) * dt;
  T mz = w[c] + T(0.5) * (f0[2] + fz[2]) * dt;
  const T hx = (T)P.half_xz, depth = (T)P.depth, y0 = (T)P.surface_y - depth;
  const T px = -hx + (i + T(0.5)) / (T)d.nx * (T(2) * hx);
  const T py = y0 + (j + T(0.5)) / (T)d.ny * depth;
  const T pz = -hx + (k + T(0.5)) / (T)d.nz * (T(2) * hx);
  const T below = (T)P.surface_y - py;
  auto ss = [](T e0, T e1, T x) {
    T t = clampv((x - e0) / (e1 - e0), T(0), T(1));
    return t * t * (T(3) - T(2) * t);
  };
  // cRbc + 0.45 concCell.g (fluid.ts:491): the lysed term with the
  // 4-channel concentration only
  const T crbc = clampv(dye[c], T(0), T(1.5)) + (lys ? T(0.45) * lys[c] : T(0));
  const T gate = ss(T(0.11), T(0.36), below) * ss(T(0.0008), T(0.015), crbc);
  if (gate > T(0.001)) {
    const T ts = (T)clk->sim_time * T(0.36) + (T)clk->schlieren * T(1.7);
    const T q1 = T(6.4) * px + T(4.0) * py + T(5.2) * pz + ts;
    const T q2 = T(-5.6) * px + T(4.8) * py + T(7.2) * pz - ts * T(0.77);
    const T q3 = T(9.0) * px - T(6.0) * py - T(6.4) * pz + ts * T(1.13);
    const T c1 = cos(q1), c2 = cos(q2), s3 = T(0.42) * sin(q3);
    const T sx = T(-6.25) * c1 + T(-6.30) * c2 + T(6.40) * s3;
    const T sy = T(2.85) * c1 + T(-1.26) * c2 + T(-0.96) * s3;
    const T sz = T(5.50) * c1 + T(-4.06) * c2 + T(9.90) * s3;
    const T a = T(0.0028) * gate * dt;
    mx += sx * a;
    my += sy * ss(T(0.16), T(0.42), below) * a;
    mz += sz * a;
  }
  mx = clampv(mx, T(-1.8), T(1.8));
  my = clampv(my, T(-1.8), T(1.8));
  mz = clampv(mz, T(-1.8), T(1.8));
  if (i >= d.nx - 1) mx = T(0);
  if (k >= d.nz - 1) mz = T(0);
  u[c] = mx;
  v[c] = my;
  w[c] = mz;
}

}  // namespace mac3d
