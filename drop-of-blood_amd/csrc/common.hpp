// common.hpp — shared device helpers of the MAC step kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mac3d {

typedef _Float16 half_t;

struct Dims {
  int nx, ny, nz;
  __host__ __device__ size_t cells() const { return (size_t)nx * ny * nz; }
};

__host__ __device__ inline size_t idx3(const Dims &d, int i, int j, int k) {
  return ((size_t)k * d.ny + j) * (size_t)d.nx + i;
}

// compute type of a storage type: FP16 and FP32 storage compute in FP32
template <class S> struct compute_of { typedef float type; };
template <> struct compute_of<double> { typedef double type; };

template <class T> __device__ inline T clampv(T x, T lo, T hi) {
  return x < lo ? lo : (x > hi ? hi : x);
}
__device__ inline int clampi(int x, int lo, int hi) {
  return x < lo ? lo : (x > hi ? hi : x);
}
template <class T> __device__ inline T mixv(T a, T b, T t) {
  return a * (T(1) - t) + b * t;
}

// store with saturation for FP16 (keeps an out-of-range correction finite)
template <class S, class C> __device__ inline S store_as(C v) { return (S)v; }
template <> __device__ inline half_t store_as<half_t, float>(float v) {
  const float m = 65504.0f;
  return (half_t)(v > m ? m : (v < -m ? -m : v));
}

// Trilinear sample of a cell-centred field at normalised coordinates with
// the clamp-to-edge texel rule of a WebGPU linear sampler:
// t = uvw*N - 0.5, floor / frac, indices clamped to [0, N-1].
template <class T>
__device__ inline T sample_cell(const T *__restrict__ f, const Dims &d, T ux,
                                T uy, T uz) {
  T tx = ux * (T)d.nx - T(0.5), ty = uy * (T)d.ny - T(0.5),
    tz = uz * (T)d.nz - T(0.5);
  T bx = floor(tx), by = floor(ty), bz = floor(tz);
  T fx = tx - bx, fy = ty - by, fz = tz - bz;
  int ib = (int)bx, jb = (int)by, kb = (int)bz;
  int i0 = clampi(ib, 0, d.nx - 1), i1 = clampi(ib + 1, 0, d.nx - 1);
  int j0 = clampi(jb, 0, d.ny - 1), j1 = clampi(jb + 1, 0, d.ny - 1);
  int k0 = clampi(kb, 0, d.nz - 1), k1 = clampi(kb + 1, 0, d.nz - 1);
  size_t r00 = ((size_t)k0 * d.ny + j0) * d.nx, r10 = ((size_t)k0 * d.ny + j1) * d.nx;
  size_t r01 = ((size_t)k1 * d.ny + j0) * d.nx, r11 = ((size_t)k1 * d.ny + j1) * d.nx;
  T x00 = mixv(f[r00 + i0], f[r00 + i1], fx);
  T x10 = mixv(f[r10 + i0], f[r10 + i1], fx);
  T x01 = mixv(f[r01 + i0], f[r01 + i1], fx);
  T x11 = mixv(f[r11 + i0], f[r11 + i1], fx);
  return mixv(mixv(x00, x10, fy), mixv(x01, x11, fy), fz);
}

// Cell-centred velocity is held as one 4-vector per cell (x, y, z, unused) so
// a trilinear sample of all three components is eight vector loads.
template <class T> struct vec4_of;
template <> struct vec4_of<float> { typedef float4 type; };
template <> struct vec4_of<double> { typedef double4 type; };

template <class T>
__device__ inline void sample_vec(const typename vec4_of<T>::type *__restrict__ f, const Dims &d,
                                  T ux, T uy, T uz, T out[3]) {
  T tx = ux * (T)d.nx - T(0.5), ty = uy * (T)d.ny - T(0.5), tz = uz * (T)d.nz - T(0.5);
  T bx = floor(tx), by = floor(ty), bz = floor(tz);
  T fx = tx - bx, fy = ty - by, fz = tz - bz;
  int ib = (int)bx, jb = (int)by, kb = (int)bz;
  int i0 = clampi(ib, 0, d.nx - 1), i1 = clampi(ib + 1, 0, d.nx - 1);
  int j0 = clampi(jb, 0, d.ny - 1), j1 = clampi(jb + 1, 0, d.ny - 1);
  int k0 = clampi(kb, 0, d.nz - 1), k1 = clampi(kb + 1, 0, d.nz - 1);
  size_t r00 = ((size_t)k0 * d.ny + j0) * d.nx, r10 = ((size_t)k0 * d.ny + j1) * d.nx;
  size_t r01 = ((size_t)k1 * d.ny + j0) * d.nx, r11 = ((size_t)k1 * d.ny + j1) * d.nx;
  auto a000 = f[r00 + i0], a100 = f[r00 + i1], a010 = f[r10 + i0], a110 = f[r10 + i1];
  auto a001 = f[r01 + i0], a101 = f[r01 + i1], a011 = f[r11 + i0], a111 = f[r11 + i1];
#define MAC3D_TRI(C)                                                            \
  mixv(mixv(mixv(a000.C, a100.C, fx), mixv(a010.C, a110.C, fx), fy),            \
       mixv(mixv(a001.C, a101.C, fx), mixv(a011.C, a111.C, fx), fy), fz)
  out[0] = MAC3D_TRI(x);
  out[1] = MAC3D_TRI(y);
  out[2] = MAC3D_TRI(z);
#undef MAC3D_TRI
}

// The same trilinear sample of three fields at one point (the cell-centred
// velocity components): index and weight arithmetic shared.
template <class T>
__device__ inline void sample_cell3(const T *__restrict__ f0, const T *__restrict__ f1,
                                    const T *__restrict__ f2, const Dims &d, T ux, T uy,
                                    T uz, T out[3]) {
  T tx = ux * (T)d.nx - T(0.5), ty = uy * (T)d.ny - T(0.5), tz = uz * (T)d.nz - T(0.5);
  T bx = floor(tx), by = floor(ty), bz = floor(tz);
  T fx = tx - bx, fy = ty - by, fz = tz - bz;
  int ib = (int)bx, jb = (int)by, kb = (int)bz;
  int i0 = clampi(ib, 0, d.nx - 1), i1 = clampi(ib + 1, 0, d.nx - 1);
  int j0 = clampi(jb, 0, d.ny - 1), j1 = clampi(jb + 1, 0, d.ny - 1);
  int k0 = clampi(kb, 0, d.nz - 1), k1 = clampi(kb + 1, 0, d.nz - 1);
  size_t r00 = ((size_t)k0 * d.ny + j0) * d.nx, r10 = ((size_t)k0 * d.ny + j1) * d.nx;
  size_t r01 = ((size_t)k1 * d.ny + j0) * d.nx, r11 = ((size_t)k1 * d.ny + j1) * d.nx;
  const T *fs[3] = {f0, f1, f2};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const T *f = fs[a];
    T x00 = mixv(f[r00 + i0], f[r00 + i1], fx);
    T x10 = mixv(f[r10 + i0], f[r10 + i1], fx);
    T x01 = mixv(f[r01 + i0], f[r01 + i1], fx);
    T x11 = mixv(f[r11 + i0], f[r11 + i1], fx);
    out[a] = mixv(mixv(x00, x10, fy), mixv(x01, x11, fy), fz);
  }
}

// wavefront (64 lanes) reductions
template <class T> __device__ inline T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
template <class T> __device__ inline T wave_max(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    T w = __shfl_xor(v, o, 64);
    v = w > v ? w : v;
  }
  return v;
}

}  // namespace mac3d
