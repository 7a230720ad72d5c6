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

// Cell-centred velocity is held as one 3-vector per cell (x, y, z; 12 B in
// FP32, one dwordx3 load) so a trilinear sample of all three components is
// eight vector loads.
template <class T> struct cvec_of;
template <> struct cvec_of<float> { typedef float3 type; };
template <> struct cvec_of<double> { typedef double3 type; };
template <class T> struct vec2_of;
template <> struct vec2_of<float> { typedef float2 type; };
template <> struct vec2_of<double> { typedef double2 type; };

template <class T>
__device__ inline void sample_vec(const typename cvec_of<T>::type *__restrict__ f, const Dims &d,
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

// ---------------------------------------------------------------------------
// Sample points in texel space kept as (integer texel, small offset) per
// axis: t = b + o with t = uvw * N - 1/2.  The reference's normalised
// coordinates hold a point to about N * 6e-8 texels in FP32 (1e-5 texel at
// N = 192); an integer base plus an offset of a few texels holds it to 1e-7
// texel.  Clamps to [lo, hi] texels and samples use the same clamp-to-edge
// rule as sample_cell.
template <class T> struct TPos {
  int b[3];
  T o[3];
};

// Every sample point of a cell's backtrace keeps the cell's own texel index
// as its base, so a clamp to [lo, hi] texels acts on the offset alone,
// against the bounds taken relative to that index (tpos_bounds).  Where a
// clamp takes effect the bound lies within a few texels of the index, and
// their difference is exact (Sterbenz), so the clamped point is the bound
// itself.  One median-of-three per axis in FP32 (a NaN offset clamps to a
// bound: the advection carries a non-finite trace on by itself, trace_nan),
// two compares and two selects in FP64.
template <class T> struct TBounds {
  T lo[3], hi[3];
};
template <class T>
__device__ inline TBounds<T> tpos_bounds(const T lo[3], const T hi[3], int i, int j, int k) {
  TBounds<T> r;
  const int b[3] = {i, j, k};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    r.lo[a] = lo[a] - (T)b[a];
    r.hi[a] = hi[a] - (T)b[a];
  }
  return r;
}
template <class T>
__device__ inline void tpos_clamp(TPos<T> &p, const TBounds<T> &r) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if constexpr (sizeof(T) == 4) p.o[a] = __builtin_amdgcn_fmed3f(p.o[a], r.lo[a], r.hi[a]);
    else p.o[a] = clampv(p.o[a], r.lo[a], r.hi[a]);
  }
}
// 0, or NaN when a velocity of a backtrace is not finite: added to the
// advected value, so a non-finite velocity yields a non-finite result even
// where its trace was clamped back into the grid
template <class T> __device__ inline T trace_nan(const T a[3], const T b[3]) {
  return T(0) * (((a[0] + a[1]) + (a[2] + b[0])) + (b[1] + b[2]));
}

template <class T, class I = size_t> struct TriIdx {
  I r00, r10, r01, r11;
  int i0, i1;
  T fx, fy, fz;
};

// corner rows and weights of a trilinear sample at p; I: the index type (int
// when the field is addressed with 32-bit element indices)
template <class T, class I = size_t>
__device__ inline TriIdx<T, I> tpos_index(const Dims &d, const TPos<T> &p) {
  TriIdx<T, I> t;
  T fl0 = floor(p.o[0]), fl1 = floor(p.o[1]), fl2 = floor(p.o[2]);
  int ib = p.b[0] + (int)fl0, jb = p.b[1] + (int)fl1, kb = p.b[2] + (int)fl2;
  t.fx = p.o[0] - fl0;
  t.fy = p.o[1] - fl1;
  t.fz = p.o[2] - fl2;
  t.i0 = clampi(ib, 0, d.nx - 1);
  t.i1 = clampi(ib + 1, 0, d.nx - 1);
  int j0 = clampi(jb, 0, d.ny - 1), j1 = clampi(jb + 1, 0, d.ny - 1);
  int k0 = clampi(kb, 0, d.nz - 1), k1 = clampi(kb + 1, 0, d.nz - 1);
  t.r00 = ((I)k0 * d.ny + j0) * d.nx;
  t.r10 = ((I)k0 * d.ny + j1) * d.nx;
  t.r01 = ((I)k1 * d.ny + j0) * d.nx;
  t.r11 = ((I)k1 * d.ny + j1) * d.nx;
  return t;
}

// Element readers for the samplers: PtrRead takes 64-bit element indices
// into a plain pointer; BufRead (below) takes 32-bit ones into a buffer
// resource, so a gather needs no 64-bit index or address arithmetic.
template <class V> struct PtrRead {
  typedef size_t index;
  static constexpr bool bytes = false;  // offsets count elements
  static constexpr int unit = 1;
  const V *__restrict__ p;
  __device__ V operator()(size_t i) const { return p[i]; }
  __device__ V ld(size_t i) const { return p[i]; }
  // offset of cell (i, j, k)
  __device__ size_t cell(const Dims &d, int i, int j, int k) const { return idx3(d, i, j, k); }
};

// Corner byte offsets of a trilinear sample at p for a field of ES-byte
// elements read through a buffer resource: o + {0, dx} + {0, dy} + {0, dz}.
// Every sample point is clamped inside [-1/2, N - 1/2) texels on each axis,
// so its lower corner index lies in [-1, N - 1] and clamp-to-edge only merges
// the two corners at -1 and at N - 1 (d = 0 there).  The offsets come from
// 24-bit multiplies (full rate; 32-bit ones are quarter rate), which need a
// plane of the field under 2^24 bytes (the host checks).
template <class T> struct TriOff {
  int o, dx, dy, dz;
  T fx, fy, fz;
};
template <class T, int ES>
__device__ inline TriOff<T> tpos_off(const Dims &d, const TPos<T> &p) {
  TriOff<T> t;
  const T fl0 = floor(p.o[0]), fl1 = floor(p.o[1]), fl2 = floor(p.o[2]);
  const int ib = p.b[0] + (int)fl0, jb = p.b[1] + (int)fl1, kb = p.b[2] + (int)fl2;
  t.fx = p.o[0] - fl0;
  t.fy = p.o[1] - fl1;
  t.fz = p.o[2] - fl2;
  const int sy = d.nx * ES, sz = d.nx * d.ny * ES;
  t.dx = (unsigned)ib < (unsigned)(d.nx - 1) ? ES : 0;
  t.dy = (unsigned)jb < (unsigned)(d.ny - 1) ? sy : 0;
  t.dz = (unsigned)kb < (unsigned)(d.nz - 1) ? sz : 0;
  t.o = __umul24((unsigned)max(kb, 0), (unsigned)sz) + __umul24((unsigned)max(jb, 0), (unsigned)sy) +
        __umul24((unsigned)max(ib, 0), (unsigned)ES);
  return t;
}

template <class T, class R>
__device__ inline T sample_tp_r(const R &f, const Dims &d, const TPos<T> &p) {
  if constexpr (R::bytes) {
    const auto t = tpos_off<T, R::unit>(d, p);
    const int o1 = t.o + t.dy, o2 = t.o + t.dz, o3 = o2 + t.dy;
    T x00 = mixv(f.ld(t.o), f.ld(t.o + t.dx), t.fx);
    T x10 = mixv(f.ld(o1), f.ld(o1 + t.dx), t.fx);
    T x01 = mixv(f.ld(o2), f.ld(o2 + t.dx), t.fx);
    T x11 = mixv(f.ld(o3), f.ld(o3 + t.dx), t.fx);
    return mixv(mixv(x00, x10, t.fy), mixv(x01, x11, t.fy), t.fz);
  }
  auto t = tpos_index<T, typename R::index>(d, p);
  T x00 = mixv(f(t.r00 + t.i0), f(t.r00 + t.i1), t.fx);
  T x10 = mixv(f(t.r10 + t.i0), f(t.r10 + t.i1), t.fx);
  T x01 = mixv(f(t.r01 + t.i0), f(t.r01 + t.i1), t.fx);
  T x11 = mixv(f(t.r11 + t.i0), f(t.r11 + t.i1), t.fx);
  return mixv(mixv(x00, x10, t.fy), mixv(x01, x11, t.fy), t.fz);
}
template <class T>
__device__ inline T sample_tp(const T *__restrict__ f, const Dims &d, const TPos<T> &p) {
  return sample_tp_r<T>(PtrRead<T>{f}, d, p);
}

// sample_tp that also gives the least and greatest of its 8 texels: the
// bounds of the MacCormack limiter around the same point (the limiter takes
// the same clamped corners, in this fmin/fmax order)
template <class T, class R>
__device__ inline T sample_tp_mm_r(const R &f, const Dims &d, const TPos<T> &p, T &mn, T &mx) {
  T s0, s1, s2, s3, s4, s5, s6, s7, fx, fy, fz;
  if constexpr (R::bytes) {
    const auto t = tpos_off<T, R::unit>(d, p);
    const int o1 = t.o + t.dy, o2 = t.o + t.dz, o3 = o2 + t.dy;
    s0 = f.ld(t.o); s1 = f.ld(t.o + t.dx); s2 = f.ld(o1); s3 = f.ld(o1 + t.dx);
    s4 = f.ld(o2); s5 = f.ld(o2 + t.dx); s6 = f.ld(o3); s7 = f.ld(o3 + t.dx);
    fx = t.fx; fy = t.fy; fz = t.fz;
  } else {
    auto t = tpos_index<T, typename R::index>(d, p);
    s0 = f(t.r00 + t.i0); s1 = f(t.r00 + t.i1); s2 = f(t.r10 + t.i0); s3 = f(t.r10 + t.i1);
    s4 = f(t.r01 + t.i0); s5 = f(t.r01 + t.i1); s6 = f(t.r11 + t.i0); s7 = f(t.r11 + t.i1);
    fx = t.fx; fy = t.fy; fz = t.fz;
  }
  mn = fmin(fmin(fmin(s0, s1), fmin(s2, s3)), fmin(fmin(s4, s5), fmin(s6, s7)));
  mx = fmax(fmax(fmax(s0, s1), fmax(s2, s3)), fmax(fmax(s4, s5), fmax(s6, s7)));
  T x00 = mixv(s0, s1, fx);
  T x10 = mixv(s2, s3, fx);
  T x01 = mixv(s4, s5, fx);
  T x11 = mixv(s6, s7, fx);
  return mixv(mixv(x00, x10, fy), mixv(x01, x11, fy), fz);
}

template <class T, class R>
__device__ inline void sample_vec_tp_r(const R &f, const Dims &d, const TPos<T> &p, T out[3]) {
  typedef typename cvec_of<T>::type V3;
  V3 a000, a100, a010, a110, a001, a101, a011, a111;
  T fx, fy, fz;
  if constexpr (R::bytes) {
    const auto t = tpos_off<T, R::unit>(d, p);
    const int o1 = t.o + t.dy, o2 = t.o + t.dz, o3 = o2 + t.dy;
    a000 = f.ld(t.o); a100 = f.ld(t.o + t.dx); a010 = f.ld(o1); a110 = f.ld(o1 + t.dx);
    a001 = f.ld(o2); a101 = f.ld(o2 + t.dx); a011 = f.ld(o3); a111 = f.ld(o3 + t.dx);
    fx = t.fx; fy = t.fy; fz = t.fz;
  } else {
    auto t = tpos_index<T, typename R::index>(d, p);
    a000 = f(t.r00 + t.i0); a100 = f(t.r00 + t.i1); a010 = f(t.r10 + t.i0); a110 = f(t.r10 + t.i1);
    a001 = f(t.r01 + t.i0); a101 = f(t.r01 + t.i1); a011 = f(t.r11 + t.i0); a111 = f(t.r11 + t.i1);
    fx = t.fx; fy = t.fy; fz = t.fz;
  }
#define MAC3D_TRI(C)                                                                    \
  mixv(mixv(mixv(a000.C, a100.C, fx), mixv(a010.C, a110.C, fx), fy),                    \
       mixv(mixv(a001.C, a101.C, fx), mixv(a011.C, a111.C, fx), fy), fz)
  out[0] = MAC3D_TRI(x);
  out[1] = MAC3D_TRI(y);
  out[2] = MAC3D_TRI(z);
#undef MAC3D_TRI
}
template <class T>
__device__ inline void sample_vec_tp(const typename cvec_of<T>::type *__restrict__ f,
                                     const Dims &d, const TPos<T> &p, T out[3]) {
  sample_vec_tp_r<T>(PtrRead<typename cvec_of<T>::type>{f}, d, p, out);
}

// Raw buffer resource over `bytes` bytes from `base` (stride 0; word 3 is the
// gfx9 default, 0x00020000).  A load whose byte offset lies outside
// [0, bytes) returns zeros: the buffer unit's range check stands in for a
// branch, so the loads issue unconditionally and the waits on them count
// exactly how many are still in flight.
typedef __amdgpu_buffer_rsrc_t BufRsrc;
__device__ inline BufRsrc buf_rsrc(const void *base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes,
                                           0x00020000);
}
template <class V> __device__ inline V buf_load(BufRsrc r, int off);
template <> __device__ inline float buf_load<float>(BufRsrc r, int off) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
template <> __device__ inline double buf_load<double>(BufRsrc r, int off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
template <> __device__ inline float2 buf_load<float2>(BufRsrc r, int off) {
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
template <> __device__ inline double2 buf_load<double2>(BufRsrc r, int off) {
  return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}

template <> __device__ inline float3 buf_load<float3>(BufRsrc r, int off) {
  auto v = __builtin_amdgcn_raw_buffer_load_b96(r, off, 0, 0);
  return make_float3(__int_as_float(v[0]), __int_as_float(v[1]), __int_as_float(v[2]));
}
template <> __device__ inline double3 buf_load<double3>(BufRsrc r, int off) {
  return make_double3(buf_load<double>(r, off), buf_load<double>(r, off + 8),
                      buf_load<double>(r, off + 16));
}
template <class V> struct BufRead {
  typedef int index;
  static constexpr bool bytes = true;  // offsets count bytes
  static constexpr int unit = (int)sizeof(V);
  BufRsrc r;
  __device__ V operator()(int i) const { return buf_load<V>(r, i * (int)sizeof(V)); }
  __device__ V ld(int off) const { return buf_load<V>(r, off); }
  // byte offset of cell (i, j, k) (24-bit multiplies: a plane is under 2^24 bytes)
  __device__ int cell(const Dims &d, int i, int j, int k) const {
    return __umul24((unsigned)k, (unsigned)(d.nx * d.ny * unit)) +
           __umul24((unsigned)j, (unsigned)(d.nx * unit)) + __umul24((unsigned)i, (unsigned)unit);
  }
};
// reader of n elements at p: a buffer resource when BUF (the host checks
// that the field is under 2^31 bytes and a plane under 2^24), else the plain
// pointer
template <bool BUF, class V> struct Reader { typedef PtrRead<V> type; };
template <class V> struct Reader<true, V> { typedef BufRead<V> type; };
template <bool BUF, class V>
__device__ inline typename Reader<BUF, V>::type make_reader(const V *p, size_t n) {
  if constexpr (BUF) return BufRead<V>{buf_rsrc(p, (uint32_t)(n * sizeof(V)))};
  else return PtrRead<V>{p};
}

// An (even, odd) x-pair of S cells as float2 (FP16 widens exactly), and its
// store from float2 (FP16 as store_as rounds it).
typedef half_t half2v __attribute__((ext_vector_type(2)));
template <class S> __device__ inline float2 buf_load_pair(BufRsrc r, int off);
template <> __device__ inline float2 buf_load_pair<float>(BufRsrc r, int off) {
  return buf_load<float2>(r, off);
}
template <> __device__ inline float2 buf_load_pair<half_t>(BufRsrc r, int off) {
  half2v h = __builtin_bit_cast(half2v, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
  return make_float2((float)h.x, (float)h.y);
}
__device__ inline void store_pair(float *p, float2 v) { *(float2 *)p = v; }
__device__ inline void store_pair(half_t *p, float2 v) {
  half2v h;
  h.x = store_as<half_t, float>(v.x);
  h.y = store_as<half_t, float>(v.y);
  *(half2v *)p = h;
}
// in-plane byte offset of a staged position outside the grid: with every
// range below 2^30 bytes, it stays negative (out of range) for any plane
// offset the marching kernels add to it
#define MAC3D_OFF_NONE (-(1 << 30))

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

// Blocks are numbered on a 1-D grid and remapped so that each of the 8 XCDs
// (blocks are dealt round-robin to XCDs) gets a contiguous run of tiles: the
// halo planes and rows a tile shares with its neighbours then sit in the
// same XCD's L2.
__device__ inline int xcd_remap(int b, int nblocks) {
  if (nblocks % 8) return b;
  return (b % 8) * (nblocks / 8) + b / 8;
}

}  // namespace mac3d
