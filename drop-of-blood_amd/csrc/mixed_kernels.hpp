// mixed_kernels.hpp — FP64 iterative-refinement kernels of the mixed mode.
//
// The pressure lives in FP64 with an FP64 right-hand side.  Each refinement
// iteration computes r = b - L0 p in FP64, hands r to an FP32 V-cycle as its
// right-hand side (scaled by 1/max|r| so the FP16 coarse levels stay in
// range) and adds the scaled correction back in FP64.
#pragma once

#include "common.hpp"
#include "mg_kernels.hpp"

namespace mac3d {

// r = b - L0 q in FP64, where q is a pointwise function of stored fields
// (Src::value): q is written to qout, r to r (FP32), and each block writes one
// (max|r|, sum r^2, sum b^2) triple.  Used with
//  - the warm-start extrapolation q = c0 p_n + c1 p_{n-1} + c2 p_{n-2}, and
//  - the refinement update q = p + max|r| e,
// so the new pressure is formed, stored and tested in one pass (qout must
// not alias the inputs: neighbours read the old values).  The operator is
// the level-0 stencil of evalMgStencil (src/shaders/fluid.ts:561-603): the
// lid adds 1 to the diagonal whether or not j is the top row, neighbours
// summed x-, x+, y-, y+, z-, z+.
//
// A 512-thread block owns a 64 x 8 column tile and a z-chunk [k0, k1).  Each
// plane's q is formed once per cell (tile plus a one-cell ring, staged in a
// three-slot LDS ring), the thread's own column keeps q at k-1, k, k+1 in
// registers, and one triple per chunk keeps the fold to a few hundred.
// Inputs are read with buffer loads whose ranges cover the chunk's planes
// k0-1 .. k1: positions outside the grid load as 0, which the Neumann sum
// omits anyway.
#define R64_TX 64
#define R64_TY 8
#define R64_T (R64_TX * R64_TY)
#define R64_EX (R64_TX + 2)
#define R64_EY (R64_TY + 2)
#define R64_PL (R64_EX * R64_EY)
#define R64_LPT ((R64_PL + R64_T - 1) / R64_T)

// value sources: the warm-start extrapolation and the refinement update
struct R64Extrap {
  BufRsrc r0, r1, r2;
  double c0, c1, c2;
  struct Reg { double a, b, c; };
  __device__ Reg load(int off8, int off4) const {
    (void)off4;
    return Reg{buf_load<double>(r0, off8), buf_load<double>(r1, off8), buf_load<double>(r2, off8)};
  }
  __device__ double value(const Reg &g) const { return c0 * g.a + c1 * g.b + c2 * g.c; }
  static constexpr bool wide = false;
};
struct R64Update {
  BufRsrc rp, re;
  double s;
  struct Reg { double p; float e; };
  __device__ Reg load(int off8, int off4) const {
    return Reg{buf_load<double>(rp, off8), buf_load<float>(re, off4)};
  }
  __device__ double value(const Reg &g) const { return g.p + s * (double)g.e; }
  static constexpr bool wide = true;
};

template <class Src>
__device__ inline void residual64_zm(const Dims &d, const Src &src, int kbase, int pb8, int pb4,
                                     BufRsrc rbv, double *__restrict__ qout,
                                     float *__restrict__ r, double *__restrict__ part, int k0,
                                     int k1, int x0, int y0) {
  __shared__ double sq[3][R64_EY][R64_EX];
  __shared__ double sred[3][R64_T / 64];
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * R64_TX + tx;
  const size_t plane = (size_t)nx * ny;
  int so8[R64_LPT], so4[R64_LPT], sr[R64_LPT], sc[R64_LPT];
#pragma unroll
  for (int q = 0; q < R64_LPT; ++q) {
    int v = tid + q * R64_T;
    sr[q] = v / R64_EX;
    sc[q] = v % R64_EX;
    int gx = x0 - 1 + sc[q], gy = y0 - 1 + sr[q];
    bool ok = v < R64_PL && gx >= 0 && gx < nx && gy >= 0 && gy < ny;
    so8[q] = ok ? (gy * nx + gx) * 8 : MAC3D_OFF_NONE;
    so4[q] = ok ? (gy * nx + gx) * 4 : MAC3D_OFF_NONE;
    if (v >= R64_PL) sr[q] = sc[q] = 0;
  }
  const int gx = x0 + tx, gy = y0 + ty;
  const bool inside = gx < nx && gy < ny;
  const int bo = inside ? (gy * nx + gx) * 8 : MAC3D_OFF_NONE;
  // two register sets of staged values: plane k + 1 is staged from one while
  // plane k + 2's loads are in flight in the other
  typename Src::Reg pre[2][R64_LPT];
  auto fetch = [&](int k, int h) {
#pragma unroll
    for (int q = 0; q < R64_LPT; ++q)
      pre[h][q] = src.load(so8[q] + (k - kbase) * pb8, so4[q] + (k - kbase) * pb4);
  };
  auto stage = [&](int slot, int h) {
#pragma unroll
    for (int q = 0; q < R64_LPT; ++q)
      if (tid + q * R64_T < R64_PL) sq[slot][sr[q]][sc[q]] = src.value(pre[h][q]);
  };
  // diagonal at this column: the lid adds 1 whether or not j is the top row
  const double dgxy = 1.0 + (gx > 0) + (gx < nx - 1) + (gy > 0);
  double mx = 0.0, rr = 0.0, bb = 0.0;
  // prologue: planes k0 - 1 and k0 staged, k0 + 1 and k0 + 2 in flight
  fetch(k0 - 1, 0);
  fetch(k0, 1);
  stage(0, 0);
  stage(1, 1);
  fetch(k0 + 1, 1);
  fetch(k0 + 2, 0);
  double bnext = buf_load<double>(rbv, bo + (k0 - kbase) * pb8);
  __syncthreads();
  double qp = sq[0][ty + 1][tx + 1], qc = sq[1][ty + 1][tx + 1];
  int s0 = 0, s1 = 1, s2 = 2;  // slots of planes k-1, k, k+1
  // iteration k stages plane k + 1 from set (k - k0 + 1) & 1 and refills it
  // with plane k + 3
  auto body = [&](int k, int h) {
    stage(s2, h);
    fetch(k + 3, h);
    double bc = bnext;
    bnext = buf_load<double>(rbv, bo + (k + 1 - kbase) * pb8);
    __syncthreads();
    double qn = sq[s2][ty + 1][tx + 1];
    if (inside) {
      double sn = 0.0;
      sn += sq[s1][ty + 1][tx];
      sn += sq[s1][ty + 1][tx + 2];
      sn += sq[s1][ty][tx + 1];
      sn += sq[s1][ty + 2][tx + 1];
      sn += qp;
      sn += qn;
      double dg = dgxy + (k > 0) + (k < nz - 1);
      double res = bc - (sn - dg * qc);
      size_t c = (size_t)k * plane + (size_t)gy * nx + gx;
      qout[c] = qc;
      r[c] = (float)res;
      mx = fmax(mx, fabs(res));
      rr += res * res;
      bb += bc * bc;
    }
    qp = qc;
    qc = qn;
    // three slots: the next iteration stages into the slot of plane k - 1,
    // which no thread reads in this one, so no second barrier
    int t = s0;
    s0 = s1;
    s1 = s2;
    s2 = t;
  };
  int k = k0;
  for (; k + 1 < k1; k += 2) {
    body(k, 1);
    body(k + 1, 0);
  }
  if (k < k1) body(k, 1);
  mx = wave_max(mx);
  rr = wave_sum(rr);
  bb = wave_sum(bb);
  if ((tid & 63) == 0) {
    sred[0][tid >> 6] = mx;
    sred[1][tid >> 6] = rr;
    sred[2][tid >> 6] = bb;
  }
  __syncthreads();
  if (tid == 0) {
    double a = sred[0][0], e = sred[1][0], f = sred[2][0];
    for (int q = 1; q < R64_T / 64; ++q) {
      a = sred[0][q] > a ? sred[0][q] : a;
      e += sred[1][q];
      f += sred[2][q];
    }
    part[3 * blockIdx.x] = a;
    part[3 * blockIdx.x + 1] = e;
    part[3 * blockIdx.x + 2] = f;
  }
}

// block -> (tile, chunk) and the chunk's buffer ranges (planes k0-1 .. k1);
// blocks in XCD-contiguous order (one z-chunk of tiles per XCD), so a
// tile's ring lines are its neighbours' lines in the same L2
struct R64Chunk {
  int x0, y0, k0, k1, kb, n;
};
__device__ inline R64Chunk r64_chunk(const Dims &d, int cz, int ncx, int ncy) {
  R64Chunk c;
  int lb = xcd_remap(blockIdx.x, gridDim.x);
  c.x0 = (lb % ncx) * R64_TX;
  c.y0 = ((lb / ncx) % ncy) * R64_TY;
  c.k0 = (lb / (ncx * ncy)) * cz;
  c.k1 = min(c.k0 + cz, d.nz);
  c.kb = max(c.k0 - 1, 0);
  c.n = min(c.k1 + 1, d.nz) - c.kb;  // planes in range
  return c;
}

__global__ void __launch_bounds__(R64_T)
k_extrap_residual64_zm(Dims d, const double *__restrict__ p0, const double *__restrict__ p1,
                       const double *__restrict__ p2, double c0, double c1, double c2,
                       const double *__restrict__ b, double *__restrict__ qout,
                       float *__restrict__ r, double *__restrict__ part, int cz, int ncx,
                       int ncy) {
  R64Chunk ch = r64_chunk(d, cz, ncx, ncy);
  const size_t plane = (size_t)d.nx * d.ny;
  const int pb8 = (int)(plane * 8);
  const uint32_t bytes = (uint32_t)ch.n * pb8;
  R64Extrap src{buf_rsrc(p0 + ch.kb * plane, bytes), buf_rsrc(p1 + ch.kb * plane, bytes),
                buf_rsrc(p2 + ch.kb * plane, bytes), c0, c1, c2};
  residual64_zm(d, src, ch.kb, pb8, pb8 / 2, buf_rsrc(b + ch.kb * plane, bytes), qout, r, part,
                ch.k0, ch.k1, ch.x0, ch.y0);
}

__global__ void __launch_bounds__(R64_T)
k_update_residual64_zm(Dims d, const double *__restrict__ p, const float *__restrict__ e,
                       const double *__restrict__ scale, const double *__restrict__ b,
                       double *__restrict__ qout, float *__restrict__ r,
                       double *__restrict__ part, int cz, int ncx, int ncy) {
  R64Chunk ch = r64_chunk(d, cz, ncx, ncy);
  const size_t plane = (size_t)d.nx * d.ny;
  const int pb8 = (int)(plane * 8);
  const uint32_t bytes = (uint32_t)ch.n * pb8;
  R64Update src{buf_rsrc(p + ch.kb * plane, bytes), buf_rsrc(e + ch.kb * plane, bytes / 2),
                scale[1]};
  residual64_zm(d, src, ch.kb, pb8, pb8 / 2, buf_rsrc(b + ch.kb * plane, bytes), qout, r, part,
                ch.k0, ch.k1, ch.x0, ch.y0);
}

// Fold n (max, sum, sum) triples; one block of 256.  Also sets the scale of
// the next correction V-cycle: scale[0] = 1/max|r|, scale[1] = max|r|.
__global__ void __launch_bounds__(256)
k_fold3(const double *__restrict__ part, size_t n, double *__restrict__ out,
        double *__restrict__ scale) {
  double mx = 0.0, e = 0.0, f = 0.0;
  // four of the thread's triples loaded together (indices past n clamped to
  // n - 1, not summed), then summed in the order q = t, t + 256, ...
  for (size_t q0 = threadIdx.x; q0 < n; q0 += 4 * 256) {
    double a[4], b[4], c[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const size_t q = min(q0 + (size_t)u * 256, n - 1);
      a[u] = part[3 * q];
      b[u] = part[3 * q + 1];
      c[u] = part[3 * q + 2];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (q0 + (size_t)u * 256 < n) {
        mx = a[u] > mx ? a[u] : mx;
        e += b[u];
        f += c[u];
      }
  }
  mx = wave_max(mx);
  e = wave_sum(e);
  f = wave_sum(f);
  __shared__ double s[3][4];
  int t = threadIdx.x;
  if ((t & 63) == 0) { s[0][t >> 6] = mx; s[1][t >> 6] = e; s[2][t >> 6] = f; }
  __syncthreads();
  if (t == 0) {
    double a = s[0][0], g = s[1][0], h = s[2][0];
    for (int q = 1; q < 4; ++q) { a = s[0][q] > a ? s[0][q] : a; g += s[1][q]; h += s[2][q]; }
    out[0] = a;
    out[1] = g;
    out[2] = h;
    // right-hand-side scale of the next correction cycle and its inverse
    scale[0] = a > 0.0 ? 1.0 / a : 1.0;
    scale[1] = a;
  }
}

}  // namespace mac3d
