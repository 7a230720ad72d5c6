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

// r = b - L0 q in FP64 where q is a pointwise function of stored fields
// (value(c) gives q at cell c): q is written to qout at the centre, r to r32,
// per-block partials max|r|, sum r^2, sum b^2.  Used with
//  - the warm-start extrapolation q = c0 p_n + c1 p_{n-1} + c2 p_{n-2}, and
//  - the refinement update q = p + max|r| e,
// so the updated pressure is formed, stored and tested in one pass (qout must
// not alias the inputs: neighbours read the old values).
template <class F>
__device__ inline void residual64_of(const Dims &d, F value, const double *__restrict__ b,
                                     double *__restrict__ qout, float *__restrict__ r,
                                     double *__restrict__ part) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int k = blockIdx.z;
  double mx = 0.0, rr = 0.0, bb = 0.0;
  if (i < d.nx && j < d.ny) {
    size_t c = idx3(d, i, j, k);
    const size_t sy = (size_t)d.nx, sz = (size_t)d.nx * d.ny;
    double q = value(c);
    double sn = 0.0, dg = 1.0;  // level 0: the lid adds 1 whether or not j is the top row
    if (i > 0) { sn += value(c - 1); dg += 1.0; }
    if (i < d.nx - 1) { sn += value(c + 1); dg += 1.0; }
    if (j > 0) { sn += value(c - sy); dg += 1.0; }
    if (j < d.ny - 1) { sn += value(c + sy); }
    if (k > 0) { sn += value(c - sz); dg += 1.0; }
    if (k < d.nz - 1) { sn += value(c + sz); dg += 1.0; }
    double bc = b[c];
    double res = bc - (sn - dg * q);
    qout[c] = q;
    r[c] = (float)res;
    mx = fabs(res);
    rr = res * res;
    bb = bc * bc;
  }
  mx = wave_max(mx);
  rr = wave_sum(rr);
  bb = wave_sum(bb);
  __shared__ double s[3][4];
  int t = threadIdx.y * blockDim.x + threadIdx.x;
  if ((t & 63) == 0) { s[0][t >> 6] = mx; s[1][t >> 6] = rr; s[2][t >> 6] = bb; }
  __syncthreads();
  if (t == 0) {
    double a = s[0][0], e = s[1][0], f = s[2][0];
    for (int q = 1; q < 4; ++q) { a = s[0][q] > a ? s[0][q] : a; e += s[1][q]; f += s[2][q]; }
    size_t blk = ((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
    part[3 * blk] = a;
    part[3 * blk + 1] = e;
    part[3 * blk + 2] = f;
  }
}

__global__ void __launch_bounds__(256)
k_extrap_residual64(Dims d, const double *__restrict__ p0, const double *__restrict__ p1,
                    const double *__restrict__ p2, double c0, double c1, double c2,
                    const double *__restrict__ b, double *__restrict__ qout,
                    float *__restrict__ r, double *__restrict__ part) {
  residual64_of(d, [&](size_t c) { return c0 * p0[c] + c1 * p1[c] + c2 * p2[c]; }, b, qout, r,
                part);
}

__global__ void __launch_bounds__(256)
k_update_residual64(Dims d, const double *__restrict__ p, const float *__restrict__ e,
                    const double *__restrict__ scale, const double *__restrict__ b,
                    double *__restrict__ qout, float *__restrict__ r, double *__restrict__ part) {
  const double s = scale[1];
  residual64_of(d, [&](size_t c) { return p[c] + s * (double)e[c]; }, b, qout, r, part);
}

// Fold n (max, sum, sum) triples; one block of 256.  Also sets the scale of
// the next correction V-cycle: scale[0] = 1/max|r|, scale[1] = max|r|.
__global__ void __launch_bounds__(256)
k_fold3(const double *__restrict__ part, size_t n, double *__restrict__ out,
        double *__restrict__ scale) {
  double mx = 0.0, e = 0.0, f = 0.0;
  for (size_t q = threadIdx.x; q < n; q += 256) {
    double a = part[3 * q];
    mx = a > mx ? a : mx;
    e += part[3 * q + 1];
    f += part[3 * q + 2];
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
