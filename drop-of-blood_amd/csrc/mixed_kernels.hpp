// mixed_kernels.hpp — FP64 iterative-refinement kernels of the mixed mode.
//
// The pressure lives in FP64 (p64) with an FP64 right-hand side (b64).  Each
// refinement iteration computes r = b - L0 p in FP64, hands r to an FP32
// V-cycle as its right-hand side (scaled by 1/max|r| so the FP16 coarse
// levels stay in range) and adds the scaled correction back in FP64.
#pragma once

#include "common.hpp"
#include "mg_kernels.hpp"

namespace mac3d {

// r = b - (sum - diag p) on level 0 in FP64; r stored FP32.  Per-block
// partials: max|r|, sum r^2, sum b^2.
__global__ void __launch_bounds__(256)
k_residual64(Dims d, const double *__restrict__ p, const double *__restrict__ b,
             float *__restrict__ r, double *__restrict__ part) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int k = blockIdx.z;
  double mx = 0.0, rr = 0.0, bb = 0.0;
  if (i < d.nx && j < d.ny) {
    size_t c = idx3(d, i, j, k);
    double sn, dg;
    mg_stencil<double, double>(p, d, 1.0, i, j, k, c, sn, dg);
    double bc = b[c];
    double res = bc - (sn - dg * p[c]);
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

// p64 += max|r| * e: the correction of one refinement iteration, solved
// for the right-hand side r / max|r|
__global__ void __launch_bounds__(256)
k_update64(size_t n, double *__restrict__ p, const float *__restrict__ e,
           const double *__restrict__ scale) {
  size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t < n) p[t] += scale[1] * (double)e[t];
}

}  // namespace mac3d
