// mg_kernels.hpp — geometric multigrid kernels of the pressure solve.
//
// Level l holds cell-centred unknowns x, right-hand side b and residual r in
// storage type S (FP32, FP64, or FP16 on the coarse levels of the mixed
// mode); arithmetic runs in compute_of<S> (FP32 for FP16 storage).
// The operator on level l is (sum of neighbours - diag x) / h_l^2 with
// h_l^2 = 4^l (level-0 spacing 1), Neumann walls and a zero-pressure lid
// (evalMgStencil, src/shaders/fluid.ts:561-603).
#pragma once

#include "common.hpp"

namespace mac3d {

// Neighbour sum and diagonal of evalMgStencil.  topc is the top row's lid
// coefficient on this level (1 on level 0).
template <class S, class C>
__device__ inline void mg_stencil(const S *__restrict__ x, const Dims &d, C topc,
                                  int i, int j, int k, size_t c, C &sum, C &diag) {
  const size_t sy = (size_t)d.nx, sz = (size_t)d.nx * d.ny;
  C sn = C(0), dg = (j < d.ny - 1) ? C(1) : topc;
  if (i > 0) { sn += (C)x[c - 1]; dg += C(1); }
  if (i < d.nx - 1) { sn += (C)x[c + 1]; dg += C(1); }
  if (j > 0) { sn += (C)x[c - sy]; dg += C(1); }
  if (j < d.ny - 1) { sn += (C)x[c + sy]; }
  if (k > 0) { sn += (C)x[c - sz]; dg += C(1); }
  if (k < d.nz - 1) { sn += (C)x[c + sz]; dg += C(1); }
  sum = sn;
  diag = dg;
}

// One colour pass of red-black Gauss-Seidel, in place
// (MG_RBGS_RED / MG_RBGS_BLACK, fluid.ts:645-681): cells with
// (i+j+k)&1 == parity get x = (sum - h^2 b bscale)/diag.  Threads map to the
// cells of one colour: thread t of a row takes i = 2t + ((parity+j+k)&1).
// zero_guess: the other colour is known to be 0 and is written as 0, this
// colour's neighbours are not read (MG_RBGS_RED_ZERO, fluid.ts:605-641).
template <class S, bool ZERO>
__global__ void __launch_bounds__(256)
k_rbgs(Dims d, S *__restrict__ x, const S *__restrict__ b,
       typename compute_of<S>::type h2, typename compute_of<S>::type topc,
       typename compute_of<S>::type bscale, int parity) {
  typedef typename compute_of<S>::type C;
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int k = blockIdx.z;
  if (j >= d.ny) return;
  int off = (parity + j + k) & 1;
  int i = 2 * t + off;
  if (ZERO) {
    // both colours of this thread's pair are written: the pair is (2t, 2t+1)
    int ia = 2 * t;
    if (ia >= d.nx) return;
    for (int q = 0; q < 2; ++q) {
      int ii = ia + q;
      if (ii >= d.nx) break;
      size_t c = idx3(d, ii, j, k);
      if (((ii + j + k) & 1) == parity) {
        C dg = (j < d.ny - 1) ? C(1) : topc;
        if (ii > 0) dg += C(1);
        if (ii < d.nx - 1) dg += C(1);
        if (j > 0) dg += C(1);
        if (k > 0) dg += C(1);
        if (k < d.nz - 1) dg += C(1);
        x[c] = store_as<S, C>((-h2 * ((C)b[c] * bscale)) / dg);
      } else {
        x[c] = S(0);
      }
    }
    return;
  }
  if (i >= d.nx) return;
  size_t c = idx3(d, i, j, k);
  C sn, dg;
  mg_stencil<S, C>(x, d, topc, i, j, k, c, sn, dg);
  x[c] = store_as<S, C>((sn - h2 * ((C)b[c] * bscale)) / dg);
}

// MG_RESIDUAL (fluid.ts:744-770): r = b bscale - (sum - diag x)/h^2
template <class S>
__global__ void __launch_bounds__(256)
k_residual(Dims d, const S *__restrict__ x, const S *__restrict__ b, S *__restrict__ r,
           typename compute_of<S>::type h2, typename compute_of<S>::type topc,
           typename compute_of<S>::type bscale) {
  typedef typename compute_of<S>::type C;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int k = blockIdx.z;
  if (i >= d.nx || j >= d.ny) return;
  size_t c = idx3(d, i, j, k);
  C sn, dg;
  mg_stencil<S, C>(x, d, topc, i, j, k, c, sn, dg);
  r[c] = store_as<S, C>((C)b[c] * bscale - (sn - dg * (C)x[c]) / h2);
}

// MG_RESTRICT (fluid.ts:775-800): coarse b = 1/8 sum of the 2x2x2 children
template <class Sf, class Sc>
__global__ void __launch_bounds__(256)
k_restrict(Dims df, Dims dc, const Sf *__restrict__ rf, Sc *__restrict__ bc) {
  typedef typename compute_of<Sf>::type C;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int k = blockIdx.z;
  if (i >= dc.nx || j >= dc.ny) return;
  const size_t sy = (size_t)df.nx, sz = (size_t)df.nx * df.ny;
  size_t f = idx3(df, 2 * i, 2 * j, 2 * k);
  C s = (C)rf[f] + (C)rf[f + 1];
  s += (C)rf[f + sy] + (C)rf[f + sy + 1];
  s += (C)rf[f + sz] + (C)rf[f + sz + 1];
  s += (C)rf[f + sz + sy] + (C)rf[f + sz + sy + 1];
  bc[idx3(dc, i, j, k)] = store_as<Sc, C>(s * C(0.125));
}

// MG_PROLONG_ADD (fluid.ts:806-855): x_f += cell-centred trilinear
// interpolation of the coarse correction; coarse indices clamped at the
// walls, 0 above the top row.  Fine cell i sits at coarse coordinate
// i/2 - 1/4, so the weights are 3/4 (own parent) and 1/4 (next parent
// towards the fine cell).
template <class Sc>
__device__ inline typename compute_of<Sc>::type
coarse_at(const Sc *__restrict__ e, const Dims &dc, int i, int j, int k) {
  typedef typename compute_of<Sc>::type C;
  if (j >= dc.ny) return C(0);
  i = clampi(i, 0, dc.nx - 1);
  j = clampi(j, 0, dc.ny - 1);
  k = clampi(k, 0, dc.nz - 1);
  return (C)e[idx3(dc, i, j, k)];
}

template <class Sc, class Sf>
__global__ void __launch_bounds__(256)
k_prolong_add(Dims df, Dims dc, const Sc *__restrict__ ec, Sf *__restrict__ xf) {
  typedef typename compute_of<Sf>::type C;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int k = blockIdx.z;
  if (i >= df.nx || j >= df.ny) return;
  // base = floor(i/2 - 1/4), frac = 3/4 for even i, 1/4 for odd i
  int bi = (i & 1) ? (i >> 1) : (i >> 1) - 1;
  int bj = (j & 1) ? (j >> 1) : (j >> 1) - 1;
  int bk = (k & 1) ? (k >> 1) : (k >> 1) - 1;
  C fx = (i & 1) ? C(0.25) : C(0.75);
  C fy = (j & 1) ? C(0.25) : C(0.75);
  C fz = (k & 1) ? C(0.25) : C(0.75);
  C c000 = (C)coarse_at(ec, dc, bi, bj, bk), c100 = (C)coarse_at(ec, dc, bi + 1, bj, bk);
  C c010 = (C)coarse_at(ec, dc, bi, bj + 1, bk), c110 = (C)coarse_at(ec, dc, bi + 1, bj + 1, bk);
  C c001 = (C)coarse_at(ec, dc, bi, bj, bk + 1), c101 = (C)coarse_at(ec, dc, bi + 1, bj, bk + 1);
  C c011 = (C)coarse_at(ec, dc, bi, bj + 1, bk + 1), c111 = (C)coarse_at(ec, dc, bi + 1, bj + 1, bk + 1);
  C x00 = mixv(c000, c100, fx), x10 = mixv(c010, c110, fx);
  C x01 = mixv(c001, c101, fx), x11 = mixv(c011, c111, fx);
  C corr = mixv(mixv(x00, x10, fy), mixv(x01, x11, fy), fz);
  size_t c = idx3(df, i, j, k);
  xf[c] = store_as<Sf, C>((C)xf[c] + corr);
}

// ---------------------------------------------------------------------------
// Exact coarsest-level solve by separation of variables.  The level operator
// is a Kronecker sum A = Ax (+) Ay (+) Az of 1-D tridiagonal operators
// (Neumann-Neumann in x and z; Neumann bottom and the lid coefficient on top
// in y), so A = Q diag(lx_a + ly_b + lz_c) Q^T with Q = Qx (x) Qy (x) Qz the
// orthonormal 1-D eigenvectors.  The solve of (sum - diag x) = h^2 b is
// x = -h^2 A^{-1} b, done as three kernels:
//   1. per y-plane: t = (Qx^T (x) Qz^T) b           (x, z transforms)
//   2. per z-slab: y transform, divide by the eigenvalue sum, inverse y
//   3. per y-plane: x = -h^2 (Qx (x) Qz) t          (inverse x, z)
// Q[m * n + a] is component m of eigenvector a.  Work buffer: FP64.
// LDS: two planes of at most COARSE_MAX_PLANE values each.
#define COARSE_MAX_PLANE 2048

template <class S, class C>
__global__ void __launch_bounds__(256)
k_coarse_xz_fwd(Dims d, const S *__restrict__ b, const double *__restrict__ qx,
                const double *__restrict__ qz, double *__restrict__ work) {
  __shared__ C s0[COARSE_MAX_PLANE], s1[COARSE_MAX_PLANE];
  int j = blockIdx.x;
  int nx = d.nx, nz = d.nz, np = nx * nz;
  for (int t = threadIdx.x; t < np; t += 256) {
    int i = t % nx, k = t / nx;
    s0[t] = (C)b[idx3(d, i, j, k)];
  }
  __syncthreads();
  // s1[k][a] = sum_m Qx[m][a] s0[k][m]
  for (int t = threadIdx.x; t < np; t += 256) {
    int a = t % nx, k = t / nx;
    C acc = C(0);
    for (int m = 0; m < nx; ++m) acc += (C)qx[m * nx + a] * s0[k * nx + m];
    s1[t] = acc;
  }
  __syncthreads();
  // out[c][a] = sum_n Qz[n][c] s1[n][a]
  for (int t = threadIdx.x; t < np; t += 256) {
    int a = t % nx, cc = t / nx;
    C acc = C(0);
    for (int n = 0; n < nz; ++n) acc += (C)qz[n * nz + cc] * s1[n * nx + a];
    work[idx3(d, a, j, cc)] = (double)acc;
  }
}

template <class C>
__global__ void __launch_bounds__(256)
k_coarse_y(Dims d, const double *__restrict__ qy, const double *__restrict__ lx,
           const double *__restrict__ ly, const double *__restrict__ lz, double h2,
           double *__restrict__ work) {
  __shared__ C s0[COARSE_MAX_PLANE], s1[COARSE_MAX_PLANE];
  int cz = blockIdx.x;
  int nx = d.nx, ny = d.ny, np = nx * ny;
  for (int t = threadIdx.x; t < np; t += 256) {
    int a = t % nx, m = t / nx;
    s0[t] = (C)work[idx3(d, a, m, cz)];
  }
  __syncthreads();
  // s1[bb][a] = (sum_m Qy[m][bb] s0[m][a]) / (lx_a + ly_bb + lz_cz) * (-h2)
  for (int t = threadIdx.x; t < np; t += 256) {
    int a = t % nx, bb = t / nx;
    C acc = C(0);
    for (int m = 0; m < ny; ++m) acc += (C)qy[m * ny + bb] * s0[m * nx + a];
    s1[t] = acc * (C)(-h2) / (C)(lx[a] + ly[bb] + lz[cz]);
  }
  __syncthreads();
  // back: work[m][a] = sum_bb Qy[m][bb] s1[bb][a]
  for (int t = threadIdx.x; t < np; t += 256) {
    int a = t % nx, m = t / nx;
    C acc = C(0);
    for (int bb = 0; bb < ny; ++bb) acc += (C)qy[m * ny + bb] * s1[bb * nx + a];
    work[idx3(d, a, m, cz)] = (double)acc;
  }
}

template <class S, class C>
__global__ void __launch_bounds__(256)
k_coarse_xz_inv(Dims d, const double *__restrict__ work, const double *__restrict__ qx,
                const double *__restrict__ qz, S *__restrict__ x) {
  __shared__ C s0[COARSE_MAX_PLANE], s1[COARSE_MAX_PLANE];
  int j = blockIdx.x;
  int nx = d.nx, nz = d.nz, np = nx * nz;
  for (int t = threadIdx.x; t < np; t += 256) {
    int a = t % nx, cc = t / nx;
    s0[t] = (C)work[idx3(d, a, j, cc)];
  }
  __syncthreads();
  // s1[n][a] = sum_c Qz[n][c] s0[c][a]
  for (int t = threadIdx.x; t < np; t += 256) {
    int a = t % nx, n = t / nx;
    C acc = C(0);
    for (int cc = 0; cc < nz; ++cc) acc += (C)qz[n * nz + cc] * s0[cc * nx + a];
    s1[t] = acc;
  }
  __syncthreads();
  // x[n][m] = sum_a Qx[m][a] s1[n][a]
  for (int t = threadIdx.x; t < np; t += 256) {
    int m = t % nx, n = t / nx;
    C acc = C(0);
    for (int a = 0; a < nx; ++a) acc += (C)qx[m * nx + a] * s1[n * nx + a];
    x[idx3(d, m, j, n)] = store_as<S, C>(acc);
  }
}

}  // namespace mac3d
