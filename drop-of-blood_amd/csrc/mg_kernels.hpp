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

// MG_RESIDUAL (fluid.ts:744-770): r = b scale - (sum - diag x)/h^2, with
// scale = scale[0] (nullable: 1)
template <class S>
__global__ void __launch_bounds__(256)
k_residual(Dims d, const S *__restrict__ x, const S *__restrict__ b, S *__restrict__ r,
           typename compute_of<S>::type h2, typename compute_of<S>::type topc,
           const double *__restrict__ scale) {
  typedef typename compute_of<S>::type C;
  const C bscale = scale ? (C)scale[0] : C(1);
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

// MG_RESIDUAL followed by MG_RESTRICT (fluid.ts:744-800) in one pass: each
// coarse cell sums the residuals of its 2x2x2 children, so the fine residual
// never goes to memory.  Same arithmetic as the two kernels.
template <class Sf, class Sc>
__global__ void __launch_bounds__(256)
k_residual_restrict(Dims df, Dims dc, const Sf *__restrict__ x, const Sf *__restrict__ b,
                    Sc *__restrict__ bc, typename compute_of<Sf>::type h2,
                    typename compute_of<Sf>::type topc, const double *__restrict__ scale) {
  typedef typename compute_of<Sf>::type C;
  const C bscale = scale ? (C)scale[0] : C(1);
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int k = blockIdx.z;
  if (i >= dc.nx || j >= dc.ny) return;
  C s = C(0);
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    int fi = 2 * i + (q & 1), fj = 2 * j + ((q >> 1) & 1), fk = 2 * k + (q >> 2);
    size_t c = idx3(df, fi, fj, fk);
    C sn, dg;
    mg_stencil<Sf, C>(x, df, topc, fi, fj, fk, c, sn, dg);
    s += (C)b[c] * bscale - (sn - dg * (C)x[c]) / h2;
  }
  bc[idx3(dc, i, j, k)] = store_as<Sc, C>(s * C(0.125));
}

// The same residual-restriction with one thread per fine cell (coalesced
// 7-point loads): a 64 x 2 x 2 block of fine cells sums to 32 coarse cells,
// pairs in x by a lane shuffle, the 2 x 2 (y, z) by LDS.
template <class Sf, class Sc>
__global__ void __launch_bounds__(256)
k_residual_restrict_t(Dims df, Dims dc, const Sf *__restrict__ x, const Sf *__restrict__ b,
                      Sc *__restrict__ bc, typename compute_of<Sf>::type h2,
                      typename compute_of<Sf>::type topc, const double *__restrict__ scale) {
  typedef typename compute_of<Sf>::type C;
  __shared__ C sr[4][64];
  const C bscale = scale ? (C)scale[0] : C(1);
  const int tx = threadIdx.x, ty = threadIdx.y, tz = threadIdx.z;
  const int i = blockIdx.x * 64 + tx, j = 2 * blockIdx.y + ty, k = 2 * blockIdx.z + tz;
  C r = C(0);
  if (i < df.nx) {
    size_t c = idx3(df, i, j, k);
    C sn, dg;
    mg_stencil<Sf, C>(x, df, topc, i, j, k, c, sn, dg);
    r = (C)b[c] * bscale - (sn - dg * (C)x[c]) / h2;
  }
  r += __shfl_xor(r, 1, 64);
  sr[ty * 2 + tz][tx] = r;
  __syncthreads();
  if (ty == 0 && tz == 0 && !(tx & 1)) {
    int ci = (blockIdx.x * 64 + tx) >> 1;
    if (ci < dc.nx) {
      C s = sr[0][tx] + sr[1][tx] + sr[2][tx] + sr[3][tx];
      bc[idx3(dc, ci, blockIdx.y, blockIdx.z)] = store_as<Sc, C>(s * C(0.125));
    }
  }
}

// Residual-restriction marching in z: a 512-thread block owns a 64 x 8 fine
// column tile and an even z-chunk; planes of x are staged in a three-plane
// LDS ring with a one-cell ring in x and y (each x value read about 1.3x
// instead of ~3x by independent blocks), the 2 x 2 x 2 child sums are formed
// in registers (z), a lane shuffle (x) and LDS (y).  Same arithmetic as
// k_residual followed by k_restrict.
#define RR_TX 64
#define RR_TY 8
#define RR_EX (RR_TX + 2)
#define RR_EY (RR_TY + 2)
#define RR_T (RR_TX * RR_TY)
#define RR_PL (RR_EX * RR_EY)
#define RR_LPT ((RR_PL + RR_T - 1) / RR_T)

template <class Sf, class Sc>
__global__ void __launch_bounds__(RR_T)
k_residual_restrict_zm(Dims df, Dims dc, const Sf *__restrict__ x, const Sf *__restrict__ b,
                       Sc *__restrict__ bc, typename compute_of<Sf>::type h2,
                       typename compute_of<Sf>::type topc, const double *__restrict__ scale,
                       int cz, int ncx, int ncy) {
  typedef typename compute_of<Sf>::type C;
  __shared__ C sx[3][RR_EY][RR_EX];
  __shared__ C sred[RR_TY][RR_TX];
  const C bscale = scale ? (C)scale[0] : C(1);
  const int nx = df.nx, ny = df.ny, nz = df.nz;
  const int lb = blockIdx.x;
  const int x0 = (lb % ncx) * RR_TX, y0 = ((lb / ncx) % ncy) * RR_TY;
  const int k0 = (lb / (ncx * ncy)) * cz, k1 = min(k0 + cz, nz);
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * RR_TX + tx;
  const size_t plane = (size_t)nx * ny;
  const int gx = x0 + tx, gy = y0 + ty;
  const bool inside = gx < nx && gy < ny;
  long st_g[RR_LPT];
  int st_r[RR_LPT], st_c[RR_LPT];
#pragma unroll
  for (int q = 0; q < RR_LPT; ++q) {
    int v = tid + q * RR_T;
    st_r[q] = v / RR_EX;
    st_c[q] = v % RR_EX;
    int sxg = x0 - 1 + st_c[q], syg = y0 - 1 + st_r[q];
    st_g[q] = (v < RR_PL && sxg >= 0 && sxg < nx && syg >= 0 && syg < ny) ? (long)syg * nx + sxg : -1;
  }
  C pre[RR_LPT];
  auto fetch = [&](int k) {
#pragma unroll
    for (int q = 0; q < RR_LPT; ++q)
      pre[q] = (st_g[q] >= 0 && k >= 0 && k < nz) ? (C)x[k * plane + st_g[q]] : C(0);
  };
  auto stage = [&](int k) {
#pragma unroll
    for (int q = 0; q < RR_LPT; ++q)
      if (tid + q * RR_T < RR_PL) sx[(k + 3) % 3][st_r[q]][st_c[q]] = pre[q];
  };
  for (int k = k0 - 1; k <= k0; ++k) {
    fetch(k);
    stage(k);
  }
  fetch(k0 + 1);
  C rsum = C(0);
  for (int k = k0; k < k1; ++k) {
    stage(k + 1);
    __syncthreads();
    fetch(k + 2);
    if (inside) {
      const int r = ty + 1, c = tx + 1, s0 = k % 3;
      C sn = C(0), dg = (gy < ny - 1) ? C(1) : topc;
      if (gx > 0) { sn += sx[s0][r][c - 1]; dg += C(1); }
      if (gx < nx - 1) { sn += sx[s0][r][c + 1]; dg += C(1); }
      if (gy > 0) { sn += sx[s0][r - 1][c]; dg += C(1); }
      if (gy < ny - 1) { sn += sx[s0][r + 1][c]; }
      if (k > 0) { sn += sx[(k + 2) % 3][r][c]; dg += C(1); }
      if (k < nz - 1) { sn += sx[(k + 1) % 3][r][c]; dg += C(1); }
      rsum += (C)b[k * plane + (size_t)gy * nx + gx] * bscale - (sn - dg * sx[s0][r][c]) / h2;
    }
    if (k & 1) {
      C pair = rsum + __shfl_xor(rsum, 1, 64);
      sred[ty][tx] = pair;
      rsum = C(0);
    }
    __syncthreads();
    if ((k & 1) && !(ty & 1) && !(tx & 1) && inside) {
      C s = sred[ty][tx] + sred[ty + 1][tx];
      bc[idx3(dc, gx >> 1, gy >> 1, k >> 1)] = store_as<Sc, C>(s * C(0.125));
    }
  }
}

// MG_PROLONG_ADD with one thread per 2 x 2 x 2 block of fine cells (the
// children of one coarse cell), which share a 3 x 3 x 3 coarse
// neighbourhood: child offset 0 along an axis takes 3/4 of the parent and
// 1/4 of the parent's lower neighbour, offset 1 takes 3/4 and 1/4 of the
// upper neighbour (cell-centred weights of fluid.ts:806-855; clamped at the
// walls, 0 above the top row).
template <class Sc, class Sf>
__global__ void __launch_bounds__(256)
k_prolong_add_2x(Dims df, Dims dc, const Sc *__restrict__ ec, Sf *__restrict__ xf) {
  typedef typename compute_of<Sf>::type C;
  int I = blockIdx.x * blockDim.x + threadIdx.x;
  int J = blockIdx.y * blockDim.y + threadIdx.y;
  int K = blockIdx.z;
  if (I >= dc.nx || J >= dc.ny) return;
  C e[3][3][3];  // [dz][dy][dx], offsets -1..1
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int bq = 0; bq < 3; ++bq)
#pragma unroll
      for (int q = 0; q < 3; ++q) e[a][bq][q] = (C)coarse_at(ec, dc, I + q - 1, J + bq - 1, K + a - 1);
  // weights along one axis for child offset o: (lower, mid, upper)
  const C wl[2] = {C(0.25), C(0)}, wm = C(0.75), wu[2] = {C(0), C(0.25)};
#pragma unroll
  for (int oz = 0; oz < 2; ++oz)
#pragma unroll
    for (int oy = 0; oy < 2; ++oy) {
      C yz[3];  // interpolated in z and y, per x offset
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        C zc[3];
#pragma unroll
        for (int bq = 0; bq < 3; ++bq)
          zc[bq] = wl[oz] * e[0][bq][q] + wm * e[1][bq][q] + wu[oz] * e[2][bq][q];
        yz[q] = wl[oy] * zc[0] + wm * zc[1] + wu[oy] * zc[2];
      }
      size_t base = idx3(df, 2 * I, 2 * J + oy, 2 * K + oz);
#pragma unroll
      for (int ox = 0; ox < 2; ++ox) {
        C corr = wl[ox] * yz[0] + wm * yz[1] + wu[ox] * yz[2];
        xf[base + ox] = store_as<Sf, C>((C)xf[base + ox] + corr);
      }
    }
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

namespace mac3d {

// ---------------------------------------------------------------------------
// One colour pass of red-black Gauss-Seidel in place (MG_RBGS_RED /
// MG_RBGS_BLACK, fluid.ts:645-681): cells with (i+j+k)&1 == parity get
// x = (sum - h^2 b)/diag.  Used on the coarse levels, where a grid holds too
// few planes for the z-marching sweep to pay.  Thread t of a row takes the
// pair (2t, 2t+1) and updates the one of this colour.  ZERO: the sweep starts
// from a zero guess (MG_RBGS_RED_ZERO, fluid.ts:605-641): the other colour is
// written as 0 and no neighbour is read.  omega != 1: over-relaxed update
// x + omega (gs - x) (the mixed mode's smoother; the reference's is omega 1).
template <class S, bool ZERO>
__global__ void __launch_bounds__(256)
k_rbgs_pass(Dims d, S *__restrict__ x, const S *__restrict__ b,
            typename compute_of<S>::type h2, typename compute_of<S>::type topc,
            typename compute_of<S>::type omega, int parity) {
  typedef typename compute_of<S>::type C;
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int k = blockIdx.z;
  int ia = 2 * t;
  if (j >= d.ny || ia >= d.nx) return;
  int i = ia + ((ia + j + k + parity) & 1);
  if (ZERO) {
    int io = ia + ((ia + j + k + parity + 1) & 1);
    if (io < d.nx) x[idx3(d, io, j, k)] = S(0);
  }
  if (i >= d.nx) return;
  size_t c = idx3(d, i, j, k);
  C sn, dg;
  if (ZERO) {
    sn = C(0);
    dg = (j < d.ny - 1) ? C(1) : topc;
    if (i > 0) dg += C(1);
    if (i < d.nx - 1) dg += C(1);
    if (j > 0) dg += C(1);
    if (k > 0) dg += C(1);
    if (k < d.nz - 1) dg += C(1);
  } else {
    mg_stencil<S, C>(x, d, topc, i, j, k, c, sn, dg);
  }
  C v = (sn - h2 * (C)b[c]) / dg;
  if (omega != C(1)) {
    C old = ZERO ? C(0) : (C)x[c];
    v = old + omega * (v - old);
  }
  x[c] = store_as<S, C>(v);
}

// ---------------------------------------------------------------------------
// Level-0 red-black sweep: both colour passes in one kernel, out of place
// (xin -> xout): 12 B per FP32 cell per sweep instead of 24 for two in-place
// colour passes.  Each 512-thread block owns a 64 x 16 column tile and
// marches over a z-chunk [k0, k1).  Planes of xin and b are staged in LDS
// with a two-cell ring (8-byte pair loads: x0 - 2 is even); red is recomputed
// on the tile plus a one-cell ring (what the tile's black cells read), so no
// block waits on another.  Threads work on (even, odd) x-pairs, each pair
// holding one cell of each colour, so no lane idles on a parity test.
// Registers prefetch two planes ahead.  ZERO: zero initial guess (xin is not
// read).  scale (nullable): the right-hand side is b * scale[0].  omega != 1:
// over-relaxed updates x + omega (gs - x).
#define SW_TX 64
#define SW_TY 16
#define SW_EX (SW_TX + 4)                 // staged row, two-cell ring
#define SW_EY (SW_TY + 4)
#define SW_THREADS 512
#define SW_NV (SW_EX / 2 * SW_EY)         // 680 staged pairs per plane
#define SW_LPT ((SW_NV + SW_THREADS - 1) / SW_THREADS)   // 2
#define SW_RP ((SW_TX + 2) / 2 * (SW_TY + 2))           // 594 red pairs
#define SW_RPT ((SW_RP + SW_THREADS - 1) / SW_THREADS)  // 2

template <class S> struct vec2_of;
template <> struct vec2_of<float> { typedef float2 type; };
template <> struct vec2_of<double> { typedef double2 type; };

__device__ inline float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline double fast_rcp(double x) { return 1.0 / x; }

// Blocks are numbered on a 1-D grid and remapped so that each of the 8 XCDs
// (blocks are dealt round-robin to XCDs) gets a contiguous run of tiles: the
// halo planes and rows a tile shares with its neighbours then sit in the
// same XCD's L2.
__device__ inline int xcd_remap(int b, int nblocks) {
  if (nblocks % 8) return b;
  return (b % 8) * (nblocks / 8) + b / 8;
}

template <class S, bool ZERO>
__global__ void __launch_bounds__(SW_THREADS)
k_rbgs_sweep(Dims d, const S *__restrict__ xin, S *__restrict__ xout, const S *__restrict__ b,
             typename compute_of<S>::type h2, typename compute_of<S>::type topc,
             typename compute_of<S>::type omega, const double *__restrict__ scale, int cz,
             int ncx, int ncy) {
  typedef typename compute_of<S>::type C;
  typedef typename vec2_of<S>::type V2;
  __shared__ C sp[4][SW_EY][SW_EX];
  __shared__ C sb[2][SW_EY][SW_EX];
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  const int lb = xcd_remap(blockIdx.x, gridDim.x);
  const int x0 = (lb % ncx) * SW_TX, y0 = ((lb / ncx) % ncy) * SW_TY;
  const int k0 = (lb / (ncx * ncy)) * cz;
  const int k1 = min(k0 + cz, nz);
  const int tid = threadIdx.x;
  const size_t plane = (size_t)nx * ny;
  const C bs = scale ? (C)scale[0] : C(1);

  // staged pairs of this thread (fixed over the march)
  int st_r[SW_LPT], st_c[SW_LPT];
  long st_g[SW_LPT];  // in-plane offset, -1 when outside the grid or the stage
#pragma unroll
  for (int q = 0; q < SW_LPT; ++q) {
    int v = tid + q * SW_THREADS;
    st_r[q] = v / (SW_EX / 2);
    st_c[q] = 2 * (v % (SW_EX / 2));
    int gx = x0 - 2 + st_c[q], gy = y0 - 2 + st_r[q];
    st_g[q] = (v < SW_NV && gx >= 0 && gx < nx && gy >= 0 && gy < ny) ? (long)gy * nx + gx : -1;
  }
  auto fetch = [&](const S *__restrict__ f, int k, V2 (&pre)[SW_LPT]) {
#pragma unroll
    for (int q = 0; q < SW_LPT; ++q) {
      V2 v;
      v.x = S(0);
      v.y = S(0);
      if (st_g[q] >= 0 && k >= 0 && k < nz) v = *(const V2 *)(f + k * plane + st_g[q]);
      pre[q] = v;
    }
  };
  auto stage_p = [&](int k, const V2 (&pre)[SW_LPT]) {
#pragma unroll
    for (int q = 0; q < SW_LPT; ++q)
      if (tid + q * SW_THREADS < SW_NV) {
        sp[k & 3][st_r[q]][st_c[q]] = (C)pre[q].x;
        sp[k & 3][st_r[q]][st_c[q] + 1] = (C)pre[q].y;
      }
  };
  auto stage_b = [&](int k, const V2 (&pre)[SW_LPT]) {
#pragma unroll
    for (int q = 0; q < SW_LPT; ++q)
      if (tid + q * SW_THREADS < SW_NV) {
        sb[k & 1][st_r[q]][st_c[q]] = (C)pre[q].x * bs;
        sb[k & 1][st_r[q]][st_c[q] + 1] = (C)pre[q].y * bs;
      }
  };

  // red pairs of this thread on the tile plus one-cell ring: left cell at
  // grid (rgx, rgy), stage (rr, rc); flags per cell of the pair
  int rr[SW_RPT], rc[SW_RPT], rgx[SW_RPT], rgy[SW_RPT];
  bool rv[SW_RPT];
#pragma unroll
  for (int q = 0; q < SW_RPT; ++q) {
    int v = tid + q * SW_THREADS;
    int row = v / ((SW_TX + 2) / 2), pc = v % ((SW_TX + 2) / 2);
    rr[q] = row + 1;
    rc[q] = 2 * pc + 1;
    rgx[q] = x0 - 1 + 2 * pc;
    rgy[q] = y0 - 1 + row;
    rv[q] = v < SW_RP && rgy[q] >= 0 && rgy[q] < ny;
  }
  // stencil update of the cell at stage (r, c), grid (gx, gy, k)
  // (nbrs false: red cells of a zero-guess sweep, whose neighbours are 0)
  auto update = [&](int r, int c, int gx, int gy, int k, bool nbrs) {
    C sn = C(0), dg = (gy < ny - 1) ? C(1) : topc;
    if (gx > 0) { if (nbrs) sn += sp[k & 3][r][c - 1]; dg += C(1); }
    if (gx < nx - 1) { if (nbrs) sn += sp[k & 3][r][c + 1]; dg += C(1); }
    if (gy > 0) { if (nbrs) sn += sp[k & 3][r - 1][c]; dg += C(1); }
    if (gy < ny - 1) { if (nbrs) sn += sp[k & 3][r + 1][c]; }
    if (k > 0) { if (nbrs) sn += sp[(k - 1) & 3][r][c]; dg += C(1); }
    if (k < nz - 1) { if (nbrs) sn += sp[(k + 1) & 3][r][c]; dg += C(1); }
    C v = (sn - h2 * sb[k & 1][r][c]) * fast_rcp(dg);
    if (omega != C(1)) {
      C old = ZERO ? C(0) : sp[k & 3][r][c];
      v = old + omega * (v - old);
    }
    return v;
  };
  auto red = [&](int k) {
    if (k < 0 || k >= nz) return;
#pragma unroll
    for (int q = 0; q < SW_RPT; ++q) {
      if (!rv[q]) continue;
      int w = (rgx[q] + rgy[q] + k) & 1;  // 0: left cell red, 1: right cell red
      int gx = rgx[q] + w;
      if (gx < 0 || gx >= nx) continue;
      sp[k & 3][rr[q]][rc[q] + w] = update(rr[q], rc[q] + w, gx, rgy[q], k, !ZERO);
    }
  };
  // black pair of this thread on the tile
  const int bty = tid / (SW_TX / 2), bpx = tid % (SW_TX / 2);
  const int bgx = x0 + 2 * bpx, bgy = y0 + bty;
  const int br = bty + 2, bc = 2 * bpx + 2;
  const bool binside = bgx < nx && bgy < ny;
  auto black_and_write = [&](int k) {
    if (!binside) return;
    int w = (bgx + bgy + k + 1) & 1;  // 0: left cell black, 1: right cell black
    C vb = update(br, bc + w, bgx + w, bgy, k, true);
    C vr = sp[k & 3][br][bc + 1 - w];
    V2 o;
    o.x = store_as<S, C>(w ? vr : vb);
    o.y = store_as<S, C>(w ? vb : vr);
    *(V2 *)(xout + k * plane + (size_t)bgy * nx + bgx) = o;
  };

  // Register ring of depth two: at the start of iteration k, ring k&1 holds
  // plane k+2 of xin and plane k+1 of b; the plane four ahead is fetched into
  // it as soon as it is staged, so two planes of loads are in flight.
  V2 pA[SW_LPT], pB[SW_LPT], bA[SW_LPT], bB[SW_LPT];
  auto iter = [&](int k, V2 (&pr)[SW_LPT], V2 (&br_)[SW_LPT]) {
    if (!ZERO) stage_p(k + 2, pr);
    stage_b(k + 1, br_);
    __syncthreads();
    if (!ZERO) fetch(xin, k + 4, pr);
    fetch(b, k + 3, br_);
    red(k + 1);
    __syncthreads();
    black_and_write(k);
    __syncthreads();
  };

  // prologue: planes k0-2 .. k0+1 of xin, b planes k0-1, k0; red on k0-1, k0
  {
    V2 t0[SW_LPT], t1[SW_LPT], t2[SW_LPT], t3[SW_LPT], u0[SW_LPT], u1[SW_LPT];
    if (!ZERO) {
      fetch(xin, k0 - 2, t0);
      fetch(xin, k0 - 1, t1);
      fetch(xin, k0, t2);
      fetch(xin, k0 + 1, t3);
    }
    fetch(b, k0 - 1, u0);
    fetch(b, k0, u1);
    if (!ZERO) {
      stage_p(k0 - 2, t0);
      stage_p(k0 - 1, t1);
      stage_p(k0, t2);
      stage_p(k0 + 1, t3);
    }
    stage_b(k0 - 1, u0);
    stage_b(k0, u1);
  }
  if (!ZERO) {
    fetch(xin, k0 + 2, pA);
    fetch(xin, k0 + 3, pB);
  }
  fetch(b, k0 + 1, bA);
  fetch(b, k0 + 2, bB);
  __syncthreads();
  red(k0 - 1);
  red(k0);
  __syncthreads();
  for (int k = k0; k < k1; k += 2) {
    iter(k, pA, bA);
    if (k + 1 < k1) iter(k + 1, pB, bB);
  }
}

}  // namespace mac3d
