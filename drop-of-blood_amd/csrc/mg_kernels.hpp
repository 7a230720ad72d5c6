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

#include <type_traits>

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
// (Acc: arithmetic type, FP64 on level 0 of the plain FP32 step as in
// k_residual_restrict_col)
template <class S, class Acc = typename compute_of<S>::type>
__global__ void __launch_bounds__(256)
k_residual(Dims d, const S *__restrict__ x, const S *__restrict__ b, S *__restrict__ r,
           typename compute_of<S>::type h2, typename compute_of<S>::type topc,
           const double *__restrict__ scale) {
  typedef Acc C;
  const C bscale = scale ? (C)scale[0] : C(1);
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int k = blockIdx.z;
  if (i >= d.nx || j >= d.ny) return;
  size_t c = idx3(d, i, j, k);
  C sn, dg;
  mg_stencil<S, C>(x, d, (C)topc, i, j, k, c, sn, dg);
  r[c] = store_as<S, C>((C)b[c] * bscale - (sn - dg * (C)x[c]) / (C)h2);
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

// MG_RESIDUAL followed by MG_RESTRICT (fluid.ts:744-800) in one pass, so the
// fine residual never goes to memory (same arithmetic as the two kernels),
// one thread per fine cell (coalesced 7-point loads): a 64 x 2 x 2 block of
// fine cells sums to 32 coarse cells,
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
  const int i = blockIdx.x * blockDim.x + tx, j = 2 * blockIdx.y + ty, k = 2 * blockIdx.z + tz;
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
    int ci = (blockIdx.x * blockDim.x + tx) >> 1;
    if (ci < dc.nx) {
      C s = sr[0][tx] + sr[1][tx] + sr[2][tx] + sr[3][tx];
      bc[idx3(dc, ci, blockIdx.y, blockIdx.z)] = store_as<Sc, C>(s * C(0.125));
    }
  }
}

// Residual-restriction by wave columns: each wavefront owns 64 x 2 fine
// cells (32 x-pairs by the two rows of one coarse row) and marches an even
// z-chunk with no LDS and no barrier.  The z neighbours of its cells stay in
// registers; x and y neighbours come from loads that hit the cache lines of
// this and the adjacent waves; the two rows of a coarse cell meet in one lane
// shuffle.  Loads run two planes ahead.  Same arithmetic as k_residual
// followed by k_restrict (the children summed z, then x, then y), except that
// the diagonal's terms are summed in another order (the same value on level
// 0, where all are whole numbers; within an ulp where the lid coefficient is
// not) and the division by h^2 is a product with its reciprocal (within an
// ulp).
//
// Acc is the arithmetic type of the residual and the restriction sum.  With
// FP32 storage on level 0 of the plain FP32 step it is FP64: the stored
// values convert exactly, the neighbour sum of a smooth field and diag x
// cancel to ~1/N^2 of their size, and an FP32 residual would carry rounding
// of ~eps |x| that the coarse levels amplify by 1/lambda_min in the smoothest
// mode (a relative pressure error of ~1e-2 at 256 x 384 x 256).  In FP64 the
// residual of the stored iterate is exact to rounding of its own size, so the
// cycles converge to the FP32-representable solution.
#define RC_WAVES 4

template <class Sf, class Sc, class Acc = typename compute_of<Sf>::type>
__global__ void __launch_bounds__(64 * RC_WAVES)
k_residual_restrict_col(Dims df, Dims dc, const Sf *__restrict__ x, const Sf *__restrict__ b,
                        Sc *__restrict__ bc, typename compute_of<Sf>::type h2,
                        typename compute_of<Sf>::type topc, const double *__restrict__ scale,
                        int cz, int ncx, int nwaves) {
  typedef Acc C;
  typedef typename vec2_of<Sf>::type V2;
  const int wv = __builtin_amdgcn_readfirstlane(xcd_remap(blockIdx.x, gridDim.x) * RC_WAVES +
                                                (int)(threadIdx.x >> 6));
  if (wv >= nwaves) return;
  const int lane = threadIdx.x & 63;
  const int nx = df.nx, ny = df.ny, nz = df.nz, ncy = ny >> 1;
  const int cx = wv % ncx, cy = (wv / ncx) % ncy, ch = wv / (ncx * ncy);
  const int k0 = ch * cz, k1 = min(k0 + cz, nz);
  const int gx = cx * 64 + 2 * (lane & 31), gy = 2 * cy + (lane >> 5);
  const bool inside = gx < nx;
  constexpr int es = (int)sizeof(Sf);
  // in-plane byte offsets of the pair and of its x-, x+, y-, y+ neighbours
  // (cells outside the grid load as 0)
  const int oc = inside ? (gy * nx + gx) * es : MAC3D_OFF_NONE;
  const int ol = inside && gx > 0 ? oc - es : MAC3D_OFF_NONE;
  const int orr = inside && gx + 2 < nx ? oc + 2 * es : MAC3D_OFF_NONE;
  const int od = inside && gy > 0 ? oc - nx * es : MAC3D_OFF_NONE;
  const int ou = inside && gy + 1 < ny ? oc + nx * es : MAC3D_OFF_NONE;
  // buffer ranges: x planes k0-1 .. k1, b planes k0 .. k1-1
  const int pb = (int)((size_t)nx * ny * es);
  const int kx0 = max(k0 - 1, 0), kx1 = min(k1 + 1, nz);
  const BufRsrc rx = buf_rsrc(x + (size_t)kx0 * nx * ny, (uint32_t)(kx1 - kx0) * pb);
  const BufRsrc rb = buf_rsrc(b + (size_t)k0 * nx * ny, (uint32_t)(k1 - k0) * pb);
  const C bscale = scale ? (C)scale[0] : C(1);
  const C rh2 = C(1) / (C)h2;
  // in-plane part of the diagonal of the pair's two cells
  const C top = (gy < ny - 1) ? C(1) : (C)topc;
  const C dga = top + C((gx > 0) + (gx < nx - 1) + (gy > 0));
  const C dgb = top + C((gx + 1 > 0) + (gx + 1 < nx - 1) + (gy > 0));
  // loads for plane p: x pair of plane p+1; x-, x+, y-, y+ and b of plane p
  struct Set { V2 cn, u, d, bb; Sf l, r; };
  auto load_set = [&](int p, Set &S) {
    const int zo = (p - kx0) * pb;
    S.cn = buf_load<V2>(rx, oc + zo + pb);
    S.l = buf_load<Sf>(rx, ol + zo);
    S.r = buf_load<Sf>(rx, orr + zo);
    S.d = buf_load<V2>(rx, od + zo);
    S.u = buf_load<V2>(rx, ou + zo);
    S.bb = buf_load<V2>(rb, oc + (p - k0) * pb);
  };
  V2 cm = buf_load<V2>(rx, oc + (k0 - 1 - kx0) * pb);  // plane k0-1 (0 below the grid)
  V2 cc = buf_load<V2>(rx, oc + (k0 - kx0) * pb);
  Set A, B;
  load_set(k0, A);
  load_set(k0 + 1, B);
  C sa = C(0), sb = C(0);
  auto plane = [&](int k, Set &S, auto ODDc) {
    constexpr bool odd = decltype(ODDc)::value;
    const C zt = C((k > 0) + (k < nz - 1));
    const C xa = (C)cc.x, xb = (C)cc.y;
    // the six neighbours in the stencil's order x-, x+, y-, y+, z-, z+
    C na = (C)S.l + xb;
    na += (C)S.d.x;
    na += (C)S.u.x;
    na += (C)cm.x;
    na += (C)S.cn.x;
    C nb = xa + (C)S.r;
    nb += (C)S.d.y;
    nb += (C)S.u.y;
    nb += (C)cm.y;
    nb += (C)S.cn.y;
    const C ra = (C)S.bb.x * bscale - (na - (dga + zt) * xa) * rh2;
    const C rbv = (C)S.bb.y * bscale - (nb - (dgb + zt) * xb) * rh2;
    cm = cc;
    cc = S.cn;
    load_set(k + 2, S);
    if (!odd) {
      sa = ra;
      sb = rbv;
    } else {
      sa += ra;
      sb += rbv;
      C s = sa + sb;
      s += __shfl_xor(s, 32, 64);
      if (lane < 32 && inside)
        bc[idx3(dc, gx >> 1, cy, k >> 1)] = store_as<Sc, C>(s * C(0.125));
    }
  };
  for (int k = k0; k < k1; k += 2) {
    plane(k, A, std::false_type{});
    plane(k + 1, B, std::true_type{});
  }
}

// Coarse correction value for MG_PROLONG_ADD (fluid.ts:806-855): coarse
// indices clamped at the walls, 0 above the top row.
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

// MG_PROLONG_ADD (fluid.ts:806-855): x_f += cell-centred trilinear
// interpolation of the coarse correction.  Fine cell i sits at coarse
// coordinate i/2 - 1/4.  One thread per 2 x 2 x 2 block of fine cells (the
// children of one coarse cell), which share a 3 x 3 x 3 coarse
// neighbourhood: child offset 0 along an axis takes 3/4 of the parent and
// 1/4 of the parent's lower neighbour, offset 1 takes 3/4 and 1/4 of the
// upper neighbour (cell-centred weights of fluid.ts:806-855; clamped at the
// walls, 0 above the top row).  A 32 x 8 block of coarse cells first stages
// its neighbourhood (34 x 10 x 3 values) in LDS, so each thread reads its 27
// from LDS rather than issuing 27 gathers.
#define PR_BX 32
#define PR_BY 8
// PRE: the thread's fine cells are loaded before the staging, so the two
// round trips overlap (grids of few blocks per CU, where no other block
// hides them; on large grids the early loads cost more than they save).
template <class Sc, class Sf, bool PRE>
__global__ void __launch_bounds__(PR_BX * PR_BY)
k_prolong_add_2x(Dims df, Dims dc, const Sc *__restrict__ ec, Sf *__restrict__ xf) {
  typedef typename compute_of<Sf>::type C;
  __shared__ C se[3][PR_BY + 2][PR_BX + 2];
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * PR_BX + tx;
  const int I0 = blockIdx.x * PR_BX, J0 = blockIdx.y * PR_BY, K = blockIdx.z;
  const int I = I0 + tx, J = J0 + ty;
  const bool in = I < dc.nx && J < dc.ny;
  struct alignas(2 * sizeof(Sf)) P2 { Sf a, b; };
  P2 fv[2][2];
  if (PRE && in) {
#pragma unroll
    for (int oz = 0; oz < 2; ++oz)
#pragma unroll
      for (int oy = 0; oy < 2; ++oy)
        fv[oz][oy] = *(const P2 *)(xf + idx3(df, 2 * I, 2 * J + oy, 2 * K + oz));
  }
  for (int t = tid; t < 3 * (PR_BY + 2) * (PR_BX + 2); t += PR_BX * PR_BY) {
    const int q = t % (PR_BX + 2), r = (t / (PR_BX + 2)) % (PR_BY + 2);
    const int a = t / ((PR_BX + 2) * (PR_BY + 2));
    se[a][r][q] = (C)coarse_at(ec, dc, I0 + q - 1, J0 + r - 1, K + a - 1);
  }
  __syncthreads();
  if (!in) return;
  C e[3][3][3];  // [dz][dy][dx], offsets -1..1
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int bq = 0; bq < 3; ++bq)
#pragma unroll
      for (int q = 0; q < 3; ++q) e[a][bq][q] = se[a][ty + bq][tx + q];
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
      // the two x-children as one aligned pair (2I is even)
      P2 *px = (P2 *)(xf + idx3(df, 2 * I, 2 * J + oy, 2 * K + oz));
      P2 f = PRE ? fv[oz][oy] : *px;
      C c0 = wl[0] * yz[0] + wm * yz[1] + wu[0] * yz[2];
      C c1 = wl[1] * yz[0] + wm * yz[1] + wu[1] * yz[2];
      f.a = store_as<Sf, C>((C)f.a + c0);
      f.b = store_as<Sf, C>((C)f.b + c1);
      *px = f;
    }
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

// The 1-D eigenvector matrices are read nx (or ny, nz) times per output, so
// with QL they are first copied to LDS (dynamic shared memory, converted to
// the compute type C, as every use converts them): the dot products then
// chain LDS reads rather than global loads.  The host passes QL when they
// fit in 48 KB.
//
// The three kernels run COARSE_T threads per block.  Each stages its inputs
// (the eigenvector matrices and one plane) with every thread's first loads
// issued together before its LDS stores, so the prologue costs one memory
// round trip rather than one per array (arrays of up to COARSE_T values, as
// at 24 x 32 x 24, need no more; larger ones copy the rest afterwards).
#define COARSE_T 1024

// a thread's first element of the block-wide copy of n >= 1 values f(t),
// loaded and stored with no branch (the index clamped to n - 1: threads past
// n store element n - 1's own value there), so that the compiler keeps the
// loads of all the copies together; the elements past COARSE_T follow in a loop
template <class C> struct CoarseFirst {
  C v;
  template <class F> __device__ void load(int n, F f) { v = f(min((int)threadIdx.x, n - 1)); }
  template <class F> __device__ void store(C *dst, int n, F f) const {
    dst[min((int)threadIdx.x, n - 1)] = v;
    for (int t = (int)threadIdx.x + COARSE_T; t < n; t += COARSE_T) dst[t] = f(t);
  }
};

// sum of f(0) .. f(n-1) in four interleaved partial sums (a quarter of the
// dependent chain of one running sum), combined pairwise
template <class F>
__device__ inline auto dot4(int n, F f) -> decltype(f(0)) {
  typedef decltype(f(0)) C;
  C a0 = C(0), a1 = C(0), a2 = C(0), a3 = C(0);
  int m = 0;
  for (; m + 4 <= n; m += 4) {
    a0 += f(m);
    a1 += f(m + 1);
    a2 += f(m + 2);
    a3 += f(m + 3);
  }
  for (; m < n; ++m) a0 += f(m);
  return (a0 + a1) + (a2 + a3);
}

template <class S, class C, bool QL>
__global__ void __launch_bounds__(COARSE_T)
k_coarse_xz_fwd(Dims d, const S *__restrict__ b, const double *__restrict__ qx,
                const double *__restrict__ qz, double *__restrict__ work) {
  __shared__ C s0[COARSE_MAX_PLANE], s1[COARSE_MAX_PLANE];
  extern __shared__ unsigned char dyn_q[];
  int j = blockIdx.x;
  int nx = d.nx, nz = d.nz, np = nx * nz;
  C *const lqx = (C *)dyn_q, *const lqz = lqx + nx * nx;
  auto QX = [&](int r) { return QL ? lqx[r] : (C)qx[r]; };
  auto QZ = [&](int r) { return QL ? lqz[r] : (C)qz[r]; };
  auto fx = [&](int t) { return (C)qx[t]; };
  auto fz = [&](int t) { return (C)qz[t]; };
  auto fb = [&](int t) { return (C)b[idx3(d, t % nx, j, t / nx)]; };
  CoarseFirst<C> hx, hz, hb;
  if (QL) {
    hx.load(nx * nx, fx);
    hz.load(nz * nz, fz);
  }
  hb.load(np, fb);
  if (QL) {
    hx.store(lqx, nx * nx, fx);
    hz.store(lqz, nz * nz, fz);
  }
  hb.store(s0, np, fb);
  __syncthreads();
  // s1[k][a] = sum_m Qx[m][a] s0[k][m]
  for (int t = threadIdx.x; t < np; t += blockDim.x) {
    int a = t % nx, k = t / nx;
    s1[t] = dot4(nx, [&](int m) { return QX(m * nx + a) * s0[k * nx + m]; });
  }
  __syncthreads();
  // out[c][a] = sum_n Qz[n][c] s1[n][a]
  for (int t = threadIdx.x; t < np; t += blockDim.x) {
    int a = t % nx, cc = t / nx;
    work[idx3(d, a, j, cc)] = (double)dot4(nz, [&](int n) { return QZ(n * nz + cc) * s1[n * nx + a]; });
  }
}

template <class C, bool QL>
__global__ void __launch_bounds__(COARSE_T)
k_coarse_y(Dims d, const double *__restrict__ qy, const double *__restrict__ lx,
           const double *__restrict__ ly, const double *__restrict__ lz, double h2,
           double *__restrict__ work) {
  __shared__ C s0[COARSE_MAX_PLANE], s1[COARSE_MAX_PLANE];
  extern __shared__ unsigned char dyn_q[];
  int cz = blockIdx.x;
  int nx = d.nx, ny = d.ny, np = nx * ny;
  C *const lqy = (C *)dyn_q;
  auto QY = [&](int r) { return QL ? lqy[r] : (C)qy[r]; };
  auto fy = [&](int t) { return (C)qy[t]; };
  auto fw = [&](int t) { return (C)work[idx3(d, t % nx, t / nx, cz)]; };
  // the eigenvalue sum of output t = a + nx bb
  auto fe = [&](int t) { return (C)(lx[t % nx] + ly[t / nx] + lz[cz]); };
  CoarseFirst<C> hy, hw, he;
  if (QL) hy.load(ny * ny, fy);
  hw.load(np, fw);
  he.load(np, fe);
  if (QL) hy.store(lqy, ny * ny, fy);
  hw.store(s0, np, fw);
  __syncthreads();
  // s1[bb][a] = (sum_m Qy[m][bb] s0[m][a]) / (lx_a + ly_bb + lz_cz) * (-h2)
  for (int t = threadIdx.x; t < np; t += COARSE_T) {
    int a = t % nx, bb = t / nx;
    C acc = dot4(ny, [&](int m) { return QY(m * ny + bb) * s0[m * nx + a]; });
    s1[t] = acc * (C)(-h2) / (t == (int)threadIdx.x ? he.v : fe(t));
  }
  __syncthreads();
  // back: work[m][a] = sum_bb Qy[m][bb] s1[bb][a]
  for (int t = threadIdx.x; t < np; t += blockDim.x) {
    int a = t % nx, m = t / nx;
    work[idx3(d, a, m, cz)] = (double)dot4(ny, [&](int bb) { return QY(m * ny + bb) * s1[bb * nx + a]; });
  }
}

template <class S, class C, bool QL>
__global__ void __launch_bounds__(COARSE_T)
k_coarse_xz_inv(Dims d, const double *__restrict__ work, const double *__restrict__ qx,
                const double *__restrict__ qz, S *__restrict__ x) {
  __shared__ C s0[COARSE_MAX_PLANE], s1[COARSE_MAX_PLANE];
  extern __shared__ unsigned char dyn_q[];
  int j = blockIdx.x;
  int nx = d.nx, nz = d.nz, np = nx * nz;
  C *const lqx = (C *)dyn_q, *const lqz = lqx + nx * nx;
  auto QX = [&](int r) { return QL ? lqx[r] : (C)qx[r]; };
  auto QZ = [&](int r) { return QL ? lqz[r] : (C)qz[r]; };
  auto fx = [&](int t) { return (C)qx[t]; };
  auto fz = [&](int t) { return (C)qz[t]; };
  auto fw = [&](int t) { return (C)work[idx3(d, t % nx, j, t / nx)]; };
  CoarseFirst<C> hx, hz, hw;
  if (QL) {
    hx.load(nx * nx, fx);
    hz.load(nz * nz, fz);
  }
  hw.load(np, fw);
  if (QL) {
    hx.store(lqx, nx * nx, fx);
    hz.store(lqz, nz * nz, fz);
  }
  hw.store(s0, np, fw);
  __syncthreads();
  // s1[n][a] = sum_c Qz[n][c] s0[c][a]
  for (int t = threadIdx.x; t < np; t += blockDim.x) {
    int a = t % nx, n = t / nx;
    s1[t] = dot4(nz, [&](int cc) { return QZ(n * nz + cc) * s0[cc * nx + a]; });
  }
  __syncthreads();
  // x[n][m] = sum_a Qx[m][a] s1[n][a]
  for (int t = threadIdx.x; t < np; t += blockDim.x) {
    int m = t % nx, n = t / nx;
    x[idx3(d, m, j, n)] = store_as<S, C>(dot4(nx, [&](int a) { return QX(m * nx + a) * s1[n * nx + a]; }));
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
#ifndef SW_TX
#define SW_TX 64
#endif
#ifndef SW_TY
#define SW_TY 16
#endif
#define SW_EX (SW_TX + 4)                 // staged row, two-cell ring
#define SW_EY (SW_TY + 4)
#define SW_THREADS (SW_TX / 2 * SW_TY)  // one black pair per thread
#define SW_NV (SW_EX / 2 * SW_EY)         // 680 staged pairs per plane
#define SW_LPT ((SW_NV + SW_THREADS - 1) / SW_THREADS)   // 2
#define SW_RP ((SW_TX + 2) / 2 * (SW_TY + 2))           // 594 red pairs
#define SW_RPT ((SW_RP + SW_THREADS - 1) / SW_THREADS)  // 2

__device__ inline float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline double fast_rcp(double x) { return 1.0 / x; }

// Red-black update of one cell, shared by the level-0 sweep kernels so that
// one sweep and the two-sweep pass round identically: x' = (neighbour sum -
// h^2 b) / diag (MG_RBGS_RED / MG_RBGS_BLACK, fluid.ts:645-681), the sum in
// the stencil's order x-, x+, y-, y+, z-, z+ with 0 for neighbours outside
// the grid, then x + omega (x' - x) when omega != 1.
template <class C>
__device__ inline C rb_cell_r(C xm, C xp, C ym, C yp, C zm, C zp, C bv, C old, C h2, C rdg,
                              C omega) {
  C sn = xm + xp;
  sn += ym;
  sn += yp;
  sn += zm;
  sn += zp;
  C v = (sn - h2 * bv) * rdg;
  if (omega != C(1)) v = old + omega * (v - old);
  return v;
}
// the same with omega known to be 1 at compile time (no blend to select from)
template <class C>
__device__ inline C rb_cell_r1(C xm, C xp, C ym, C yp, C zm, C zp, C bv, C h2, C rdg) {
  C sn = xm + xp;
  sn += ym;
  sn += yp;
  sn += zm;
  sn += zp;
  return (sn - h2 * bv) * rdg;
}
template <class C>
__device__ inline C rb_cell(C xm, C xp, C ym, C yp, C zm, C zp, C bv, C old, C h2, C dg,
                            C omega) {
  return rb_cell_r<C>(xm, xp, ym, yp, zm, zp, bv, old, h2, fast_rcp(dg), omega);
}
// diagonal of evalMgStencil (fluid.ts:561-603) at (gx, gy, k): one per
// neighbour inside the grid, the lid coefficient topc at the top row
template <class C>
__device__ inline C rb_diag(const Dims &d, int gx, int gy, int k, C topc) {
  return ((gy < d.ny - 1) ? C(1) : topc) +
         C((gx > 0) + (gx < d.nx - 1) + (gy > 0) + (k > 0) + (k < d.nz - 1));
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
  int st_o[SW_LPT];  // in-plane byte offset, MAC3D_OFF_NONE outside the grid or the stage
#pragma unroll
  for (int q = 0; q < SW_LPT; ++q) {
    int v = tid + q * SW_THREADS;
    st_r[q] = v / (SW_EX / 2);
    st_c[q] = 2 * (v % (SW_EX / 2));
    int gx = x0 - 2 + st_c[q], gy = y0 - 2 + st_r[q];
    st_o[q] = (v < SW_NV && gx >= 0 && gx < nx && gy >= 0 && gy < ny)
                  ? (gy * nx + gx) * (int)sizeof(S) : MAC3D_OFF_NONE;
  }
  // The planes this chunk reads, as buffer ranges: xin k0-2 .. k1+1, b
  // k0-1 .. k1 (clipped to the grid).  Planes outside them (the grid's z
  // walls, and the prefetches past the chunk's end) load as zeros.
  const int pb = (int)(plane * sizeof(S));
  const int kx0 = max(k0 - 2, 0), kx1 = min(k1 + 2, nz);
  const int kb0 = max(k0 - 1, 0), kb1 = min(k1 + 1, nz);
  const BufRsrc rx = buf_rsrc(xin + (size_t)kx0 * plane, (uint32_t)(kx1 - kx0) * pb);
  const BufRsrc rb = buf_rsrc(b + (size_t)kb0 * plane, (uint32_t)(kb1 - kb0) * pb);
  auto fetch = [&](BufRsrc r, int kbase, int k, V2 (&pre)[SW_LPT]) {
#pragma unroll
    for (int q = 0; q < SW_LPT; ++q) pre[q] = buf_load<V2>(r, st_o[q] + (k - kbase) * pb);
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
  // grid (rgx, rgy), stage (rr, rc); rv false for pairs past the ring (parked
  // on stage (1, 1), computed and dropped) and rows outside the grid
  int rr[SW_RPT], rc[SW_RPT], rgx[SW_RPT], rgy[SW_RPT];
  bool rv[SW_RPT];
#pragma unroll
  for (int q = 0; q < SW_RPT; ++q) {
    int v = tid + q * SW_THREADS;
    int row = v / ((SW_TX + 2) / 2), pc = v % ((SW_TX + 2) / 2);
    if (v >= SW_RP) row = pc = 0;
    rr[q] = row + 1;
    rc[q] = 2 * pc + 1;
    rgx[q] = x0 - 1 + 2 * pc;
    rgy[q] = y0 - 1 + row;
    rv[q] = v < SW_RP && rgy[q] >= 0 && rgy[q] < ny;
  }
  // Stencil update of the cell at stage (r, c), grid (gx, gy, k).  Staged
  // values outside the grid are zero (out-of-range loads, never updated),
  // so the six neighbours are summed without branches, in the order of the
  // reference stencil; only the diagonal depends on where the cell is.
  // (nbrs false: red cells of a zero-guess sweep, whose neighbours are 0.)
  auto update = [&](int r, int c, int gx, int gy, int k, bool nbrs) {
    C z = C(0);
    return rb_cell<C>(nbrs ? sp[k & 3][r][c - 1] : z, nbrs ? sp[k & 3][r][c + 1] : z,
                      nbrs ? sp[k & 3][r - 1][c] : z, nbrs ? sp[k & 3][r + 1][c] : z,
                      nbrs ? sp[(k - 1) & 3][r][c] : z, nbrs ? sp[(k + 1) & 3][r][c] : z,
                      sb[k & 1][r][c], ZERO ? z : sp[k & 3][r][c], h2,
                      rb_diag<C>(d, gx, gy, k, topc), omega);
  };
  auto red = [&](int k) {
    if (k < 0 || k >= nz) return;
#pragma unroll
    for (int q = 0; q < SW_RPT; ++q) {
      int w = (rgx[q] + rgy[q] + k) & 1;  // 0: left cell red, 1: right cell red
      int gx = rgx[q] + w;
      C v = update(rr[q], rc[q] + w, gx, rgy[q], k, !ZERO);
      if (rv[q] && gx >= 0 && gx < nx) sp[k & 3][rr[q]][rc[q] + w] = v;
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
  // ZERO: the x registers stay 0 and zeros are staged (the black update
  // sums all six neighbours, red ones outside the grid included)
  auto iter = [&](int k, V2 (&pr)[SW_LPT], V2 (&br_)[SW_LPT]) {
    stage_p(k + 2, pr);
    stage_b(k + 1, br_);
    __syncthreads();
    if (!ZERO) fetch(rx, kx0, k + 4, pr);
    fetch(rb, kb0, k + 3, br_);
    red(k + 1);
    __syncthreads();
    black_and_write(k);
    __syncthreads();
  };

  // prologue: planes k0-2 .. k0+1 of xin, b planes k0-1, k0; red on k0-1, k0
  {
    V2 t0[SW_LPT], t1[SW_LPT], t2[SW_LPT], t3[SW_LPT], u0[SW_LPT], u1[SW_LPT];
#pragma unroll
    for (int q = 0; q < SW_LPT; ++q) {
      t0[q].x = t0[q].y = t1[q].x = t1[q].y = S(0);
      t2[q].x = t2[q].y = t3[q].x = t3[q].y = S(0);
      pA[q].x = pA[q].y = pB[q].x = pB[q].y = S(0);
    }
    if (!ZERO) {
      fetch(rx, kx0, k0 - 2, t0);
      fetch(rx, kx0, k0 - 1, t1);
      fetch(rx, kx0, k0, t2);
      fetch(rx, kx0, k0 + 1, t3);
    }
    fetch(rb, kb0, k0 - 1, u0);
    fetch(rb, kb0, k0, u1);
    stage_p(k0 - 2, t0);
    stage_p(k0 - 1, t1);
    stage_p(k0, t2);
    stage_p(k0 + 1, t3);
    stage_b(k0 - 1, u0);
    stage_b(k0, u1);
  }
  // (issued in the order the loop issues them, so that the loop's waits
  // count the same loads in flight on entry as on every later pass)
  if (!ZERO) fetch(rx, kx0, k0 + 2, pA);
  fetch(rb, kb0, k0 + 1, bA);
  if (!ZERO) fetch(rx, kx0, k0 + 3, pB);
  fetch(rb, kb0, k0 + 2, bB);
  __syncthreads();
  red(k0 - 1);
  red(k0);
  __syncthreads();
  // both halves unconditional in the loop body, the odd plane after it
  int k = k0;
  for (; k + 1 < k1; k += 2) {
    iter(k, pA, bA);
    iter(k + 1, pB, bB);
  }
  if (k < k1) iter(k, pA, bA);
}


// ---------------------------------------------------------------------------
// Two red-black sweeps in one pass (temporal blocking), FP32 levels:
// R1, B1 (sweep 1) and R2, B2 (sweep 2) march through z together, so x and b
// are read once and x written once for two sweeps (12 B per FP32 cell per
// pass, against 24 for two passes of k_rbgs_sweep).  A block of TX / 2 * TY
// threads (one tile pair each) owns a TX x TY column tile and a z-chunk
// [k0, k1): 64 x 16, 48 x 16 on grids 48-wide tiles fit better, or 32 x 16 on
// small grids (twice the blocks).  Each colour stage is recomputed on a ring one cell wider than the next stage reads (R1 on the
// tile plus three cells, B1 plus two, R2 plus one, B2 on the tile), so no
// block waits on another and the result is exactly two sweeps, with the cell
// arithmetic of k_rbgs_sweep (rb_cell).
//
// LDS holds planes of x in a 7-slot ring (tile plus four cells, 72 x 24 for
// 64 x 16) and b in a 5-slot ring (tile plus three cells, 72 x 22), each row
// split by colour: [row][colour][pair], pair = (even, odd) x-pair.  A cell's
// six neighbours are of the other colour: in x the other cell of its pair P
// and a cell of pair P-1 or P+1 (by the cell's side), in y and z the cells at
// pair P of the rows and planes either side, so every stage reads LDS at
// stride one.  The z neighbours of a stage's first pair stay in registers
// from one plane to the next.  Iteration j of the march runs three groups
// with a barrier after each:
//   B2(j) (reads slots j-1..j+1, writes plane j of xout) with R1(j+4) (reads
//     slots j+3..j+5, writes red of j+4 and b(j+4) into its b slot);
//   B1(j+3);
//   R2(j+2) with the staging of x(j+6).
// An x slot lives from its staging to B2 of the plane above, seven
// iterations; a b slot from R1 to B2 of its plane, five.  b(j+6) and x(j+8)
// go into registers two iterations before their use, by buffer loads whose
// ranges cover only the planes this chunk reads: planes outside the grid (and
// past the chunk) load as zeros with no branch.  Each stage takes the pairs
// covering its ring and updates its colour's cell when that cell lies in the
// ring and the grid.  ZERO: zero initial guess (x is not read, zeros are
// staged).  scale (nullable): the right-hand side is b scale[0].  Storage S:
// FP32 or FP16 (widened to FP32 as it is staged; the stages compute and keep
// FP32 in LDS and only the written plane is rounded to FP16); an FP64 ring
// would not fit two blocks' LDS.
#define TB_XR 7
#define TB_BR 5

// W1: omega is 1 (the plain Gauss-Seidel sweep), so no over-relaxation blend.
template <bool ZERO, int TX, int TY, class S, bool W1>
__global__ void __launch_bounds__(TX / 2 * TY, TX == 64 ? 4 : 3)
k_rbgs_sweep2(Dims d, const S *__restrict__ xin, S *__restrict__ xout,
              const S *__restrict__ b, float h2, float topc, float omega,
              const double *__restrict__ scale, int cz, int ncx, int ncy) {
  constexpr int TB_TX = TX, TB_TY = TY;
  constexpr int TB_T = TX / 2 * TY;
  constexpr int TB_PW = (TX + 8) / 2;   // pairs per staged row: four-cell ring
  constexpr int TB_EY = TY + 8;
  constexpr int TB_BY = TY + 6;         // b rows: three-cell ring
  constexpr int CS = TB_PW * 4;         // byte offset of a row's black pairs
  constexpr int RS = 2 * CS;            // bytes per staged row, both colours
  typedef float C;
  typedef float2 V2;
  static_assert(TB_T == TB_TX / 2 * TB_TY, "one tile pair per thread");
  static_assert(TB_PW * TB_EY <= 2 * TB_T && TB_PW * TB_BY <= 2 * TB_T, "two pairs per thread");
  __shared__ C lds[(TB_XR * TB_EY + TB_BR * TB_BY) * 2 * TB_PW];
  char *const L = (char *)lds;
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  const int lb = xcd_remap(blockIdx.x, gridDim.x);
  const int x0 = (lb % ncx) * TB_TX, y0 = ((lb / ncx) % ncy) * TB_TY;
  const int k0 = (lb / (ncx * ncy)) * cz;  // even: cz is even
  const int k1 = min(k0 + cz, nz);
  const int tid = threadIdx.x;
  const C bs = scale ? (C)scale[0] : C(1);
  const int pb = (int)((size_t)nx * ny * sizeof(S));

  // pairs of the x staging (tile plus four cells) and of R1 (rows of the
  // three-cell ring, the same columns): in-plane byte offsets, and the
  // staged pair's byte offset in a slot (red cell; black CS further)
  int xo[2], bo[2], so[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    int v = tid + q * TB_T;
    int row = v / TB_PW, pc = v % TB_PW;
    int gx = x0 - 4 + 2 * pc, gy = y0 - 4 + row;
    so[q] = row * RS + 4 * pc;
    xo[q] = (v < TB_PW * TB_EY && gx >= 0 && gx < nx && gy >= 0 && gy < ny)
                ? (gy * nx + gx) * (int)sizeof(S) : MAC3D_OFF_NONE;
    gy += 1;
    bo[q] = (v < TB_PW * TB_BY && gx >= 0 && gx < nx && gy >= 0 && gy < ny)
                ? (gy * nx + gx) * (int)sizeof(S) : MAC3D_OFF_NONE;
  }
  const int kx0 = max(k0 - 4, 0), kx1 = min(k1 + 4, nz);
  const int kb0 = max(k0 - 3, 0), kb1 = min(k1 + 3, nz);
  const BufRsrc rx = buf_rsrc(xin + (size_t)kx0 * nx * ny, (uint32_t)(kx1 - kx0) * pb);
  const BufRsrc rb = buf_rsrc(b + (size_t)kb0 * nx * ny, (uint32_t)(kb1 - kb0) * pb);
  auto fetch = [&](BufRsrc r, int kbase, int k, const int (&off)[2], V2 (&pre)[2]) {
#pragma unroll
    for (int q = 0; q < 2; ++q) pre[q] = buf_load_pair<S>(r, off[q] + (k - kbase) * pb);
  };

  // ---- per-thread tables, fixed over the march ----------------------------
  // Stage slots: 0 = B2 (the thread's tile pair), 1-2 = R1, 3-4 = B1, 5-6 =
  // R2 (pairs tid and tid + TB_T of the stage's ring).  Per slot: ob, the
  // byte offset in a slot of the pair's red cell one row up (the pair's own
  // row starts RS further), and per plane parity fo, the byte offset of the
  // cell's x neighbour in the next pair over (pair - 1 for a cell on the
  // left, + 1 on the right); 1/diag on planes with both z neighbours in the
  // grid (srd) and on planes 0 and nz - 1 (srb); bits 4 st + 2 par (cell
  // valid) and 4 st + 2 par + 1 (cell on the right of its pair).  Bits 28 +
  // q: staged pair q has red on the right at even planes; bit 30: staged
  // pair 1 exists; bit 31: R1 pair 1 exists.
  int ob[7], fo[2][7];
  C srd[2][7], srb[2][7];
  unsigned bits = 0;
  auto setup = [&](int st, int W, int X, int v, int NP) {
    int PAD = (W + 1) & ~1, PC = TB_TX / 2 + PAD;
    // lanes past the stage's pairs take pair 0's addresses (they store
    // nothing, and their reads stay inside the ring)
    const bool act = v < NP;
    const int row = act ? v / PC : 0, pc = act ? v % PC : 0;
    int gy = y0 - W + row, gxl = x0 - PAD + 2 * pc;
    int xr = row + 4 - W, pcol = pc + (4 - PAD) / 2;  // staged row and pair column
    ob[st] = (xr - 1) * RS + 4 * pcol;
#pragma unroll
    for (int par = 0; par < 2; ++par) {
      int sd = (gy + par + X) & 1, gx = gxl + sd;
      fo[par][st] = xr * RS + (1 - X) * CS + 4 * (pcol + (sd ? 1 : -1));
      bool ok = act && gx >= x0 - W && gx < x0 + TB_TX + W && gx >= 0 && gx < nx && gy >= 0 &&
                gy < ny;
      // rb_diag with both z neighbours in the grid (srd) and with one (srb,
      // planes 0 and nz - 1)
      C top = (gy < ny - 1) ? C(1) : topc;
      int nxy = (gx > 0) + (gx < nx - 1) + (gy > 0);
      srd[par][st] = fast_rcp(top + C(nxy + 2));
      srb[par][st] = fast_rcp(top + C(nxy + 1));
      bits |= (ok ? 1u : 0u) << (4 * st + 2 * par);
      bits |= (unsigned)sd << (4 * st + 2 * par + 1);
    }
  };
  setup(0, 0, 1, tid, TB_T);
  setup(1, 3, 0, tid, TB_PW * TB_BY);
  setup(2, 3, 0, tid + TB_T, TB_PW * TB_BY);
  setup(3, 2, 1, tid, (TB_TX / 2 + 2) * (TB_TY + 4));
  setup(4, 2, 1, tid + TB_T, (TB_TX / 2 + 2) * (TB_TY + 4));
  setup(5, 1, 0, tid, (TB_TX / 2 + 2) * (TB_TY + 2));
  setup(6, 1, 0, tid + TB_T, (TB_TX / 2 + 2) * (TB_TY + 2));
#pragma unroll
  for (int q = 0; q < 2; ++q) bits |= (unsigned)((y0 - 4 + (tid + q * TB_T) / TB_PW) & 1) << (28 + q);
  if (tid + TB_T < TB_PW * TB_EY) bits |= 1u << 30;
  if (tid + TB_T < TB_PW * TB_BY) bits |= 1u << 31;
  // Second pairs of R1, B1 and R2 exist for the first threads only (280, 168
  // and 100 of 512 on a 64 x 16 tile): whole waves past them skip those
  // updates (a scalar branch; their lanes write nothing).
  const int wb = __builtin_amdgcn_readfirstlane(tid & ~63) + TB_T;
  const bool w2 = wb < TB_PW * TB_BY;
  const bool w4 = wb < (TB_TX / 2 + 2) * (TB_TY + 4);
  const bool w6 = wb < (TB_TX / 2 + 2) * (TB_TY + 2);
  const int tr = tid / (TB_TX / 2), tpc = tid % (TB_TX / 2);
  const int tgy = y0 + tr, tgx = x0 + 2 * tpc;
  const bool tin = tgx < nx && tgy < ny;
  S *const tout = xout + (size_t)tgy * nx + tgx;

  // ---- LDS addressing -------------------------------------------------------
  // A slot holds one plane, row by row: TB_PW red pairs, then TB_PW black
  // ones (row-interleaved), so the cell of a stage pair, the other colour's
  // cell of the pair and that colour's cells one row down and up lie at fixed
  // offsets from one base (ob), within the reach of the LDS instructions'
  // immediate offsets.  x slots come first, then b slots, whose rows start
  // one row lower than the x slot's (so ob addresses b too).
  constexpr int XSL = TB_EY * RS, BSL = TB_BY * RS, BOFF = TB_XR * XSL;
  auto ld = [&](int a) { return *(const C *)(L + a); };
  auto st_ = [&](int a, C v) { *(C *)(L + a) = v; };
  // all-ones where bit k of bits is set (a sign-extending bit extract), and
  // a select by such a mask: no predicate registers
  auto mask = [&](int k) { return ((int)(bits << (31 - k))) >> 31; };
  auto pick = [](int m, C a, C b) {  // m ? a : b
    return __int_as_float((__float_as_int(a) & m) | (__float_as_int(b) & ~m));
  };
  // the other colour's cell of stage slot st's pair in the slot at byte Sb
  // (X: the stage's colour)
  auto oc_at = [&](int st, int Sb, int X) { return ld(Sb + ob[st] + RS + (1 - X) * CS); };
  // new value of the cell (colour X) of stage slot st at plane parity par in
  // slot Sb, given the other colour's cell of its pair (sm) and its z
  // neighbours; the x neighbours add in either order (the sum is the same)
  auto upd = [&](int st, int par, bool bnd, int Sb, int X, C sm, C zm, C zp, C bv) -> C {
    const int a = Sb + ob[st], oc = (1 - X) * CS;
    const C ym = ld(a + oc), yp = ld(a + 2 * RS + oc), fx = ld(Sb + fo[par][st]);
    const C rdg = bnd ? srb[par][st] : srd[par][st];
    if (W1) return rb_cell_r1<C>(sm, fx, ym, yp, zm, zp, bv, h2, rdg);
    return rb_cell_r<C>(sm, fx, ym, yp, zm, zp, bv, ld(a + RS + X * CS), h2, rdg, omega);
  };
  // store of a stage cell: only cells in the stage's ring and the grid
  // (lanes past the stage's pairs write nothing)
  auto put = [&](int st, int par, int Sb, int X, C v) {
    if ((bits >> (4 * st + 2 * par)) & 1u) st_(Sb + ob[st] + RS + X * CS, v);
  };

  // the prologue's staging (plane p at x ring position (p - (k0 - 4)) mod 7)
  const int px = k0 - 4;
  auto stage_x_g = [&](int p, const V2 (&pre)[2]) {
    const int Sb = ((p - px) % TB_XR) * XSL;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      int v = tid + q * TB_T;
      if (v < TB_PW * TB_EY) {
        bool rs = (y0 - 4 + v / TB_PW + p) & 1;
        st_(Sb + so[q], rs ? pre[q].y : pre[q].x);
        st_(Sb + so[q] + CS, rs ? pre[q].x : pre[q].y);
      }
    }
  };

  // ---- the march --------------------------------------------------------------
  const int r1lo = max(k0 - 3, 0), r1hi = min(k1 + 3, nz);
  const int b1lo = max(k0 - 2, 0), b1hi = min(k1 + 2, nz);
  const int r2lo = max(k0 - 1, 0), r2hi = min(k1 + 1, nz);
  // iteration j (parity PAR, a compile-time constant); xsl[d] is the x slot
  // base of plane j - 1 + d (plane j + 6 reuses xsl[0]), bsl[d] the b slot
  // base of plane j + d
  int xsl[7], bsl[5];
  // z carries of the first pair of each stage (B2, R1, B1, R2): the other
  // colour's cell of the pair on the stage's plane (cm) and on the plane
  // above (cs) are that cell's z- neighbour and pair cell one iteration later
  // (no stage writes them in between), so unchecked iterations read only the
  // plane above from LDS
  C cm[4], cs[4];
  // CHK false: a steady-state iteration, in which every stage is in range and
  // no plane is 0 or nz - 1 (the checks compile away), and the stages ran in
  // the iteration before (the carries are set)
  auto iter = [&](auto PARc, auto CHKc, int j, V2 (&xr)[2], V2 (&br)[2]) {
    constexpr int par = decltype(PARc)::value;
    constexpr bool CHK = decltype(CHKc)::value;
    auto edge = [&](int p) { return CHK && (p == 0 || p == nz - 1); };
    // group 1: B2(j), R1(j + 4)
    if (!CHK || j >= k0) {
      const C zm = CHK ? oc_at(0, xsl[0], 1) : cm[0];
      const C sm = CHK ? oc_at(0, xsl[1], 1) : cs[0];
      const C zp = oc_at(0, xsl[2], 1);
      const C vb = upd(0, par, edge(j), xsl[1], 1, sm, zm, zp, ld(bsl[0] + ob[0] + CS));
      cm[0] = sm;
      cs[0] = zp;
      if (tin) {
        int rt = mask(2 * par + 1);
        V2 o;
        o.x = pick(rt, sm, vb);
        o.y = pick(rt, vb, sm);
        store_pair(tout + (size_t)j * nx * ny, o);
      }
    }
    if (!CHK || (j + 4 >= r1lo && j + 4 < r1hi)) {
      const bool bnd = edge(j + 4);
      C bred[2], bblk[2];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        C b0 = br[q].x * bs, b1 = br[q].y * bs;
        int rt = mask(4 * (1 + q) + 2 * par + 1);
        bred[q] = pick(rt, b1, b0);
        bblk[q] = pick(rt, b0, b1);
      }
      const C zm = CHK ? oc_at(1, xsl[4], 0) : cm[1];
      const C sm = CHK ? oc_at(1, xsl[5], 0) : cs[1];
      const C zp = oc_at(1, xsl[6], 0);
      const C nv0 = upd(1, par, bnd, xsl[5], 0, sm, zm, zp, bred[0]);
      cm[1] = sm;
      cs[1] = zp;
      C nv1 = C(0);
      if (w2)
        nv1 = upd(2, par, bnd, xsl[5], 0, oc_at(2, xsl[5], 0), oc_at(2, xsl[4], 0),
                  oc_at(2, xsl[6], 0), bred[1]);
      put(1, par, xsl[5], 0, nv0);
      if (w2) put(2, par, xsl[5], 0, nv1);
      st_(bsl[4] + ob[1], bred[0]);
      st_(bsl[4] + ob[1] + CS, bblk[0]);
      if (bits >> 31) {
        st_(bsl[4] + ob[2], bred[1]);
        st_(bsl[4] + ob[2] + CS, bblk[1]);
      }
    }
    fetch(rb, kb0, j + 6, bo, br);
    __syncthreads();
    // group 2: B1(j + 3)
    if (!CHK || (j + 3 >= b1lo && j + 3 < b1hi)) {
      constexpr int pp = par ^ 1;
      const bool bnd = edge(j + 3);
      const C zm = CHK ? oc_at(3, xsl[3], 1) : cm[2];
      const C sm = CHK ? oc_at(3, xsl[4], 1) : cs[2];
      const C zp = oc_at(3, xsl[5], 1);
      const C nv0 = upd(3, pp, bnd, xsl[4], 1, sm, zm, zp, ld(bsl[3] + ob[3] + CS));
      cm[2] = sm;
      cs[2] = zp;
      C nv1 = C(0);
      if (w4)
        nv1 = upd(4, pp, bnd, xsl[4], 1, oc_at(4, xsl[4], 1), oc_at(4, xsl[3], 1),
                  oc_at(4, xsl[5], 1), ld(bsl[3] + ob[4] + CS));
      put(3, pp, xsl[4], 1, nv0);
      if (w4) put(4, pp, xsl[4], 1, nv1);
    }
    __syncthreads();
    // group 3: R2(j + 2), staging of x(j + 6)
    if (!CHK || (j + 2 >= r2lo && j + 2 < r2hi)) {
      const bool bnd = edge(j + 2);
      const C zm = CHK ? oc_at(5, xsl[2], 0) : cm[3];
      const C sm = CHK ? oc_at(5, xsl[3], 0) : cs[3];
      const C zp = oc_at(5, xsl[4], 0);
      const C nv0 = upd(5, par, bnd, xsl[3], 0, sm, zm, zp, ld(bsl[2] + ob[5]));
      cm[3] = sm;
      cs[3] = zp;
      C nv1 = C(0);
      if (w6)
        nv1 = upd(6, par, bnd, xsl[3], 0, oc_at(6, xsl[3], 0), oc_at(6, xsl[2], 0),
                  oc_at(6, xsl[4], 0), ld(bsl[2] + ob[6]));
      put(5, par, xsl[3], 0, nv0);
      if (w6) put(6, par, xsl[3], 0, nv1);
    }
    {
      const int Sb = xsl[0];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        if (q == 0 || ((bits >> 30) & 1)) {
          int rs = mask(28 + q) ^ (par ? -1 : 0);
          st_(Sb + so[q], pick(rs, xr[q].y, xr[q].x));
          st_(Sb + so[q] + CS, pick(rs, xr[q].x, xr[q].y));
        }
      }
    }
    if (!ZERO) fetch(rx, kx0, j + 8, xo, xr);
    __syncthreads();
  };

  // prologue: x(k0-4 .. k0-2) staged; x(k0-1), x(k0), b(k0-3), b(k0-2) in
  // registers for the first two iterations (issued in the loop's order)
  V2 xA[2], xB[2], bA[2], bB[2];
  {
    V2 t0[2], t1[2], t2[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) xA[q] = xB[q] = t0[q] = t1[q] = t2[q] = make_float2(0.f, 0.f);
    if (!ZERO) {
      fetch(rx, kx0, k0 - 4, xo, t0);
      fetch(rx, kx0, k0 - 3, xo, t1);
      fetch(rx, kx0, k0 - 2, xo, t2);
    }
    fetch(rb, kb0, k0 - 3, bo, bA);
    if (!ZERO) fetch(rx, kx0, k0 - 1, xo, xA);
    fetch(rb, kb0, k0 - 2, bo, bB);
    if (!ZERO) fetch(rx, kx0, k0, xo, xB);
    stage_x_g(k0 - 4, t0);
    stage_x_g(k0 - 3, t1);
    stage_x_g(k0 - 2, t2);
  }
  __syncthreads();
  // j = k0 - 7 is odd.  Ring slots: plane p in x slot (p - (k0 - 4)) mod 7,
  // in b slot (p - (k0 - 3)) mod 5: at j, planes j-1 .. j+5 in x slots 3, 4,
  // 5, 6, 0, 1, 2 and planes j .. j+4 in b slots 1, 2, 3, 4, 0.
#pragma unroll
  for (int q = 0; q < 7; ++q) xsl[q] = ((3 + q) % TB_XR) * XSL;
#pragma unroll
  for (int q = 0; q < 5; ++q) bsl[q] = BOFF + ((1 + q) % TB_BR) * BSL;
  auto rotate = [&]() {
    int t = xsl[0];
#pragma unroll
    for (int q = 0; q < 6; ++q) xsl[q] = xsl[q + 1];
    xsl[6] = t;
    t = bsl[0];
#pragma unroll
    for (int q = 0; q < 4; ++q) bsl[q] = bsl[q + 1];
    bsl[4] = t;
  };
  // Steady state: iterations j in [max(k0, 1), min(k1 - 1, nz - 5)) have
  // every stage in range and no plane at 0 or nz - 1; they run unchecked, from
  // the first odd j there (the loop pairs odd and even iterations).
  const std::integral_constant<int, 1> odd;
  const std::integral_constant<int, 0> even;
  const std::true_type chk;
  const std::false_type nochk;
  const int js = max(k0, 1) | 1, je = min(k1 - 1, nz - 5);
  int j = k0 - 7;
  for (; j + 1 < k1 && j < js; j += 2) {
    iter(odd, chk, j, xA, bA);
    rotate();
    iter(even, chk, j + 1, xB, bB);
    rotate();
  }
  for (; j + 1 < je; j += 2) {
    iter(odd, nochk, j, xA, bA);
    rotate();
    iter(even, nochk, j + 1, xB, bB);
    rotate();
  }
  for (; j + 1 < k1; j += 2) {
    iter(odd, chk, j, xA, bA);
    rotate();
    iter(even, chk, j + 1, xB, bB);
    rotate();
  }
  if (j < k1) iter(odd, chk, j, xA, bA);
}

// ---------------------------------------------------------------------------
// Two red-black sweeps on one 3-D tile per block, for the coarse levels: the
// same result as k_rbgs_sweep2 bit for bit (the cell arithmetic of rb_cell,
// the x neighbours summed pair cell first), without its z-march.  On a grid
// of 1-3 M cells a marching pass is a chain of ~13 dependent plane iterations
// (7 of them filling the stage pipeline) per block; here a block loads its
// tile of x with a four-cell ring in every direction (TX+8 x TY+8 x TZ+8
// cells) into LDS at once, runs R1 (tile plus three cells), B1 (plus two), R2
// (plus one) and B2 (the tile) with one barrier between stages, and writes
// the tile: one memory round trip and four short stages.  Cells outside the
// grid stage as 0 and are never updated, so the stencil sums its six
// neighbours without branches; only the diagonal depends on where the cell is.
//
// In a stage each thread owns one column of the stage's region, a (pair,
// row), and walks its planes: everything about the column (LDS slot, side of
// the stage's colour on even and odd planes, in-grid test, 1/diag inside and
// on the z walls) is formed once; the plane enters only as an immediate LDS
// offset and a wave-uniform z test.  The right-hand sides of a thread's
// updates, one per plane of each stage, are loaded into registers with the
// staging.  LDS holds each staged row split by colour (TX/2+4 red pairs, then
// as many black ones): a cell's six neighbours are of the other colour, at
// pair p (the pair cell, and the rows and planes either side) and at pair p-1
// or p+1 (by the cell's side).  ZERO: zero initial guess (x is not read).
// scale (nullable): the right-hand side is b scale[0].  Storage S: FP32 or
// FP16 (widened as staged, rounded once on the store).

// Stage geometry of k_rbgs_tile2: a stage on the tile plus an R-cell ring
// takes pairs [PLO, PLO + NPP) of a staged row and rows [4 - R, 4 + TY + R)
// (NC columns), planes [4 - R, 4 + TZ + R) (NZ).
template <int TX, int TY, int TZ, int R> struct Tile2Stage {
  static constexpr int PLO = (4 - R) / 2, NPP = (TX + 3 + R) / 2 + 1 - PLO;
  static constexpr int NRY = TY + 2 * R, NC = NPP * NRY, NZ = TZ + 2 * R;
};

template <bool ZERO, class S, bool W1, int TX, int TY, int TZ, int NT>
__global__ void __launch_bounds__(NT)
k_rbgs_tile2(Dims d, const S *__restrict__ xin, S *__restrict__ xout, const S *__restrict__ b,
             float h2, float topc, float omega, const double *__restrict__ scale, int ncx,
             int ncy) {
  typedef float C;
  constexpr int PX = TX / 2 + 4;           // staged pairs per row (four-cell ring)
  constexpr int EY = TY + 8, EZ = TZ + 8;  // staged rows and planes
  constexpr int RS = 2 * PX;               // floats per staged row: red, then black
  constexpr int PS = EY * RS;              // floats per staged plane
  constexpr int es = (int)sizeof(S);
  static_assert(TX % 2 == 0 && TZ % 2 == 0, "x-pairs; tiles start on even planes");
  static_assert(Tile2Stage<TX, TY, TZ, 3>::NC <= NT, "one column per thread and stage");
  __shared__ C lds[EZ * PS];
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  const int lb = xcd_remap(blockIdx.x, gridDim.x);
  const int x0 = (lb % ncx) * TX, y0 = ((lb / ncx) % ncy) * TY, z0 = (lb / (ncx * ncy)) * TZ;
  const int tid = threadIdx.x;
  const C bs = scale ? (C)scale[0] : C(1);
  const int pst = nx * ny * es;  // bytes per plane
  const uint32_t bytes = (uint32_t)(d.cells() * es);
  const BufRsrc rx = buf_rsrc(xin, bytes), rb = buf_rsrc(b, bytes);

  // ---- per stage: this thread's column ----------------------------------------
  // a: LDS slot of the column's colour-X cell on plane 0; fo[h]: offset of its
  // x neighbour in the next pair over on planes of parity h (sides alternate
  // plane by plane); ok bit h: the cell lies in the ring and in the grid in x
  // and y; rd[h] / rw[h]: 1/diag on planes with both z neighbours / with one;
  // bo[h]: byte offset of the cell in its plane (b's loads)
  struct Col { int a, fo[2], bo[2]; unsigned ok; C rd[2], rw[2]; bool act; };
  auto column = [&](auto Xc, auto Rc) -> Col {
    constexpr int X = decltype(Xc)::value, R = decltype(Rc)::value;
    typedef Tile2Stage<TX, TY, TZ, R> T;
    Col c;
    c.act = tid < T::NC;
    const int w = c.act ? tid : 0;
    const int pp = T::PLO + w % T::NPP, py = 4 - R + w / T::NPP;
    const int gy = y0 - 4 + py, gxl = x0 - 4 + 2 * pp;
    c.a = py * RS + X * PX + pp;
    c.ok = 0;
    const C top = (gy < ny - 1) ? C(1) : (C)topc;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      // plane parity h: gz = z0 - 4 + pz with z0 even, so gz & 1 = pz & 1
      const int sd = ((gy + h) & 1) ^ X;  // side of the colour-X cell in its pair
      const int gx = gxl + sd;
      c.fo[h] = sd ? 1 : -1;
      bool in = c.act && gx >= 0 && gx < nx && gy >= 0 && gy < ny;
      if (R & 1) in = in && gx >= x0 - R && gx < x0 + TX + R;
      c.ok |= (in ? 1u : 0u) << h;
      const int nxy = (gx > 0) + (gx < nx - 1) + (gy > 0);
      c.rd[h] = fast_rcp(top + C(nxy + 2));
      c.rw[h] = fast_rcp(top + C(nxy + 1));
      c.bo[h] = (gy * nx + gx) * es;
    }
    return c;
  };
  const std::integral_constant<int, 0> red, r0;
  const std::integral_constant<int, 1> black, r1;
  const std::integral_constant<int, 2> r2;
  const std::integral_constant<int, 3> r3;
  const Col c1 = column(red, r3), c2 = column(black, r2), c3 = column(red, r1),
            c4 = column(black, r0);

  // ---- loads: the staged x pairs by columns (thread t stages the column of
  // pairs (t mod PX, t / PX) through all planes), then the right-hand side of
  // every update
  constexpr int NXC = PX * EY;
  static_assert(NXC <= NT, "one staged column per thread");
  const bool xact = tid < NXC;
  const int xpp = xact ? tid % PX : 0, xpy = xact ? tid / PX : 0;
  const int xgx = x0 - 4 + 2 * xpp, xgy = y0 - 4 + xpy;
  // in-plane byte offset of the staged pair, or out of every plane's range
  // (with the level under 2^30 bytes, no plane offset brings it back)
  const int xo = (xgx >= 0 && xgx < nx && xgy >= 0 && xgy < ny) ? (xgy * nx + xgx) * es
                                                                 : MAC3D_OFF_NONE;
  float2 xs[ZERO ? 1 : EZ];
  if (!ZERO) {
#pragma unroll
    for (int pz = 0; pz < EZ; ++pz) {
      const int gz = z0 - 4 + pz;
      // planes outside the grid lie outside the buffer's range (zeros)
      xs[pz] = buf_load_pair<S>(rx, xo + gz * pst);
    }
  }
  // b of a column's cell on plane pz (0 outside the grid in z: the plane's
  // offset then lies outside the buffer's range)
  auto ld_b = [&](const Col &c, int pz) -> C {
    const int gz = z0 - 4 + pz;
    const int o = c.bo[pz & 1] + gz * pst;
    if constexpr (sizeof(S) == 4) return buf_load<float>(rb, o) * bs;
    else
      return (C)__builtin_bit_cast(
                 half_t, (unsigned short)__builtin_amdgcn_raw_buffer_load_b16(rb, o, 0, 0)) *
             bs;
  };
  C b1[TZ + 6], b2[TZ + 4], b3[TZ + 2], b4[TZ];
#pragma unroll
  for (int z = 0; z < TZ + 6; ++z) b1[z] = ld_b(c1, 1 + z);
#pragma unroll
  for (int z = 0; z < TZ + 4; ++z) b2[z] = ld_b(c2, 2 + z);
#pragma unroll
  for (int z = 0; z < TZ + 2; ++z) b3[z] = ld_b(c3, 3 + z);
#pragma unroll
  for (int z = 0; z < TZ; ++z) b4[z] = ld_b(c4, 4 + z);
  // staging: pair (x, y) to its row's red and black slots; red on the right
  // where (gy + gz) is odd (the ring offsets are even)
  if (xact) {
    const int a0 = xpy * RS + xpp, rs0 = (y0 + xpy + z0) & 1;
#pragma unroll
    for (int pz = 0; pz < EZ; ++pz) {
      const int a = a0 + pz * PS;
      if (ZERO) {
        lds[a] = C(0);
        lds[a + PX] = C(0);
      } else {
        const bool rs = rs0 ^ (pz & 1);
        lds[a] = rs ? xs[pz].y : xs[pz].x;
        lds[a + PX] = rs ? xs[pz].x : xs[pz].y;
      }
    }
  }
  __syncthreads();

  // One colour stage over the column's planes: x neighbours (pair cell first,
  // as k_rbgs_sweep2 sums them), y-, y+, z-, z+, then the update.  OUT (B2):
  // the new black cell goes to xout with the pair's red cell from R2.
  auto stage = [&](auto Xc, auto Rc, auto OUTc, const Col &c, const auto &bv) {
    constexpr int X = decltype(Xc)::value, R = decltype(Rc)::value;
    constexpr bool OUT = decltype(OUTc)::value;
    typedef Tile2Stage<TX, TY, TZ, R> T;
    // whole waves past the stage's columns skip it (a scalar branch)
    if (__builtin_amdgcn_readfirstlane(tid & ~63) >= T::NC) return;
    constexpr int oc = X ? -PX : PX;  // the pair's other-colour cell
#pragma unroll
    for (int z = 0; z < T::NZ; ++z) {
      const int pz = 4 - R + z, h = pz & 1;
      const int gz = z0 - 4 + pz;  // uniform over the block
      const int a = c.a + pz * PS, o = a + oc;
      const C sm = lds[o], fx = lds[o + c.fo[h]];
      const C ym = lds[o - RS], yp = lds[o + RS], zm = lds[o - PS], zp = lds[o + PS];
      const C rdg = (gz > 0 && gz < nz - 1) ? c.rd[h] : c.rw[h];
      C nv;
      if (W1) nv = rb_cell_r1<C>(sm, fx, ym, yp, zm, zp, bv[z], (C)h2, rdg);
      else nv = rb_cell_r<C>(sm, fx, ym, yp, zm, zp, bv[z], lds[a], (C)h2, rdg, (C)omega);
      if (gz < 0 || gz >= nz || !((c.ok >> h) & 1u)) continue;
      if (OUT) {
        const int sd = c.fo[h] > 0;
        float2 ov;
        ov.x = sd ? sm : nv;
        ov.y = sd ? nv : sm;
        store_pair(xout + (size_t)gz * nx * ny + ((c.bo[h] / es) - sd), ov);
      } else {
        lds[a] = nv;
      }
    }
  };
  stage(red, r3, std::false_type{}, c1, b1);
  __syncthreads();
  stage(black, r2, std::false_type{}, c2, b2);
  __syncthreads();
  stage(red, r1, std::false_type{}, c3, b3);
  __syncthreads();
  stage(black, r0, std::true_type{}, c4, b4);
}

}  // namespace mac3d
