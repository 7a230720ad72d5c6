// mac3d.hip — context, step orchestration and C ABI of the MAC step.
//
// The context owns all device buffers and one HIP stream (default priority).
// Host code only sequences kernel launches on that stream; every field
// operation runs in a hand-written kernel of step_kernels.hpp,
// mg_kernels.hpp or mixed_kernels.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/mac3d.h"
#include "common.hpp"
#include "mg_kernels.hpp"
#include "mixed_kernels.hpp"
#include "step_kernels.hpp"
#include "adjoint_kernels.hpp"
#include "wave_kernels.hpp"
#include "parcel_kernels.hpp"
#include "conc_kernels.hpp"
#include "profile.hpp"
#include "release.hpp"

using namespace mac3d;

namespace {

enum { TF16 = 0, TF32 = 1, TF64 = 2 };
const int kMaxLev = 8;
// refinement iterations a mixed step may run before its first convergence test
const int kMaxPreIters = 15;

size_t type_size(int t) { return t == TF16 ? 2 : (t == TF32 ? 4 : 8); }

struct Buf {
  void *ptr = nullptr;
  int type = TF32;
  size_t n = 0;
};

template <class F> void with_type(int t, F &&f) {
  switch (t) {
    case TF16: f((half_t *)nullptr); break;
    case TF32: f((float *)nullptr); break;
    default: f((double *)nullptr); break;
  }
}

// step fields are FP32 or FP64 only
template <class F> void with_step_type(int t, F &&f) {
  if (t == TF64) f((double *)nullptr); else f((float *)nullptr);
}

struct Ctx {
  mac3d_config cfg;
  int L = 0;
  int prec = 0;
  int t0 = TF32;  // type of the level-0 step fields
  Dims d[kMaxLev + 1];
  size_t n[kMaxLev + 1];
  hipStream_t stream = nullptr;
  int ncu = 256;                     // the device's compute units
  hipEvent_t norms_ready = nullptr;  // mixed mode: the residual norms reached the host
  // cvel holds the centre velocity of the current u, v, w (set by a step's
  // last face-to-centre pass, cleared by anything else that may change them)
  bool cvel_fresh = false;
  // dmm holds the limiter bounds of the dye and centre velocity of the last
  // dye predict, and nothing has changed either since (resolve may use them)
  bool dmm_valid = false;
  std::string err;
  // level-0 step fields
  Buf u, v, w, u2, v2, w2, p, dye, dye2, dpred, cvel, rhs;  // cvel: 3 per cell
  Buf dmm;  // limiter bounds of the dye predict: 2 per cell
  // multigrid: x, b, r per level (level 0 x/b alias p/rhs except in mixed);
  // sweeps run out of place between mgx and its alternate mgxa
  Buf mgx[kMaxLev + 1], mgxa[kMaxLev + 1], mgb[kMaxLev + 1], mgr[kMaxLev + 1];
  // mixed-mode FP32 L0 correction and its right-hand side
  Buf e0, r0;
  // mixed mode: FP64 pressure ring; role 0 = current, 1 = previous step,
  // 2 = the one before, 3 = spare; hist = converged pressures held (<= 2)
  Buf pr[4];
  int role[4] = {0, 1, 2, 3};
  int hist = 0;
  // refinement iterations of the last few mixed steps (iter_n of them held):
  // a step runs the fewest of them before its first convergence test
  int iter_log[4] = {0, 0, 0, 0};
  int iter_n = 0;
  double omega = 1.0;
  int extrap = 0;
  // right-hand-side scale of level 0: [0] = multiplier, [1] = its inverse
  double *scale = nullptr;
  // every device allocation, freed at destroy (the Bufs above are views)
  std::vector<void *> allocs;
  // home pointers of the ping-pong buffers (parity 0)
  void *u_home = nullptr, *dye_home = nullptr, *mgx_home[kMaxLev + 1] = {};
  // captured launch sequences, keyed by kind and buffer parities
  struct Graph {
    hipGraphExec_t exec = nullptr;
    double dt = 0.0;
    uint64_t after = 0;  // parity bits after the sequence
  };
  std::map<uint64_t, Graph> graphs;
  bool use_graphs = true;
  // adjoint work space (allocated on first use)
  struct Adj {
    bool ready = false;
    void *c1 = nullptr, *vela[3] = {}, *lvel[3] = {}, *lvelc[3] = {}, *lvela[3] = {}, *lnew[3] = {};
    void *lc[3] = {}, *ldye = nullptr, *ldyen = nullptr, *lpred = nullptr, *target = nullptr;
    Buf lx, lxa, lb;  // adjoint pressure solve on level 0
    double *part = nullptr, *acc = nullptr;  // partial sums; acc[0] = d loss / d rho_b
    size_t part_n = 0;
    // states (u, v, w, p, dye) of the checkpoints and of one segment
    std::vector<void *> ckpt, seg;
    int ckpt_cap = 0, seg_cap = 0;
  } adj;
  bool adj_solving = false;
  // coarse solve
  double *cq = nullptr;  // qx | qy | qz | lx | ly | lz
  size_t cq_off[6] = {0, 0, 0, 0, 0, 0};
  double *cwork = nullptr;
  // reductions
  double *part = nullptr;
  size_t part_n = 0;
  double *red = nullptr;     // device fold output (3 doubles)
  double *red_host = nullptr; // pinned
  // free-surface coupling (off until mac3d_wave_enable): wave state (4 per
  // texel), the complex spectrum (2 per texel), the force pass's film, the lid
  // pressure (nx * nz), the drop slots and the FFT twiddles on the device
  bool wave_on = false;
  bool wave_scene = false;  // also the advection's scene terms
  // the reference's own physics (lattice Casson, tanh buoyancy, clamp-to-edge
  // neighbours, the projection clamp, the dye predictor's branches)
  bool ref_physics = false;
  double casson_scale = 1.0, buoyancy = 0.48;
  // the reference's 4-channel concentration (planes: intact, lysed,
  // oxygenation, floor bed; FP32) and its twin on their own grid dcg, conc_r
  // times the MAC grid per axis; the lysed channel's predictor there; with
  // conc_r 2 the channels as the MAC cells read them (cmac: .r and .g at the
  // cell centres, then at the top y-faces, n[0] each); the parcels' pending
  // emission (device memory, read by the emission kernel of a parcel stage)
  bool conc_on = false;
  int conc_r = 0;
  Dims dcg{};
  Buf conc, conc2, cpred, cmac;
  void *conc_home = nullptr;
  EmitCtl *emit = nullptr;
  WaveParams wp{};
  Buf wave, wz, wfilm, ptop;
  Buf curl;  // scene: curl of the post-advection centre velocity, 4 per cell
  SceneClock clk_host{0.f, 0.032f, 1.02f, 0.088f};  // host copy of the device clock
  // Lagrangian parcels and their two splat counters (intact, lysed; one per
  // cell of the splat grid, room for splat_cap cells each)
  Parcel *parcels = nullptr;
  int np = 0, np_cap = 0;
  unsigned *splat = nullptr;
  size_t splat_cap = 0;
  WaveDrop *drops = nullptr;  // WV_MAX_DROPS slots, then the SceneClock
  float2 *wtw = nullptr;
  // staging for upload/download (stage_n doubles)
  double *stage = nullptr;
  size_t stage_n = 0;
  mac3d::BlobDev *blobs = nullptr;
  int blobs_cap = 0;
  Profiler prof;
  bool dead = false;
};

#define HIPCHK(call)                                                     \
  do {                                                                   \
    hipError_t e_ = (call);                                              \
    if (e_ != hipSuccess) {                                              \
      ctx->err = std::string(#call) + ": " + hipGetErrorString(e_);      \
      return MAC3D_EHIP;                                                 \
    }                                                                    \
  } while (0)

#define LAUNCHCHK()                                                      \
  do {                                                                   \
    hipError_t e_ = hipGetLastError();                                   \
    if (e_ != hipSuccess) {                                              \
      ctx->err = std::string("kernel launch: ") + hipGetErrorString(e_); \
      return MAC3D_EHIP;                                                 \
    }                                                                    \
  } while (0)

int fail(Ctx *ctx, int code, const std::string &msg) {
  ctx->err = msg;
  return code;
}

dim3 cell_grid(const Dims &d) { return dim3((d.nx + 63) / 64, (d.ny + 3) / 4, d.nz); }
const dim3 kCellBlock(64, 4, 1);
// the semi-Lagrangian gather kernels run 32 x 4 blocks: their 8-texel
// samples reach one row and plane either side, and the narrower block keeps
// more of them in one CU's L1 (measured fastest of 64x4, 32x8, 16x16, 32x4)
dim3 gather_grid(const Dims &d) { return dim3((d.nx + 31) / 32, (d.ny + 3) / 4, d.nz); }
const dim3 kGatherBlock(32, 4, 1);

double top_coeff(int level) {
  double s = (double)(1 << level);
  return 2.0 * s / (s + 1.0);
}
double level_h2(int level) {
  double s = (double)(1 << level);
  return s * s;
}

int dev_alloc(Ctx *ctx, void **p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(ctx, MAC3D_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  ctx->allocs.push_back(*p);
  return MAC3D_OK;
}

// frees a dev_alloc'd block before the context ends (the caller drops any
// graph that holds it first)
void dev_free(Ctx *ctx, void *p) {
  if (!p) return;
  auto it = std::find(ctx->allocs.begin(), ctx->allocs.end(), p);
  if (it != ctx->allocs.end()) ctx->allocs.erase(it);
  (void)hipFree(p);
}

int alloc(Ctx *ctx, Buf &b, int type, size_t n) {
  b.type = type;
  b.n = n;
  int rc = dev_alloc(ctx, &b.ptr, type_size(type) * n);
  if (rc) return rc;
  hipError_t e = hipMemsetAsync(b.ptr, 0, type_size(type) * n, ctx->stream);
  if (e != hipSuccess) return fail(ctx, MAC3D_EHIP, hipGetErrorString(e));
  return MAC3D_OK;
}

// Symmetric tridiagonal 1-D operator of one axis and its eigen-pairs by
// cyclic Jacobi rotations (n <= a few hundred, host FP64, once per context).
void axis_eigen(int n, bool lid, double topc, std::vector<double> &Q, std::vector<double> &lam) {
  std::vector<double> A(n * n, 0.0);
  for (int i = 0; i < n; ++i) {
    double dg = (i > 0 ? 1.0 : 0.0);
    if (i < n - 1) dg += 1.0; else if (lid) dg += topc;
    A[i * n + i] = dg;
    if (i > 0) A[i * n + i - 1] = A[(i - 1) * n + i] = -1.0;
  }
  Q.assign(n * n, 0.0);
  for (int i = 0; i < n; ++i) Q[i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) off += A[p * n + q] * A[p * n + q];
    if (off < 1e-30) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        double apq = A[p * n + q];
        if (std::fabs(apq) < 1e-300) continue;
        double theta = (A[q * n + q] - A[p * n + p]) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int r = 0; r < n; ++r) {  // A <- A J
          double arp = A[r * n + p], arq = A[r * n + q];
          A[r * n + p] = c * arp - s * arq;
          A[r * n + q] = s * arp + c * arq;
        }
        for (int r = 0; r < n; ++r) {  // A <- J^T A
          double apr = A[p * n + r], aqr = A[q * n + r];
          A[p * n + r] = c * apr - s * aqr;
          A[q * n + r] = s * apr + c * aqr;
        }
        for (int r = 0; r < n; ++r) {  // Q <- Q J
          double qrp = Q[r * n + p], qrq = Q[r * n + q];
          Q[r * n + p] = c * qrp - s * qrq;
          Q[r * n + q] = s * qrp + c * qrq;
        }
      }
  }
  lam.resize(n);
  for (int i = 0; i < n; ++i) lam[i] = A[i * n + i];
}

// ---------------------------------------------------------------------------
// kernel launchers
int launch_face_to_centre(Ctx *ctx) {
  const Dims &d = ctx->d[0];
  ctx->dmm_valid = false;  // the limiter bounds were traced on the old cvel
  ProfScope ps(ctx->prof, ctx->stream, KC_F2C);
  with_step_type(ctx->t0, [&](auto tag) {
    using T = std::remove_pointer_t<decltype(tag)>;
    auto kern = ctx->wave_on && ctx->wave_scene ? k_face_to_centre<T, true>
                                                : k_face_to_centre<T, false>;
    hipLaunchKernelGGL(kern, cell_grid(d), kCellBlock, 0, ctx->stream, d,
                       (const T *)ctx->u.ptr, (const T *)ctx->v.ptr, (const T *)ctx->w.ptr,
                       (typename cvec_of<T>::type *)ctx->cvel.ptr, ctx->wp,
                       (const WaveDrop *)ctx->drops);
  });
  LAUNCHCHK();
  return MAC3D_OK;
}

StepParams step_params(const Ctx *ctx, double dt) {
  const mac3d_config &c = ctx->cfg;
  StepParams P;
  P.dt = dt; P.dx = c.dx; P.rho_w = c.rho_w; P.mu_w = c.mu_w; P.rho_b = c.rho_b;
  P.tau_y = c.tau_y; P.mu_inf = c.mu_inf; P.g = c.g; P.gamma_reg = c.gamma_reg;
  P.shear_eps2 = c.shear_eps2; P.alpha_max = c.alpha_max; P.vel_damp = c.vel_damp;
  P.maccormack = c.maccormack;
  P.casson_scale = ctx->casson_scale;
  P.buoyancy = ctx->buoyancy;
  P.surface_y = ctx->wp.surface_y;
  P.depth = ctx->wp.depth;
  return P;
}

// the coupling is on with the reference's own physics
bool ref_physics(const Ctx *ctx) { return ctx->wave_on && ctx->ref_physics; }

// The splat grid (the concentration grid while it is on, else the MAC grid),
// which also holds what the parcels read: the intact-RBC concentration (the
// dye, or the 4-channel concentration's .r) and the lysed one (.g; null
// without it).
const Dims &splat_dims(const Ctx *ctx) { return ctx->conc_on ? ctx->dcg : ctx->d[0]; }
size_t splat_cells(const Ctx *ctx) { return splat_dims(ctx).cells(); }
const void *intact_conc(const Ctx *ctx) { return ctx->conc_on ? ctx->conc.ptr : ctx->dye.ptr; }
const float *lysed_conc(const Ctx *ctx) {
  return ctx->conc_on ? (const float *)ctx->conc.ptr + ctx->conc.n / 4 : nullptr;
}
// The same two as the MAC cells' Casson pass, buoyancy and unfurling gate
// read them, and at the cells' top y-faces (null: the mean of the cell and
// the cell above): on the MAC grid the fields themselves, from a finer
// concentration grid the samples k_conc_at_mac forms.
const void *intact_mac(const Ctx *ctx) {
  return ctx->conc_on && ctx->conc_r != 1 ? ctx->cmac.ptr : intact_conc(ctx);
}
const float *lysed_mac(const Ctx *ctx) {
  return ctx->conc_on && ctx->conc_r != 1 ? (const float *)ctx->cmac.ptr + ctx->n[0]
                                          : lysed_conc(ctx);
}
const float *face_mac(const Ctx *ctx, int which) {
  return ctx->conc_on && ctx->conc_r != 1 ? (const float *)ctx->cmac.ptr + (2 + which) * ctx->n[0]
                                          : nullptr;
}
// the two splat counters hold at least the splat grid's cells each
// (allocated zeroed; the caller drops graphs that hold the old block)
int ensure_splat(Ctx *ctx) {
  const size_t need = splat_cells(ctx);
  if (ctx->splat && ctx->splat_cap >= need) return MAC3D_OK;
  dev_free(ctx, ctx->splat);
  ctx->splat = nullptr;
  ctx->splat_cap = 0;
  int rc;
  if ((rc = dev_alloc(ctx, (void **)&ctx->splat, 2 * need * sizeof(unsigned)))) return rc;
  HIPCHK(hipMemsetAsync(ctx->splat, 0, 2 * need * sizeof(unsigned), ctx->stream));
  ctx->splat_cap = need;
  return MAC3D_OK;
}
ConcScene conc_scene(const Ctx *ctx) {
  ConcScene S;
  S.half_xz = ctx->wp.half_xz;
  S.y_max = ctx->wp.surface_y;
  S.y_min = ctx->wp.surface_y - ctx->wp.depth;
  S.ext = ctx->wp.ext;
  S.wave_n = ctx->wp.n;
  return S;
}

void swap_vel(Ctx *ctx) {
  std::swap(ctx->u, ctx->u2);
  std::swap(ctx->v, ctx->v2);
  std::swap(ctx->w, ctx->w2);
}

// The gather kernels read their fields through buffer resources with 32-bit
// byte offsets (formed by 24-bit multiplies) when the largest field (the
// packed centre velocity) is under 2^31 bytes and a plane of it under 2^24,
// else through plain pointers.
template <class T> bool gather_buf(const Dims &d) {
  return 3 * d.cells() * sizeof(T) < ((size_t)1 << 31) &&
         3 * (size_t)d.nx * d.ny * sizeof(T) < ((size_t)1 << 24);
}

// the wave coupling's parameters with the step's dt
WaveParams wave_params_dt(const Ctx *ctx, double dt) {
  WaveParams p = ctx->wp;
  p.dt = (float)std::min(std::max(dt, 0.0), 0.025);  // viewportAndDt.z clamp (wave.ts:83)
  return p;
}

// the drop slots' allocation: WV_MAX_DROPS WaveDrop, the SceneClock, then
// WV_MAX_DROPS DropPhase
const size_t kDropBytes =
    WV_MAX_DROPS * sizeof(WaveDrop) + sizeof(SceneClock) + WV_MAX_DROPS * sizeof(DropPhase);
const SceneClock *scene_clock(const Ctx *ctx) {
  return (const SceneClock *)(ctx->drops + WV_MAX_DROPS);
}
const DropPhase *drop_phases(const Ctx *ctx) {
  return (const DropPhase *)(scene_clock(ctx) + 1);
}
// One parcel step (BloodDropEngine.ts:2073-2093): both splat counters
// cleared, then every parcel advanced and splatted.
int launch_parcels(Ctx *ctx, double dt) {
  const Dims &d = ctx->d[0], &ds = splat_dims(ctx);
  if (ctx->t0 != TF32) return fail(ctx, MAC3D_EINVAL, "parcels need the FP32 or mixed step");
  const size_t ns = ds.cells(), n2 = 2 * ns;
  if (!ctx->splat || ctx->splat_cap < ns) return fail(ctx, MAC3D_ESTATE, "no splat counters");
  ProfScope ps(ctx->prof, ctx->stream, KC_PARCELS);
  hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)((n2 / 4 + 3 + 255) / 256)), dim3(256), 0,
                     ctx->stream, ctx->splat, n2);  // (hipMalloc: 256-byte aligned)
  LAUNCHCHK();
  if (ctx->np == 0) return MAC3D_OK;
  if (ctx->emit) {  // the pending emission, then mark it consumed
    hipLaunchKernelGGL(k_parcel_emit, dim3(kMaxEmit / 256), dim3(256), 0, ctx->stream, ds,
                       conc_scene(ctx), (const EmitCtl *)ctx->emit, (const WaveDrop *)ctx->drops,
                       scene_clock(ctx), (const float4 *)ctx->wave.ptr, ctx->np, ctx->parcels,
                       ctx->splat + ns);
    hipLaunchKernelGGL(k_emit_done, dim3(1), dim3(64), 0, ctx->stream, ctx->emit);
    LAUNCHCHK();
  }
  ParcelScene S;
  S.dt = (float)std::min(std::max(dt, 0.0), 0.025);
  S.half_xz = ctx->wp.half_xz;
  S.y_max = ctx->wp.surface_y;
  S.y_min = ctx->wp.surface_y - ctx->wp.depth;
  S.sim_time = ctx->clk_host.sim_time;
  S.k_lysis = ctx->clk_host.hemolysis;
  hipLaunchKernelGGL(k_parcel_advect_splat, dim3((unsigned)((ctx->np + 255) / 256)), dim3(256), 0,
                     ctx->stream, d, ds, S, ctx->np, ctx->parcels, (const float3 *)ctx->cvel.ptr,
                     (const float *)intact_conc(ctx), lysed_conc(ctx), ctx->splat,
                     ctx->splat + ns);
  LAUNCHCHK();
  return MAC3D_OK;
}


int launch_advect(Ctx *ctx, double dt) {
  const Dims &d = ctx->d[0];
  {
    ProfScope ps(ctx->prof, ctx->stream, KC_ADVECT);
    with_step_type(ctx->t0, [&](auto tag) {
      using T = std::remove_pointer_t<decltype(tag)>;
      const bool buf = gather_buf<T>(d), scene = ctx->wave_on && ctx->wave_scene;
      if constexpr (std::is_same<T, float>::value) {
        // FP32: samples from the block's staged tile (buffer gathers where a
        // sample leaves it).  Not with the scene terms: their registers cost
        // the staged kernel more than its gathers save (315 against 273 us at
        // 192x256x192), and scene flows trace further.
        if (buf && !scene) {
          hipLaunchKernelGGL(k_advect_mac_lds<false>,
                             dim3((d.nx + AL_TX - 1) / AL_TX, (d.ny + AL_TY - 1) / AL_TY,
                                  (d.nz + AL_TZ - 1) / AL_TZ),
                             dim3(AL_TX * AL_TY), 0, ctx->stream, d, (float)(dt / ctx->cfg.dx),
                             (float)ctx->cfg.vel_damp, (const float *)ctx->u.ptr,
                             (const float *)ctx->v.ptr, (const float *)ctx->w.ptr,
                             (const float3 *)ctx->cvel.ptr, (float *)ctx->u2.ptr,
                             (float *)ctx->v2.ptr, (float *)ctx->w2.ptr, wave_params_dt(ctx, dt),
                             (const WaveDrop *)ctx->drops, drop_phases(ctx),
                             (const float4 *)ctx->wave.ptr);
          return;
        }
      }
      auto kern = scene ? (buf ? k_advect_mac<T, true, true> : k_advect_mac<T, false, true>)
                        : (buf ? k_advect_mac<T, true, false> : k_advect_mac<T, false, false>);
      hipLaunchKernelGGL(kern, gather_grid(d), kGatherBlock, 0, ctx->stream, d,
                         (T)(dt / ctx->cfg.dx), (T)ctx->cfg.vel_damp, (const T *)ctx->u.ptr, (const T *)ctx->v.ptr,
                         (const T *)ctx->w.ptr, (const typename cvec_of<T>::type *)ctx->cvel.ptr,
                         (T *)ctx->u2.ptr, (T *)ctx->v2.ptr, (T *)ctx->w2.ptr,
                         wave_params_dt(ctx, dt), (const WaveDrop *)ctx->drops, drop_phases(ctx),
                         (const float4 *)ctx->wave.ptr);
    });
  }
  LAUNCHCHK();
  swap_vel(ctx);
  return MAC3D_OK;
}

// The scene's Casson terms (BloodDropEngine.ts:1922-1934): the centre velocity
// of the advected faces, its curl, then, after the viscous update, the
// confinement and unfurling forces in place.
int launch_curl(Ctx *ctx) {
  const Dims &d = ctx->d[0];
  int rc;
  if ((rc = launch_face_to_centre(ctx))) return rc;
  {
    ProfScope ps(ctx->prof, ctx->stream, KC_CASSON);
    with_step_type(ctx->t0, [&](auto tag) {
      using T = std::remove_pointer_t<decltype(tag)>;
      hipLaunchKernelGGL(k_curl<T>, cell_grid(d), kCellBlock, 0, ctx->stream, d,
                         (const typename cvec_of<T>::type *)ctx->cvel.ptr,
                         (typename q4_of<T>::type *)ctx->curl.ptr);
    });
  }
  LAUNCHCHK();
  return MAC3D_OK;
}

int launch_casson_scene(Ctx *ctx, double dt) {
  const Dims &d = ctx->d[0];
  {
    ProfScope ps(ctx->prof, ctx->stream, KC_CASSON);
    with_step_type(ctx->t0, [&](auto tag) {
      using T = std::remove_pointer_t<decltype(tag)>;
      hipLaunchKernelGGL(k_casson_scene<T>, cell_grid(d), kCellBlock, 0, ctx->stream, d,
                         wave_params_dt(ctx, dt), scene_clock(ctx),
                         (const typename q4_of<T>::type *)ctx->curl.ptr,
                         (const T *)intact_mac(ctx), (const T *)lysed_mac(ctx),
                         (T *)ctx->u.ptr, (T *)ctx->v.ptr, (T *)ctx->w.ptr);
    });
  }
  LAUNCHCHK();
  return MAC3D_OK;
}

int launch_casson(Ctx *ctx, double dt) {
  const Dims &d = ctx->d[0];
  StepParams P = step_params(ctx, dt);
  const bool scene = ctx->wave_on && ctx->wave_scene;
  int rc;
  if (scene && (rc = launch_curl(ctx))) return rc;
  if (ctx->conc_on && ctx->conc_r != 1) {  // the finer grid's samples the pass reads
    ProfScope ps(ctx->prof, ctx->stream, KC_CASSON);
    hipLaunchKernelGGL(k_conc_at_mac, cell_grid(d), kCellBlock, 0, ctx->stream, d, ctx->dcg,
                       (const float *)ctx->conc.ptr, (float *)ctx->cmac.ptr);
    LAUNCHCHK();
  }
  {
    ProfScope ps(ctx->prof, ctx->stream, KC_CASSON);
    with_step_type(ctx->t0, [&](auto tag) {
      using T = std::remove_pointer_t<decltype(tag)>;
      // z-marching kernel: 64 x 8 tiles, about three 512-thread blocks per CU
      int ncx = (d.nx + CS_TX - 1) / CS_TX, ncy = (d.ny + CS_TY - 1) / CS_TY;
      int chunks = std::max(1, 768 / (ncx * ncy));
      int cz = (d.nz + chunks - 1) / chunks;
      dim3 grid(ncx * ncy * ((d.nz + cz - 1) / cz));
      auto kern = ref_physics(ctx) ? k_casson_zm<T, true> : k_casson_zm<T, false>;
      hipLaunchKernelGGL(kern, grid, dim3(CS_TX, CS_TY), 0, ctx->stream, d, P,
                         (const T *)ctx->u.ptr, (const T *)ctx->v.ptr, (const T *)ctx->w.ptr,
                         (const T *)intact_mac(ctx), (const T *)lysed_mac(ctx),
                         (const T *)face_mac(ctx, 0), (const T *)face_mac(ctx, 1),
                         (T *)ctx->u2.ptr, (T *)ctx->v2.ptr, (T *)ctx->w2.ptr, cz, ncx, ncy);
    });
  }
  LAUNCHCHK();
  swap_vel(ctx);
  return scene ? launch_casson_scene(ctx, dt) : MAC3D_OK;
}

// the lid's Dirichlet pressure for the divergence and projection: the wave
// coupling's, or none (the flat lid, p_top = 0)
const float *lid_pressure(const Ctx *ctx) {
  return ctx->wave_on ? (const float *)ctx->ptop.ptr : nullptr;
}

// One wave step (BloodDropEngine.ts:1886-1914): the force pass with the
// forward FFT of each row, the forward column FFT with the spectral
// propagation and the inverse column FFT, the inverse FFT of each row, and
// the finalize into the wave state (the reference's seven passes in four).
int launch_wave_step(Ctx *ctx, double dt) {
  const WaveParams P = wave_params_dt(ctx, dt);
  const int n = P.n;
  int logn = 0;
  while ((1 << logn) < n) ++logn;
  const dim3 g2((n + 15) / 16, (n + 15) / 16), b2(16, 16);
  ProfScope ps(ctx->prof, ctx->stream, KC_WAVE);
  float2 *z = (float2 *)ctx->wz.ptr;
  const float2 *tw = (const float2 *)ctx->wtw;
  const dim3 bf(std::max(64, std::min(512, n / 2)));
  const size_t sh = (size_t)n * sizeof(float2), shtw = (size_t)(n / 2) * sizeof(float2);
  with_step_type(ctx->t0, [&](auto tag) {
    using T = std::remove_pointer_t<decltype(tag)>;
    // (one thread per texel for the force pass; the FFT stages use half)
    hipLaunchKernelGGL(k_wave_force_rows<T>, dim3(n), dim3(std::max(64, std::min(512, n))),
                       sh + shtw, ctx->stream, P,
                       (const WaveDrop *)ctx->drops, (const float4 *)ctx->wave.ptr, ctx->d[0],
                       (const T *)ctx->v.ptr, z, (float *)ctx->wfilm.ptr, logn, tw);
  });
  hipLaunchKernelGGL(k_wave_colpair, dim3(n / 2 + 1), bf, 2 * sh + shtw, ctx->stream, P, z, logn,
                     tw);
  hipLaunchKernelGGL(k_wave_fft, dim3(n), bf, sh + shtw, ctx->stream, z, n, logn, 1, 1, tw);
  hipLaunchKernelGGL(k_wave_finalize, g2, b2, 0, ctx->stream, P, (const float2 *)z,
                     (const float *)ctx->wfilm.ptr, (float4 *)ctx->wave.ptr);
  LAUNCHCHK();
  return MAC3D_OK;
}

// evalTopDirichletPressure over the lid plane into ptop
int launch_top_pressure(Ctx *ctx, double dt) {
  const WaveParams P = wave_params_dt(ctx, dt);
  const Dims &d = ctx->d[0];
  ProfScope ps(ctx->prof, ctx->stream, KC_WAVE);
  hipLaunchKernelGGL(k_top_pressure, dim3((d.nx + 15) / 16, (d.nz + 15) / 16), dim3(16, 16), 0,
                     ctx->stream, P, (const WaveDrop *)ctx->drops, (const float4 *)ctx->wave.ptr,
                     d.nx, d.nz, (float *)ctx->ptop.ptr);
  LAUNCHCHK();
  return MAC3D_OK;
}

int launch_divergence(Ctx *ctx) {
  const Dims &d = ctx->d[0];
  ProfScope ps(ctx->prof, ctx->stream, KC_DIV);
  with_step_type(ctx->t0, [&](auto tag) {
    using T = std::remove_pointer_t<decltype(tag)>;
    with_step_type(ctx->rhs.type, [&](auto rtag) {
      using R = std::remove_pointer_t<decltype(rtag)>;
      hipLaunchKernelGGL((k_divergence<T, R>), cell_grid(d), kCellBlock, 0, ctx->stream, d,
                         (const T *)ctx->u.ptr, (const T *)ctx->v.ptr, (const T *)ctx->w.ptr,
                         (R *)ctx->rhs.ptr, lid_pressure(ctx));
    });
  });
  LAUNCHCHK();
  return MAC3D_OK;
}

int launch_project(Ctx *ctx) {
  const Dims &d = ctx->d[0];
  ProfScope ps(ctx->prof, ctx->stream, KC_PROJECT);
  with_step_type(ctx->t0, [&](auto tag) {
    using T = std::remove_pointer_t<decltype(tag)>;
    with_step_type(ctx->p.type, [&](auto ptag) {
      using Pt = std::remove_pointer_t<decltype(ptag)>;
      auto kern = ref_physics(ctx) ? k_project<T, Pt, true> : k_project<T, Pt, false>;
      hipLaunchKernelGGL(kern, cell_grid(d), kCellBlock, 0, ctx->stream, d,
                         (T *)ctx->u.ptr, (T *)ctx->v.ptr, (T *)ctx->w.ptr,
                         (const Pt *)ctx->p.ptr, lid_pressure(ctx));
    });
  });
  LAUNCHCHK();
  return MAC3D_OK;
}

int launch_dye_predict(Ctx *ctx, double dt) {
  const Dims &d = ctx->d[0];
  ProfScope ps(ctx->prof, ctx->stream, KC_DYE);
  with_step_type(ctx->t0, [&](auto tag) {
    using T = std::remove_pointer_t<decltype(tag)>;
    const bool buf = gather_buf<T>(d), ref = ref_physics(ctx);
    if constexpr (std::is_same<T, float>::value) {
      // FP32, plain step: samples from the block's staged tile
      if (buf && !ref) {
        hipLaunchKernelGGL(k_dye_predict_lds,
                           dim3((d.nx + AL_TX - 1) / AL_TX, (d.ny + AL_TY - 1) / AL_TY,
                                (d.nz + AL_TZ - 1) / AL_TZ),
                           dim3(AL_TX * AL_TY), 0, ctx->stream, d, (float)(dt / ctx->cfg.dx),
                           (const float3 *)ctx->cvel.ptr, (const float *)ctx->dye.ptr,
                           (float *)ctx->dpred.ptr, (float2 *)ctx->dmm.ptr);
        return;
      }
    }
    auto kern = ref ? (buf ? k_dye_predict<T, true, true> : k_dye_predict<T, false, true>)
                    : (buf ? k_dye_predict<T, true, false> : k_dye_predict<T, false, false>);
    // (the reference's branches sample elsewhere than the resolve's trace:
    // no limiter bounds from the predict there)
    hipLaunchKernelGGL(kern, gather_grid(d), kGatherBlock, 0, ctx->stream, d,
                       (T)(dt / ctx->cfg.dx),
                       (const typename cvec_of<T>::type *)ctx->cvel.ptr,
                       (const T *)ctx->dye.ptr, (T *)ctx->dpred.ptr,
                       ref ? nullptr : (typename vec2_of<T>::type *)ctx->dmm.ptr);
  });
  LAUNCHCHK();
  ctx->dmm_valid = !ref_physics(ctx);
  return MAC3D_OK;
}

int launch_dye_resolve(Ctx *ctx, double dt) {
  const Dims &d = ctx->d[0];
  {
    ProfScope ps(ctx->prof, ctx->stream, KC_DYE);
    with_step_type(ctx->t0, [&](auto tag) {
      using T = std::remove_pointer_t<decltype(tag)>;
      const T dth = (T)(dt / ctx->cfg.dx), s = (T)ctx->cfg.maccormack;
      const auto *cv = (const typename cvec_of<T>::type *)ctx->cvel.ptr;
      const bool buf = gather_buf<T>(d);
      auto kmm = buf ? k_dye_resolve_mm<T, true> : k_dye_resolve_mm<T, false>;
      auto kfull = ref_physics(ctx)
                       ? (buf ? k_dye_resolve<T, true, true> : k_dye_resolve<T, false, true>)
                       : (buf ? k_dye_resolve<T, true, false> : k_dye_resolve<T, false, false>);
      if (ctx->dmm_valid)
        hipLaunchKernelGGL(kmm, gather_grid(d), kGatherBlock, 0, ctx->stream, d, dth, s, cv,
                           (const T *)ctx->dye.ptr, (const T *)ctx->dpred.ptr,
                           (const typename vec2_of<T>::type *)ctx->dmm.ptr, (T *)ctx->dye2.ptr);
      else
        hipLaunchKernelGGL(kfull, gather_grid(d), kGatherBlock, 0, ctx->stream, d, dth, s, cv,
                           (const T *)ctx->dye.ptr, (const T *)ctx->dpred.ptr, (T *)ctx->dye2.ptr);
    });
  }
  LAUNCHCHK();
  ctx->dmm_valid = false;
  std::swap(ctx->dye, ctx->dye2);
  return MAC3D_OK;
}

// ---- multigrid
// CONC_PREDICT on the lysed plane (the dye predictor with the reference's
// branches) into dpred, then CONC_RESOLVE of the four planes into the twin
int launch_conc_predict(Ctx *ctx, double dt) {
  const Dims &dc = ctx->dcg;
  ProfScope ps(ctx->prof, ctx->stream, KC_DYE);
  hipLaunchKernelGGL(k_conc_predict, cell_grid(dc), kCellBlock, 0, ctx->stream, dc, ctx->d[0],
                     conc_scene(ctx), (float)dt, (const float3 *)ctx->cvel.ptr,
                     (const float *)ctx->conc.ptr + dc.cells(), (float *)ctx->cpred.ptr);
  LAUNCHCHK();
  return MAC3D_OK;
}
int launch_conc_resolve(Ctx *ctx, double dt) {
  const Dims &dc = ctx->dcg;
  if (!ctx->splat || ctx->splat_cap < dc.cells()) return fail(ctx, MAC3D_ESTATE, "no splat counters");
  {
    ProfScope ps(ctx->prof, ctx->stream, KC_DYE);
    hipLaunchKernelGGL(k_conc_resolve, cell_grid(dc), kCellBlock, 0, ctx->stream, dc, ctx->d[0],
                       conc_scene(ctx), (float)dt, (float)ctx->cfg.maccormack,
                       (const float3 *)ctx->cvel.ptr, (const float *)ctx->conc.ptr,
                       (const float *)ctx->cpred.ptr, (const unsigned *)ctx->splat,
                       (const unsigned *)ctx->splat + dc.cells(), (float *)ctx->conc2.ptr);
  }
  LAUNCHCHK();
  std::swap(ctx->conc, ctx->conc2);
  return MAC3D_OK;
}

void swap_x(Ctx *ctx, int l) {
  std::swap(ctx->mgx[l], ctx->mgxa[l]);
  if (l == 0 && ctx->prec != MAC3D_PREC_MIXED && !ctx->adj_solving) ctx->p = ctx->mgx[0];
}

// z-chunk of the level-0 sweep kernel: about three 512-thread blocks per CU
// (256 CUs) in one round, chunks of at least 8 planes
int sweep_cz(const Dims &d, size_t elem) {
  int cols = ((d.nx + SW_TX - 1) / SW_TX) * ((d.ny + SW_TY - 1) / SW_TY);
  int chunks = std::max(1, 768 / cols);
  int cz = std::max((d.nz + chunks - 1) / chunks, std::min(d.nz, 8));
  size_t pb = (size_t)d.nx * d.ny * elem;  // the chunk's buffer ranges stay below 2^30 bytes
  while (cz > 1 && (size_t)(cz + 8) * pb >= ((size_t)1 << 30)) --cz;
  return cz;
}

// One red-black sweep on level l.  Level 0: the fused out-of-place kernel,
// mgx -> mgxa, then the two swap roles; its right-hand side is multiplied by
// scale[0] (mixed mode: 1/max|r|).  Coarser levels: two in-place colour passes.
int mg_sweep(Ctx *ctx, int l, bool zero) {
  const Dims &d = ctx->d[l];
  if (l > 0) {
    ProfScope ps(ctx->prof, ctx->stream, KC_MG);
    dim3 grid(((d.nx + 1) / 2 + 63) / 64, (d.ny + 3) / 4, d.nz);
    with_type(ctx->mgx[l].type, [&](auto tag) {
      using S = std::remove_pointer_t<decltype(tag)>;
      using C = typename compute_of<S>::type;
      C h2 = (C)level_h2(l), tc = (C)top_coeff(l);
      if (zero)
        hipLaunchKernelGGL((k_rbgs_pass<S, true>), grid, kCellBlock, 0, ctx->stream, d,
                           (S *)ctx->mgx[l].ptr, (const S *)ctx->mgb[l].ptr, h2, tc,
                           (C)ctx->omega, 0);
      else
        hipLaunchKernelGGL((k_rbgs_pass<S, false>), grid, kCellBlock, 0, ctx->stream, d,
                           (S *)ctx->mgx[l].ptr, (const S *)ctx->mgb[l].ptr, h2, tc,
                           (C)ctx->omega, 0);
      hipLaunchKernelGGL((k_rbgs_pass<S, false>), grid, kCellBlock, 0, ctx->stream, d,
                         (S *)ctx->mgx[l].ptr, (const S *)ctx->mgb[l].ptr, h2, tc,
                         (C)ctx->omega, 1);
    });
    LAUNCHCHK();
    return MAC3D_OK;
  }
  int cz = sweep_cz(d, ctx->mgx[0].type == TF64 ? 8 : 4);
  int ncx = (d.nx + SW_TX - 1) / SW_TX, ncy = (d.ny + SW_TY - 1) / SW_TY;
  dim3 grid(ncx * ncy * ((d.nz + cz - 1) / cz));
  const double *sc = ctx->scale;
  {
    ProfScope ps(ctx->prof, ctx->stream, KC_SMOOTH0);
    with_step_type(ctx->mgx[0].type, [&](auto tag) {
      using S = std::remove_pointer_t<decltype(tag)>;
      using C = typename compute_of<S>::type;
      if (zero)
        hipLaunchKernelGGL((k_rbgs_sweep<S, true>), grid, dim3(SW_THREADS), 0, ctx->stream, d,
                           (const S *)ctx->mgx[0].ptr, (S *)ctx->mgxa[0].ptr,
                           (const S *)ctx->mgb[0].ptr, (C)1, (C)1, (C)ctx->omega, sc, cz,
                           ncx, ncy);
      else
        hipLaunchKernelGGL((k_rbgs_sweep<S, false>), grid, dim3(SW_THREADS), 0, ctx->stream, d,
                           (const S *)ctx->mgx[0].ptr, (S *)ctx->mgxa[0].ptr,
                           (const S *)ctx->mgb[0].ptr, (C)1, (C)1, (C)ctx->omega, sc, cz,
                           ncx, ncy);
    });
  }
  LAUNCHCHK();
  swap_x(ctx, 0);
  return MAC3D_OK;
}

int mg_fill_zero(Ctx *ctx, Buf &b) {
  size_t n = b.n;
  with_type(b.type, [&](auto tag) {
    using S = std::remove_pointer_t<decltype(tag)>;
    hipLaunchKernelGGL(k_fill<S>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream,
                       (S *)b.ptr, n, (S)0);
  });
  LAUNCHCHK();
  return MAC3D_OK;
}

// Tile width of the two-sweep pass: 32 on grids of up to 2^18 cells (twice
// the blocks of a 64-wide tile, measured faster on 48x64x48); 48 where 64-wide
// tiles would overhang the grid and 48-wide ones fit it (96 columns: 16.5
// against 17.5 us on 96x128x96); else 64.
static int sweep2_tx(const Dims &d) {
  if (d.cells() <= ((size_t)1 << 18)) return 32;
  if (d.nx % 64 != 0 && d.nx % 48 == 0) return 48;
  return 64;
}
#define SWEEP2_TY 16

// z-chunk of the two-sweep pass: about two blocks per CU (the LDS of two
// fills a CU), at least 6 planes (each chunk pays seven partial iterations at
// its start), even (the kernel's unrolled halves have fixed plane parities),
// and a chunk's buffer ranges below 2^30 bytes
int sweep2_cz(const Dims &d, int tx) {
  int cols = ((d.nx + tx - 1) / tx) * ((d.ny + SWEEP2_TY - 1) / SWEEP2_TY);
  int chunks = std::max(1, (64 * 512 / tx) / cols);
  // grids of up to 2^18 cells: 2-plane chunks (measured fastest on 48x64x48)
  if ((size_t)d.nx * d.ny * d.nz <= ((size_t)1 << 18)) return 2;
  int cz = std::max((d.nz + chunks - 1) / chunks, std::min(d.nz, 6));
  cz = (cz + 1) & ~1;
  size_t pb = (size_t)d.nx * d.ny * sizeof(float);
  while (cz > 2 && (size_t)(cz + 16) * pb >= ((size_t)1 << 30)) cz -= 2;
  return cz;
}

// Tile shapes of the coarse levels' two-sweep pass (k_rbgs_tile2): interior
// TX x TY x TZ cells and NT threads per block (at least one per column of the
// staged x ring and of R1's region).
struct Tile2Shape { int tx, ty, tz, nt; };
// three x-y families (32 x 8, 16 x 8, 48 x 8), each 8, 6 or 2 planes deep
static constexpr Tile2Shape kTile2Shapes[] = {
    {32, 8, 8, 320}, {16, 8, 8, 192}, {48, 8, 8, 448}, {32, 8, 6, 320}, {16, 8, 6, 192},
    {48, 8, 6, 448}, {32, 8, 2, 320}, {16, 8, 2, 192}, {48, 8, 2, 448}};

// Measured on 192x256x192's coarse levels (us per launch, given / zero guess;
// z-marching pass 16.3 / 14.5 on level 1, 11.0 / 10.1 on level 2):
//   96x128x96: 48x8x8 14.5 / 12.5, 16x8x8 16.1 / 13.4, 32x8x8 20.0 / 13.6;
//   48x64x48:  16x8x8  8.1 / 7.5, 48x8x8 9.6 / 8.6, 32x8x8 8.8 / 8.1.
// Small grids take the narrow tile (more blocks); larger ones the widest tile
// that divides the rows (less ring work per cell).  The depth is then the
// largest of 8, 6, 2 planes that still gives every CU two blocks (else 2):
// levels 1-3 of 192x256x192 per FP32 step, 643 us with 8-plane tiles, 601
// with 48x8x6 on 96x128x96 (512 blocks; 626 with 4 planes, 707 with 2) and
// 16x8x2 on 48x64x48 (576 blocks; 608 with 6 planes, 608 with 4).
static int tile2_shape(const Dims &d, int ncu) {
  const int fam = d.cells() <= ((size_t)1 << 19) ? 1 : (d.nx % 48 == 0 ? 2 : (d.nx % 32 == 0 ? 0 : 1));
  const Tile2Shape &f = kTile2Shapes[fam];
  const size_t cols = (size_t)((d.nx + f.tx - 1) / f.tx) * ((d.ny + f.ty - 1) / f.ty);
  for (int q = 0; q < 2; ++q) {  // depths 8, then 6
    const int tz = kTile2Shapes[fam + 3 * q].tz;
    if (cols * ((d.nz + tz - 1) / tz) >= (size_t)2 * ncu) return fam + 3 * q;
  }
  return fam + 6;
}

// The coarse levels' two-sweep pass: one 3-D tile per block (k_rbgs_tile2).
int mg_tile2(Ctx *ctx, int l, bool zero) {
  const Dims &d = ctx->d[l];
  const int shape = tile2_shape(d, ctx->ncu);
  const Tile2Shape sh = kTile2Shapes[shape];
  const int ncx = (d.nx + sh.tx - 1) / sh.tx, ncy = (d.ny + sh.ty - 1) / sh.ty;
  const int ncz = (d.nz + sh.tz - 1) / sh.tz;
  dim3 grid(ncx * ncy * ncz);
  {
    ProfScope ps(ctx->prof, ctx->stream, l == 0 ? KC_SMOOTH0X2 : KC_MG);
    auto launch = [&](auto stag, auto w1tag, auto ztag, auto shtag) {
      using S = std::remove_pointer_t<decltype(stag)>;
      constexpr bool W1 = decltype(w1tag)::value, Z = decltype(ztag)::value;
      constexpr int SH = decltype(shtag)::value;
      constexpr Tile2Shape T = kTile2Shapes[SH];
      hipLaunchKernelGGL((k_rbgs_tile2<Z, S, W1, T.tx, T.ty, T.tz, T.nt>), grid, dim3(T.nt), 0,
                         ctx->stream, d, (const S *)ctx->mgx[l].ptr, (S *)ctx->mgxa[l].ptr,
                         (const S *)ctx->mgb[l].ptr, (float)level_h2(l),
                         (float)(l == 0 ? 1.0 : top_coeff(l)), (float)ctx->omega,
                         l == 0 ? (const double *)ctx->scale : nullptr, ncx, ncy);
    };
    auto by_shape = [&](auto stag, auto w1tag, auto ztag) {
      switch (shape) {
        case 1: launch(stag, w1tag, ztag, std::integral_constant<int, 1>{}); break;
        case 2: launch(stag, w1tag, ztag, std::integral_constant<int, 2>{}); break;
        case 3: launch(stag, w1tag, ztag, std::integral_constant<int, 3>{}); break;
        case 4: launch(stag, w1tag, ztag, std::integral_constant<int, 4>{}); break;
        case 5: launch(stag, w1tag, ztag, std::integral_constant<int, 5>{}); break;
        case 6: launch(stag, w1tag, ztag, std::integral_constant<int, 6>{}); break;
        case 7: launch(stag, w1tag, ztag, std::integral_constant<int, 7>{}); break;
        case 8: launch(stag, w1tag, ztag, std::integral_constant<int, 8>{}); break;
        default: launch(stag, w1tag, ztag, std::integral_constant<int, 0>{}); break;
      }
    };
    auto by_zero = [&](auto stag, auto w1tag) {
      if (zero) by_shape(stag, w1tag, std::true_type{});
      else by_shape(stag, w1tag, std::false_type{});
    };
    const bool w1 = (float)ctx->omega == 1.0f;
    if (ctx->mgx[l].type == TF16) {
      if (w1) by_zero((half_t *)nullptr, std::true_type{});
      else by_zero((half_t *)nullptr, std::false_type{});
    } else {
      if (w1) by_zero((float *)nullptr, std::true_type{});
      else by_zero((float *)nullptr, std::false_type{});
    }
  }
  LAUNCHCHK();
  swap_x(ctx, l);
  return MAC3D_OK;
}

// Two red-black sweeps on level l in one pass (FP32 or FP16 storage): mgx -> mgxa,
// then the two swap roles.  zero: the first sweep starts from a zero guess.
// Levels >= 1, and level 0 of grids a z-march would not fill, take the tile
// pass; march: the z-marching pass on any level (parity tests of the two).
int mg_sweep2(Ctx *ctx, int l, bool zero, bool march = false) {
  const Dims &d = ctx->d[l];
  // (the tile pass takes x-pairs and 32-bit byte offsets, a level and its
  // ring planes under 2^30 bytes; on level 0 it takes grids of up to 2^21
  // cells, where a z-march has few planes per block to fill its pipeline:
  // 96x128x96 14.1 / 12.2 against 16.3 / 14.5 us)
  const size_t es = type_size(ctx->mgx[l].type);
  if (!march && (l > 0 || d.cells() <= ((size_t)1 << 21)) && d.nx % 2 == 0 && d.nz >= 2 &&
      (d.cells() + (size_t)8 * d.nx * d.ny) * es < ((size_t)1 << 30))
    return mg_tile2(ctx, l, zero);
  const int tx = sweep2_tx(d);
  int cz = sweep2_cz(d, tx);
  int ncx = (d.nx + tx - 1) / tx, ncy = (d.ny + SWEEP2_TY - 1) / SWEEP2_TY;
  dim3 grid(ncx * ncy * ((d.nz + cz - 1) / cz));
  {
    ProfScope ps(ctx->prof, ctx->stream, l == 0 ? KC_SMOOTH0X2 : KC_MG);
    auto launch = [&](auto stag, auto w1tag) {
      using S = std::remove_pointer_t<decltype(stag)>;
      constexpr bool W1 = decltype(w1tag)::value;
      auto kern = tx == 64   ? (zero ? k_rbgs_sweep2<true, 64, SWEEP2_TY, S, W1>
                                     : k_rbgs_sweep2<false, 64, SWEEP2_TY, S, W1>)
                  : tx == 48 ? (zero ? k_rbgs_sweep2<true, 48, SWEEP2_TY, S, W1>
                                     : k_rbgs_sweep2<false, 48, SWEEP2_TY, S, W1>)
                             : (zero ? k_rbgs_sweep2<true, 32, SWEEP2_TY, S, W1>
                                     : k_rbgs_sweep2<false, 32, SWEEP2_TY, S, W1>);
      hipLaunchKernelGGL(kern, grid, dim3(tx / 2 * SWEEP2_TY), 0, ctx->stream, d,
                         (const S *)ctx->mgx[l].ptr, (S *)ctx->mgxa[l].ptr,
                         (const S *)ctx->mgb[l].ptr, (float)level_h2(l),
                         (float)(l == 0 ? 1.0 : top_coeff(l)), (float)ctx->omega,
                         l == 0 ? (const double *)ctx->scale : nullptr, cz, ncx, ncy);
    };
    // omega exactly 1 in the kernel's FP32: the variant without the blend
    const bool w1 = (float)ctx->omega == 1.0f;
    if (ctx->mgx[l].type == TF16) {
      if (w1) launch((half_t *)nullptr, std::true_type{});
      else launch((half_t *)nullptr, std::false_type{});
    } else {
      if (w1) launch((float *)nullptr, std::true_type{});
      else launch((float *)nullptr, std::false_type{});
    }
  }
  LAUNCHCHK();
  swap_x(ctx, l);
  return MAC3D_OK;
}

// sweeps red-black sweeps on level l.  With FP32 or FP16 storage, pairs of
// sweeps run as one two-sweep pass: the z-marching pass on level 0, the tile
// pass on coarser levels of at least 512 cells with even rows (one launch
// against four colour passes: the 32x48x32 workload 1266 -> 1523 steps/s,
// 256x384x256 198.8 -> 205.6, whose levels 3 and 4 take it).
int mg_smooth(Ctx *ctx, int l, int sweeps, bool zero) {
  int rc;
  if (zero && sweeps == 0) return mg_fill_zero(ctx, ctx->mgx[l]);
  int s = 0;
  const Dims &d = ctx->d[l];
  if ((ctx->mgx[l].type == TF32 || ctx->mgx[l].type == TF16) &&
      (l == 0 || (ctx->n[l] >= 512 && d.nx % 2 == 0 && d.nz >= 2)))
    for (; s + 2 <= sweeps; s += 2)
      if ((rc = mg_sweep2(ctx, l, zero && s == 0))) return rc;
  for (; s < sweeps; ++s)
    if ((rc = mg_sweep(ctx, l, zero && s == 0))) return rc;
  return MAC3D_OK;
}

// The plain FP32 step forms the level-0 residual in FP64 arithmetic (FP32
// storage): the accuracy the cycles can reach is set by the finest level's
// residual, and an FP32 one stalls the smoothest mode at ~eps/lambda_min.
bool fine_residual_fp64(const Ctx *ctx, int l) {
  return l == 0 && ctx->prec == MAC3D_PREC_FP32 && ctx->mgx[0].type == TF32;
}

int mg_residual(Ctx *ctx, int l) {
  const Dims &d = ctx->d[l];
  ProfScope ps(ctx->prof, ctx->stream, l == 0 ? KC_RESID0 : KC_MG);
  const bool acc64 = fine_residual_fp64(ctx, l);
  with_type(ctx->mgx[l].type, [&](auto tag) {
    using S = std::remove_pointer_t<decltype(tag)>;
    using C = typename compute_of<S>::type;
    auto kern = k_residual<S>;
    if constexpr (std::is_same<S, float>::value)
      if (acc64) kern = k_residual<S, double>;
    hipLaunchKernelGGL(kern, cell_grid(d), kCellBlock, 0, ctx->stream, d,
                       (const S *)ctx->mgx[l].ptr, (const S *)ctx->mgb[l].ptr,
                       (S *)ctx->mgr[l].ptr, (C)level_h2(l),
                       (C)(l == 0 ? 1.0 : top_coeff(l)), l == 0 ? (const double *)ctx->scale : nullptr);
  });
  LAUNCHCHK();
  return MAC3D_OK;
}

int mg_residual_restrict(Ctx *ctx, int l) {
  const Dims &df = ctx->d[l], &dc = ctx->d[l + 1];
  ProfScope ps(ctx->prof, ctx->stream, l == 0 ? KC_RESID0 : KC_MG);
  if (l == 0 || (ctx->mgx[l].type != TF16 && ctx->n[l] >= ((size_t)1 << 20))) {
    // wave-column kernel (level 0, and FP32/FP64 levels of 2^20 cells and
    // up): even chunks of at least 6 planes, about 24 waves per CU
    int wcx = (df.nx + 63) / 64, cols = wcx * (df.ny / 2);
    int chunks = std::max(1, 6144 / cols);
    int cz = (df.nz + chunks - 1) / chunks;
    cz = std::max(std::min(df.nz, 6), (cz + 1) & ~1);
    // the chunk's buffer ranges and byte offsets (loads run up to four planes
    // past the chunk) stay below 2^30
    const size_t pb = (size_t)df.nx * df.ny * type_size(ctx->mgx[l].type);
    while (cz > 2 && (size_t)(cz + 8) * pb >= ((size_t)1 << 30)) cz -= 2;
    int nw = cols * ((df.nz + cz - 1) / cz);
    dim3 grid((nw + RC_WAVES - 1) / RC_WAVES);
    const bool acc64 = fine_residual_fp64(ctx, l);
    with_step_type(ctx->mgx[l].type, [&](auto ftag) {
      using Sf = std::remove_pointer_t<decltype(ftag)>;
      using C = typename compute_of<Sf>::type;
      with_type(ctx->mgb[l + 1].type, [&](auto ctag) {
        using Sc = std::remove_pointer_t<decltype(ctag)>;
        auto kern = k_residual_restrict_col<Sf, Sc>;
        if constexpr (std::is_same<Sf, float>::value && std::is_same<Sc, float>::value)
          if (acc64) kern = k_residual_restrict_col<Sf, Sc, double>;
        hipLaunchKernelGGL(kern, grid, dim3(64 * RC_WAVES), 0,
                           ctx->stream, df, dc, (const Sf *)ctx->mgx[l].ptr,
                           (const Sf *)ctx->mgb[l].ptr, (Sc *)ctx->mgb[l + 1].ptr,
                           (C)level_h2(l), (C)(l == 0 ? 1.0 : top_coeff(l)),
                           l == 0 ? (const double *)ctx->scale : nullptr, cz, wcx, nw);
      });
    });
    LAUNCHCHK();
    return MAC3D_OK;
  }
  with_type(ctx->mgx[l].type, [&](auto ftag) {
    using Sf = std::remove_pointer_t<decltype(ftag)>;
    using C = typename compute_of<Sf>::type;
    with_type(ctx->mgb[l + 1].type, [&](auto ctag) {
      using Sc = std::remove_pointer_t<decltype(ctag)>;
      hipLaunchKernelGGL((k_residual_restrict_t<Sf, Sc>),
                         dim3((df.nx + 31) / 32, dc.ny, dc.nz), dim3(32, 2, 2), 0, ctx->stream,
                         df, dc, (const Sf *)ctx->mgx[l].ptr, (const Sf *)ctx->mgb[l].ptr,
                         (Sc *)ctx->mgb[l + 1].ptr, (C)level_h2(l),
                         (C)(l == 0 ? 1.0 : top_coeff(l)),
                         l == 0 ? (const double *)ctx->scale : nullptr);
    });
  });
  LAUNCHCHK();
  return MAC3D_OK;
}

int mg_restrict(Ctx *ctx, int l) {
  const Dims &df = ctx->d[l], &dc = ctx->d[l + 1];
  ProfScope ps(ctx->prof, ctx->stream, l == 0 ? KC_RESID0 : KC_MG);
  with_type(ctx->mgr[l].type, [&](auto ftag) {
    using Sf = std::remove_pointer_t<decltype(ftag)>;
    with_type(ctx->mgb[l + 1].type, [&](auto ctag) {
      using Sc = std::remove_pointer_t<decltype(ctag)>;
      hipLaunchKernelGGL((k_restrict<Sf, Sc>), cell_grid(dc), kCellBlock, 0, ctx->stream, df, dc,
                         (const Sf *)ctx->mgr[l].ptr, (Sc *)ctx->mgb[l + 1].ptr);
    });
  });
  LAUNCHCHK();
  return MAC3D_OK;
}

int mg_prolong(Ctx *ctx, int l) {
  const Dims &df = ctx->d[l], &dc = ctx->d[l + 1];
  ProfScope ps(ctx->prof, ctx->stream, l == 0 ? KC_PROLONG0 : KC_MG);
  with_type(ctx->mgx[l + 1].type, [&](auto ctag) {
    using Sc = std::remove_pointer_t<decltype(ctag)>;
    with_type(ctx->mgx[l].type, [&](auto ftag) {
      using Sf = std::remove_pointer_t<decltype(ftag)>;
      // 32 x 8 threads over the coarse grid (coarse rows of 96 fill three
      // blocks); the fine cells loaded early on grids of under 8 blocks per CU
      // (192x256x192's levels 1-3: 6 us per step faster; onto its level 0,
      // 4608 blocks, about 1 us per launch slower)
      const dim3 grid((dc.nx + PR_BX - 1) / PR_BX, (dc.ny + PR_BY - 1) / PR_BY, dc.nz);
      const bool pre = (size_t)grid.x * grid.y * grid.z < (size_t)8 * ctx->ncu;
      auto kern = pre ? k_prolong_add_2x<Sc, Sf, true> : k_prolong_add_2x<Sc, Sf, false>;
      hipLaunchKernelGGL(kern, grid, dim3(PR_BX, PR_BY), 0, ctx->stream, df, dc,
                         (const Sc *)ctx->mgx[l + 1].ptr, (Sf *)ctx->mgx[l].ptr);
    });
  });
  LAUNCHCHK();
  return MAC3D_OK;
}

int mg_coarse(Ctx *ctx) {
  int L = ctx->L;
  const Dims &d = ctx->d[L];
  const double *qx = ctx->cq + ctx->cq_off[0], *qy = ctx->cq + ctx->cq_off[1],
               *qz = ctx->cq + ctx->cq_off[2], *lx = ctx->cq + ctx->cq_off[3],
               *ly = ctx->cq + ctx->cq_off[4], *lz = ctx->cq + ctx->cq_off[5];
  double h2 = level_h2(L);
  ProfScope ps(ctx->prof, ctx->stream, KC_COARSE);
  bool c64 = ctx->prec != MAC3D_PREC_FP32;
  // eigenvector matrices in LDS when they fit in 48 KB (as FP64, the larger)
  size_t qxz = ((size_t)d.nx * d.nx + (size_t)d.nz * d.nz), qyy = (size_t)d.ny * d.ny;
  bool ql = std::max(qxz, qyy) * sizeof(double) <= 48 * 1024;
  with_type(ctx->mgx[L].type, [&](auto tag) {
    using S = std::remove_pointer_t<decltype(tag)>;
    auto run = [&](auto ctag, auto qtag) {
      using C = std::remove_pointer_t<decltype(ctag)>;
      constexpr bool QL = decltype(qtag)::value;
      size_t sh_xz = QL ? qxz * sizeof(C) : 0, sh_y = QL ? qyy * sizeof(C) : 0;
      hipLaunchKernelGGL((k_coarse_xz_fwd<S, C, QL>), dim3(d.ny), dim3(COARSE_T), sh_xz, ctx->stream,
                         d, (const S *)ctx->mgb[L].ptr, qx, qz, ctx->cwork);
      hipLaunchKernelGGL((k_coarse_y<C, QL>), dim3(d.nz), dim3(COARSE_T), sh_y, ctx->stream, d, qy,
                         lx, ly, lz, h2, ctx->cwork);
      hipLaunchKernelGGL((k_coarse_xz_inv<S, C, QL>), dim3(d.ny), dim3(COARSE_T), sh_xz, ctx->stream,
                         d, (const double *)ctx->cwork, qx, qz, (S *)ctx->mgx[L].ptr);
    };
    if (c64) {
      if (ql) run((double *)nullptr, std::true_type());
      else run((double *)nullptr, std::false_type());
    } else {
      if (ql) run((float *)nullptr, std::true_type());
      else run((float *)nullptr, std::false_type());
    }
  });
  LAUNCHCHK();
  return MAC3D_OK;
}

// One V-cycle on (mgx[0], mgb[0]) with the schedule of
// BloodDropEngine.ts:1942-2059: smooth / residual / restrict down to the
// coarsest level, exact coarse solve, prolong / smooth back up.  Levels >= 1
// start from a zero guess.  zero0: level 0 also starts from zero.  Level 0's
// right-hand side is scaled by scale[0] (1 outside the refinement loop).
int vcycle(Ctx *ctx, bool zero0) {
  int rc, L = ctx->L;
  int nu1 = ctx->cfg.nu1, nu2 = ctx->cfg.nu2;
  if ((rc = mg_smooth(ctx, 0, nu1, zero0))) return rc;
  if ((rc = mg_residual_restrict(ctx, 0))) return rc;
  for (int l = 1; l < L; ++l) {
    if ((rc = mg_smooth(ctx, l, nu1, true))) return rc;
    if ((rc = mg_residual_restrict(ctx, l))) return rc;
  }
  if ((rc = mg_coarse(ctx))) return rc;
  for (int l = L - 1; l >= 1; --l) {
    if ((rc = mg_prolong(ctx, l))) return rc;
    if ((rc = mg_smooth(ctx, l, nu2, false))) return rc;
  }
  if ((rc = mg_prolong(ctx, 0))) return rc;
  if ((rc = mg_smooth(ctx, 0, nu2, false))) return rc;
  return MAC3D_OK;
}

size_t cell_blocks(const Dims &d) {
  dim3 g = cell_grid(d);
  return (size_t)g.x * g.y * g.z;
}

// ---- mixed mode: FP64 pressure ring
void sync_p(Ctx *ctx) {
  if (ctx->prec == MAC3D_PREC_MIXED) ctx->p = ctx->pr[ctx->role[0]];
}

// fold the residual partials, set the next correction's scale, and queue the
// norms' copy to pinned host memory
// z-chunks of the FP64 residual kernels: about three 512-thread blocks per CU,
// a chunk's buffer ranges below 2^30 bytes
struct R64Grid {
  int cz, ncx, ncy, blocks;
};
R64Grid r64_grid(const Dims &d) {
  R64Grid g;
  g.ncx = (d.nx + R64_TX - 1) / R64_TX;
  g.ncy = (d.ny + R64_TY - 1) / R64_TY;
  int chunks = std::max(1, 768 / (g.ncx * g.ncy));
  g.cz = (d.nz + chunks - 1) / chunks;
  size_t pb = (size_t)d.nx * d.ny * 8;
  while (g.cz > 1 && (size_t)(g.cz + 4) * pb >= ((size_t)1 << 30)) --g.cz;
  g.blocks = g.ncx * g.ncy * ((d.nz + g.cz - 1) / g.cz);
  return g;
}

// The residual norms (max |r|, ||r||^2, ||b||^2) go to host slot `slot`.
int fold_norms(Ctx *ctx, int slot) {
  hipLaunchKernelGGL(k_fold3, dim3(1), dim3(256), 0, ctx->stream, (const double *)ctx->part,
                     (size_t)r64_grid(ctx->d[0]).blocks, ctx->red, ctx->scale);
  LAUNCHCHK();
  HIPCHK(hipMemcpyAsync(ctx->red_host + 3 * slot, ctx->red, 3 * sizeof(double),
                        hipMemcpyDeviceToHost, ctx->stream));
  return MAC3D_OK;
}

// First residual of a step: the warm start extrapolated from the pressure
// history (order min(extrap, hist)), formed into the spare buffer, which then
// becomes current while the history shifts.  order 0 restarts from p.
int mixed_begin(Ctx *ctx, int order, int slot = 0) {
  const Dims &d = ctx->d[0];
  double c0 = 1.0, c1 = 0.0, c2 = 0.0;
  if (order >= 2) { c0 = 3.0; c1 = -3.0; c2 = 1.0; }
  else if (order == 1) { c0 = 2.0; c1 = -1.0; }
  int *r = ctx->role;
  {
    ProfScope ps(ctx->prof, ctx->stream, KC_RESID64);
    R64Grid g = r64_grid(d);
    hipLaunchKernelGGL(k_extrap_residual64_zm, dim3(g.blocks), dim3(R64_TX, R64_TY), 0,
                       ctx->stream, d, (const double *)ctx->pr[r[0]].ptr,
                       (const double *)ctx->pr[r[1]].ptr, (const double *)ctx->pr[r[2]].ptr, c0,
                       c1, c2, (const double *)ctx->rhs.ptr, (double *)ctx->pr[r[3]].ptr,
                       (float *)ctx->r0.ptr, ctx->part, g.cz, g.ncx, g.ncy);
  }
  LAUNCHCHK();
  int nr[4] = {r[3], r[0], r[1], r[2]};
  for (int q = 0; q < 4; ++q) r[q] = nr[q];
  sync_p(ctx);
  return fold_norms(ctx, slot);
}

// One refinement iteration: correction V-cycle on r/max|r| from zero, then
// p <- p + max|r| e formed into the spare buffer together with the new
// residual; spare and current swap.
int mixed_iter(Ctx *ctx, int slot = 0) {
  int rc;
  if ((rc = vcycle(ctx, true))) return rc;
  const Dims &d = ctx->d[0];
  int *r = ctx->role;
  {
    ProfScope ps(ctx->prof, ctx->stream, KC_RESID64);
    R64Grid g = r64_grid(d);
    hipLaunchKernelGGL(k_update_residual64_zm, dim3(g.blocks), dim3(R64_TX, R64_TY), 0,
                       ctx->stream, d, (const double *)ctx->pr[r[0]].ptr,
                       (const float *)ctx->mgx[0].ptr, (const double *)ctx->scale,
                       (const double *)ctx->rhs.ptr, (double *)ctx->pr[r[3]].ptr,
                       (float *)ctx->r0.ptr, ctx->part, g.cz, g.ncx, g.ncy);
  }
  LAUNCHCHK();
  std::swap(r[0], r[3]);
  sync_p(ctx);
  return fold_norms(ctx, slot);
}

int div_norms(Ctx *ctx, double *mx, double *l2) {
  const Dims &d = ctx->d[0];
  with_step_type(ctx->t0, [&](auto tag) {
    using T = std::remove_pointer_t<decltype(tag)>;
    hipLaunchKernelGGL(k_div_partials<T>, cell_grid(d), kCellBlock, 0, ctx->stream, d,
                       (const T *)ctx->u.ptr, (const T *)ctx->v.ptr, (const T *)ctx->w.ptr,
                       ctx->part);
  });
  hipLaunchKernelGGL(k_fold_partials, dim3(1), dim3(256), 0, ctx->stream,
                     (const double *)ctx->part, cell_blocks(d), ctx->red);
  LAUNCHCHK();
  HIPCHK(hipMemcpyAsync(ctx->red_host, ctx->red, 2 * sizeof(double), hipMemcpyDeviceToHost,
                        ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  *mx = ctx->red_host[0];
  *l2 = std::sqrt(ctx->red_host[1] / (double)ctx->n[0]);
  return MAC3D_OK;
}

// ---- launch sequences and their graphs
// Parity bits of the ping-pong buffers: bit 0 velocity, bit 1 dye, bit 2+l
// multigrid level l.  A captured sequence is valid for the parities it was
// captured at and leaves a fixed set of parities behind.
uint64_t parity_bits(const Ctx *ctx) {
  uint64_t b = 0;
  if (ctx->u.ptr != ctx->u_home) b |= 1;
  if (ctx->dye.ptr != ctx->dye_home) b |= 2;
  for (int l = 0; l <= ctx->L; ++l)
    if (ctx->mgx[l].ptr != ctx->mgx_home[l]) b |= (uint64_t)4 << l;
  if (ctx->conc.ptr && ctx->conc.ptr != ctx->conc_home) b |= (uint64_t)1 << 12;
  for (int q = 0; q < 4; ++q) b |= (uint64_t)ctx->role[q] << (16 + 2 * q);
  return b;
}

void set_parity_bits(Ctx *ctx, uint64_t b) {
  if (((ctx->u.ptr != ctx->u_home) ? 1u : 0u) != (b & 1)) swap_vel(ctx);
  if (((ctx->dye.ptr != ctx->dye_home) ? 1u : 0u) != ((b >> 1) & 1)) std::swap(ctx->dye, ctx->dye2);
  for (int l = 0; l <= ctx->L; ++l)
    if (((ctx->mgx[l].ptr != ctx->mgx_home[l]) ? 1u : 0u) != ((b >> (2 + l)) & 1)) swap_x(ctx, l);
  if (ctx->conc.ptr && ((ctx->conc.ptr != ctx->conc_home) ? 1u : 0u) != ((b >> 12) & 1))
    std::swap(ctx->conc, ctx->conc2);
  for (int q = 0; q < 4; ++q) ctx->role[q] = (int)((b >> (16 + 2 * q)) & 3);
  sync_p(ctx);
}

// Run a launch sequence: eagerly while profiling, otherwise through a graph
// captured on first use for the current parities and dt.
template <class F>
int run_seq(Ctx *ctx, int kind, double dt, F &&enqueue, int variant = 0) {
  if (!ctx->use_graphs || ctx->prof.enabled) return enqueue();
  uint64_t before = parity_bits(ctx);
  uint64_t key = ((uint64_t)kind << 40) | ((uint64_t)variant << 32) | before;
  auto it = ctx->graphs.find(key);
  if (it == ctx->graphs.end() || it->second.dt != dt) {
    if (it != ctx->graphs.end()) {
      (void)hipGraphExecDestroy(it->second.exec);
      ctx->graphs.erase(it);
    }
    HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue();
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      set_parity_bits(ctx, before);
      return rc;
    }
    if (e != hipSuccess) return fail(ctx, MAC3D_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
    Ctx::Graph gr;
    gr.dt = dt;
    gr.after = parity_bits(ctx);
    hipError_t ei = hipGraphInstantiate(&gr.exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) return fail(ctx, MAC3D_EHIP, std::string("graph instantiate: ") + hipGetErrorString(ei));
    set_parity_bits(ctx, before);
    it = ctx->graphs.emplace(key, gr).first;
  }
  HIPCHK(hipGraphLaunch(it->second.exec, ctx->stream));
  set_parity_bits(ctx, it->second.after);
  return MAC3D_OK;
}

void drop_graphs(Ctx *ctx) {
  for (auto &kv : ctx->graphs) (void)hipGraphExecDestroy(kv.second.exec);
  ctx->graphs.clear();
}

enum { SEQ_STEP = 1, SEQ_PRE = 2, SEQ_ITER = 3, SEQ_POST = 4 };

int seq_pre(Ctx *ctx, double dt) {
  int rc;
  if (ctx->wave_on) {  // the wave pass reads this step's starting velocity
    if ((rc = launch_wave_step(ctx, dt))) return rc;
    if ((rc = launch_top_pressure(ctx, dt))) return rc;
  }
  // the previous step's last pass already formed cvel from this velocity
  if (!ctx->cvel_fresh && (rc = launch_face_to_centre(ctx))) return rc;
  if ((rc = launch_advect(ctx, dt))) return rc;
  if ((rc = launch_casson(ctx, dt))) return rc;
  return launch_divergence(ctx);
}

int seq_post(Ctx *ctx, double dt) {
  int rc;
  if ((rc = launch_project(ctx))) return rc;
  if ((rc = launch_face_to_centre(ctx))) return rc;
  if (ctx->conc_on) {  // pass B (BloodDropEngine.ts:2073-2108)
    if ((rc = launch_parcels(ctx, dt))) return rc;
    if ((rc = launch_conc_predict(ctx, dt))) return rc;
    return launch_conc_resolve(ctx, dt);
  }
  if ((rc = launch_dye_predict(ctx, dt))) return rc;
  return launch_dye_resolve(ctx, dt);
}

// One step.  FP32/FP64: a fixed number of V-cycles on (P, RHS), warm-started
// from the previous step's P, all in one captured sequence.  Mixed: FP64
// iterative refinement around FP32 correction V-cycles until
// ||r||/||b|| <= ir_tol; each iteration is one captured sequence ending in
// the residual norms, which the host reads to decide whether to go on.
int one_step(Ctx *ctx, double dt, int *cycles, int *iters, double *rel, int *conv) {
  int rc;
  if (ctx->prec != MAC3D_PREC_MIXED) {
    rc = run_seq(ctx, SEQ_STEP, dt, [&]() {
      int r;
      if ((r = seq_pre(ctx, dt))) return r;
      for (int c = 0; c < ctx->cfg.cycles; ++c)
        if ((r = vcycle(ctx, false))) return r;
      return seq_post(ctx, dt);
    }, ctx->cvel_fresh ? 1 : 0);  // (graphs replay without running the lambda: the
    if (rc) return rc;              // host sets the flag after each successful run)
    ctx->cvel_fresh = true;
    *cycles += ctx->cfg.cycles;
    *iters = ctx->cfg.cycles;
    *rel = -1.0;
    *conv = 1;
    return MAC3D_OK;
  }
  int order = std::min(ctx->extrap, ctx->hist);
  // The first convergence test waits for the host; the iterations every
  // recent step needed run before it, in the same sequence, so that a step
  // stops for the host about once instead of once per iteration.  The norms
  // after j of them go to host slot j, so the step still learns how many it
  // needed (one that would have converged sooner refines a little further).
  int pre_it = 0;
  if (ctx->iter_n > 0) {
    pre_it = ctx->iter_log[0];
    for (int q = 1; q < ctx->iter_n; ++q) pre_it = std::min(pre_it, ctx->iter_log[q]);
    pre_it = std::min(std::min(pre_it, ctx->cfg.ir_max_iters), kMaxPreIters);
  }
  rc = run_seq(ctx, SEQ_PRE, dt, [&]() {
    int r;
    if ((r = seq_pre(ctx, dt))) return r;
    if ((r = mixed_begin(ctx, order, 0))) return r;
    for (int q = 0; q < pre_it; ++q)
      if ((r = mixed_iter(ctx, q + 1))) return r;
    return MAC3D_OK;
  }, order | (ctx->cvel_fresh ? 4 : 0) | (pre_it << 3));
  if (rc) return rc;
  auto converged = [&](const double *h, double *rr) {
    double r = h[2] > 0.0 ? std::sqrt(h[1] / h[2]) : 0.0;
    if (rr) *rr = r;
    return r <= ctx->cfg.ir_tol || h[0] == 0.0;
  };
  int it = pre_it, needed = -1;
  *conv = 0;
  for (bool first = true;; first = false) {
    // the next launch waits on this test: poll for the norms rather than
    // block in a stream synchronize, which may sleep before it wakes
    HIPCHK(hipEventRecord(ctx->norms_ready, ctx->stream));
    hipError_t q;
    while ((q = hipEventQuery(ctx->norms_ready)) == hipErrorNotReady) {
    }
    if (q != hipSuccess) return fail(ctx, MAC3D_EHIP, std::string("norms: ") + hipGetErrorString(q));
    {  // non-finite norms: the state went NaN or inf; stop refining it
      const double *h = ctx->red_host + (first ? 3 * pre_it : 0);
      if (!std::isfinite(h[0]) || !std::isfinite(h[1]) || !std::isfinite(h[2]))
        return fail(ctx, MAC3D_ENONFINITE,
                    "the pressure solve's residual norms are not finite (NaN or inf in the state)");
    }
    if (first) {
      for (int j = 0; j <= pre_it && needed < 0; ++j)
        if (converged(ctx->red_host + 3 * j, nullptr)) needed = j;
      converged(ctx->red_host + 3 * pre_it, rel);
      if (needed >= 0) { *conv = 1; break; }
    } else if (converged(ctx->red_host, rel)) {
      needed = it;
      *conv = 1;
      break;
    }
    if (it >= ctx->cfg.ir_max_iters) break;
    rc = run_seq(ctx, SEQ_ITER, 0.0, [&]() { return mixed_iter(ctx, 0); });
    if (rc) return rc;
    ++it;
  }
  *cycles += it;
  *iters = it;
  for (int q = 3; q > 0; --q) ctx->iter_log[q] = ctx->iter_log[q - 1];
  ctx->iter_log[0] = needed >= 0 ? needed : it;
  ctx->iter_n = std::min(ctx->iter_n + 1, 4);
  ctx->hist = std::min(ctx->hist + 1, 2);
  rc = run_seq(ctx, SEQ_POST, dt, [&]() { return seq_post(ctx, dt); });
  if (!rc) ctx->cvel_fresh = true;
  return rc;
}

Buf *field_buf(Ctx *ctx, int field, int level) {
  if (level < 0 || level > ctx->L) return nullptr;
  bool l0 = level == 0;
  switch (field) {
    case MAC3D_FIELD_U: return l0 ? &ctx->u : nullptr;
    case MAC3D_FIELD_V: return l0 ? &ctx->v : nullptr;
    case MAC3D_FIELD_W: return l0 ? &ctx->w : nullptr;
    case MAC3D_FIELD_P: return l0 ? &ctx->p : nullptr;
    case MAC3D_FIELD_DYE: return l0 ? &ctx->dye : nullptr;
    case MAC3D_FIELD_CU:
    case MAC3D_FIELD_CV:
    case MAC3D_FIELD_CW: return l0 ? &ctx->cvel : nullptr;
    case MAC3D_FIELD_RHS: return l0 ? &ctx->rhs : nullptr;
    case MAC3D_FIELD_DYE_PRED: return l0 ? &ctx->dpred : nullptr;
    case MAC3D_FIELD_MGX: return &ctx->mgx[level];
    case MAC3D_FIELD_MGB: return &ctx->mgb[level];
    case MAC3D_FIELD_MGR: return &ctx->mgr[level];
    case MAC3D_FIELD_WAVE: return l0 && ctx->wave_on ? &ctx->wave : nullptr;
    case MAC3D_FIELD_PTOP: return l0 && ctx->wave_on ? &ctx->ptop : nullptr;
    case MAC3D_FIELD_CURL: return l0 && ctx->curl.ptr ? &ctx->curl : nullptr;
    case MAC3D_FIELD_CONC: return l0 && ctx->conc_on ? &ctx->conc : nullptr;
    case MAC3D_FIELD_CONC_PRED: return l0 && ctx->conc_on ? &ctx->cpred : nullptr;
  }
  return nullptr;
}

int valid_config(const mac3d_config *c, std::string &why) {
  if (c->nx <= 0 || c->ny <= 0 || c->nz <= 0) { why = "grid sizes must be positive"; return 0; }
  if (c->levels < 1 || c->levels > kMaxLev) { why = "levels must be in [1, 8]"; return 0; }
  int m = 1 << c->levels;
  if (c->nx % m || c->ny % m || c->nz % m) {
    why = "every grid size must divide by 2^levels";
    return 0;
  }
  int cx = c->nx / m, cy = c->ny / m, cz = c->nz / m;
  if (cx * cz > COARSE_MAX_PLANE || cx * cy > COARSE_MAX_PLANE) {
    why = "coarsest level too large for the exact solve (nx*nz and nx*ny at the coarsest "
          "level must be <= 2048): add levels";
    return 0;
  }
  if (c->nu1 < 0 || c->nu2 < 0 || c->cycles < 0 || c->ir_max_iters < 0) {
    why = "sweep and cycle counts must be >= 0";
    return 0;
  }
  if (c->precision < 0 || c->precision > 2) { why = "unknown precision"; return 0; }
  if (c->extrapolate < -1 || c->extrapolate > 2) { why = "extrapolate must be -1, 0, 1 or 2"; return 0; }
  if (!(c->sor_omega >= 0.0 && c->sor_omega < 2.0)) { why = "sor_omega must be in [0, 2)"; return 0; }
  if (!(c->dx > 0) || !std::isfinite(c->dx)) { why = "dx must be finite and > 0"; return 0; }
  if (!(c->rho_w > 0) || !(c->rho_b > 0)) { why = "densities must be > 0"; return 0; }
  if ((size_t)c->nx * c->ny * c->nz > (size_t)1 << 31) { why = "grid too large"; return 0; }
  if ((size_t)c->nx * c->ny > (size_t)1 << 22) {
    why = "an x-y plane holds at most 2^22 cells (the level-0 sweeps address a chunk of "
          "planes with 32-bit offsets)";
    return 0;
  }
  return 1;
}

int create_impl(Ctx *ctx, const mac3d_config *cfg) {
  int rc;
  ctx->cfg = *cfg;
  ctx->L = cfg->levels;
  ctx->prec = cfg->precision;
  ctx->t0 = ctx->prec == MAC3D_PREC_FP64 ? TF64 : TF32;
  HIPCHK(hipSetDevice(cfg->device));
  HIPCHK(hipDeviceGetAttribute(&ctx->ncu, hipDeviceAttributeMultiprocessorCount, cfg->device));
  HIPCHK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  HIPCHK(hipEventCreateWithFlags(&ctx->norms_ready, hipEventDisableTiming));
  for (int l = 0; l <= ctx->L; ++l) {
    ctx->d[l] = Dims{cfg->nx >> l, cfg->ny >> l, cfg->nz >> l};
    ctx->n[l] = ctx->d[l].cells();
  }
  size_t n0 = ctx->n[0];
  int t0 = ctx->t0;
  Buf *f0[] = {&ctx->u, &ctx->v, &ctx->w, &ctx->u2, &ctx->v2, &ctx->w2, &ctx->dye,
               &ctx->dye2, &ctx->dpred};
  for (Buf *b : f0)
    if ((rc = alloc(ctx, *b, t0, n0))) return rc;
  if ((rc = alloc(ctx, ctx->cvel, t0, 3 * n0))) return rc;
  if ((rc = alloc(ctx, ctx->dmm, t0, 2 * n0))) return rc;
  bool mixed = ctx->prec == MAC3D_PREC_MIXED;
  // mixed: 1.25, measured fastest (cycles per step at 192x256x192 over 100
  // steps: 3.34 at 1.2, 3.06 at 1.24-1.25, 3.17 at 1.28)
  ctx->omega = cfg->sor_omega > 0.0 ? cfg->sor_omega : (mixed ? 1.25 : 1.0);
  ctx->extrap = mixed ? (cfg->extrapolate < 0 ? 2 : std::min(cfg->extrapolate, 2)) : 0;
  if (mixed) {
    for (int q = 0; q < 4; ++q)
      if ((rc = alloc(ctx, ctx->pr[q], TF64, n0))) return rc;
    ctx->p = ctx->pr[0];
  } else if ((rc = alloc(ctx, ctx->p, t0, n0))) return rc;
  if ((rc = alloc(ctx, ctx->rhs, mixed ? TF64 : t0, n0))) return rc;
  if (mixed) {
    if ((rc = alloc(ctx, ctx->e0, TF32, n0))) return rc;
    if ((rc = alloc(ctx, ctx->r0, TF32, n0))) return rc;
    ctx->mgx[0] = ctx->e0;
    ctx->mgb[0] = ctx->r0;
  } else {
    ctx->mgx[0] = ctx->p;
    ctx->mgb[0] = ctx->rhs;
  }
  if ((rc = alloc(ctx, ctx->mgxa[0], ctx->mgx[0].type, n0))) return rc;
  if ((rc = alloc(ctx, ctx->mgr[0], ctx->mgx[0].type, n0))) return rc;
  for (int l = 1; l <= ctx->L; ++l) {
    int t = mixed ? (l >= 2 ? TF16 : TF32) : t0;
    if ((rc = alloc(ctx, ctx->mgx[l], t, ctx->n[l]))) return rc;
    if ((rc = alloc(ctx, ctx->mgxa[l], t, ctx->n[l]))) return rc;
    if ((rc = alloc(ctx, ctx->mgb[l], t, ctx->n[l]))) return rc;
    if ((rc = alloc(ctx, ctx->mgr[l], t, ctx->n[l]))) return rc;
  }
  ctx->u_home = ctx->u.ptr;
  ctx->dye_home = ctx->dye.ptr;
  for (int l = 0; l <= ctx->L; ++l) ctx->mgx_home[l] = ctx->mgx[l].ptr;
  // coarse eigen-decomposition
  const Dims &dc = ctx->d[ctx->L];
  std::vector<double> qx, qy, qz, lx, ly, lz;
  axis_eigen(dc.nx, false, 0.0, qx, lx);
  axis_eigen(dc.ny, true, top_coeff(ctx->L), qy, ly);
  axis_eigen(dc.nz, false, 0.0, qz, lz);
  std::vector<double> all;
  std::vector<double> *parts[6] = {&qx, &qy, &qz, &lx, &ly, &lz};
  for (int a = 0; a < 6; ++a) {
    ctx->cq_off[a] = all.size();
    all.insert(all.end(), parts[a]->begin(), parts[a]->end());
  }
  if ((rc = dev_alloc(ctx, (void **)&ctx->cq, all.size() * sizeof(double)))) return rc;
  HIPCHK(hipMemcpyAsync(ctx->cq, all.data(), all.size() * sizeof(double), hipMemcpyHostToDevice,
                        ctx->stream));
  if ((rc = dev_alloc(ctx, (void **)&ctx->cwork, ctx->n[ctx->L] * sizeof(double)))) return rc;
  ctx->part_n = cell_blocks(ctx->d[0]);
  if ((rc = dev_alloc(ctx, (void **)&ctx->part, 3 * ctx->part_n * sizeof(double)))) return rc;
  if ((rc = dev_alloc(ctx, (void **)&ctx->red, 4 * sizeof(double)))) return rc;
  if ((rc = dev_alloc(ctx, (void **)&ctx->scale, 2 * sizeof(double)))) return rc;
  HIPCHK(hipHostMalloc(&ctx->red_host, 3 * (kMaxPreIters + 1) * sizeof(double),
                       hipHostMallocDefault));
  if ((rc = dev_alloc(ctx, (void **)&ctx->stage, n0 * sizeof(double)))) return rc;
  ctx->stage_n = n0;
  hipLaunchKernelGGL(k_fill<double>, dim3(1), dim3(256), 0, ctx->stream, ctx->scale, (size_t)2, 1.0);
  LAUNCHCHK();
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if ((rc = ctx->prof.init(ctx->err))) return rc;
  return MAC3D_OK;
}

void destroy_impl(Ctx *ctx) {
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  drop_graphs(ctx);
  for (void *p : ctx->allocs) (void)hipFree(p);
  ctx->allocs.clear();
  if (ctx->red_host) (void)hipHostFree(ctx->red_host);
  ctx->red_host = nullptr;
  if (ctx->blobs) (void)hipFree(ctx->blobs);
  ctx->blobs = nullptr;
  ctx->prof.destroy();
  if (ctx->norms_ready) (void)hipEventDestroy(ctx->norms_ready);
  ctx->norms_ready = nullptr;
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  ctx->stream = nullptr;
}

// test entry points run level 0 with an unscaled right-hand side
int reset_scale(Ctx *ctx) {
  hipLaunchKernelGGL(k_fill<double>, dim3(1), dim3(256), 0, ctx->stream, ctx->scale, (size_t)2, 1.0);
  LAUNCHCHK();
  return MAC3D_OK;
}

int init_blobs_impl(Ctx *ctx, const mac3d_blob *blobs, int nblobs) {
  ctx->cvel_fresh = false;
  ctx->dmm_valid = false;
  ctx->hist = 0;
  ctx->iter_n = 0;
  if (nblobs < 0 || (nblobs > 0 && !blobs)) return fail(ctx, MAC3D_EINVAL, "bad blob list");
  for (int q = 0; q < nblobs; ++q) {
    const mac3d_blob &b = blobs[q];
    if (!std::isfinite(b.cx) || !std::isfinite(b.cy) || !std::isfinite(b.cz) ||
        !std::isfinite(b.w0) || !(b.radius > 0.5) || !std::isfinite(b.radius))
      return fail(ctx, MAC3D_EINVAL, "blob fields must be finite with radius > 0.5");
  }
  if (nblobs > ctx->blobs_cap) {
    if (ctx->blobs) HIPCHK(hipFree(ctx->blobs));
    ctx->blobs = nullptr;
    HIPCHK(hipMalloc(&ctx->blobs, sizeof(BlobDev) * nblobs));
    ctx->blobs_cap = nblobs;
  }
  std::vector<BlobDev> hb(nblobs);
  for (int q = 0; q < nblobs; ++q)
    hb[q] = BlobDev{blobs[q].cx, blobs[q].cy, blobs[q].cz, blobs[q].radius, blobs[q].w0};
  if (nblobs)
    HIPCHK(hipMemcpyAsync(ctx->blobs, hb.data(), sizeof(BlobDev) * nblobs, hipMemcpyHostToDevice,
                          ctx->stream));
  Buf *zero[] = {&ctx->u, &ctx->w, &ctx->p, &ctx->e0, &ctx->cvel};
  for (Buf *b : zero)
    if (b->ptr) HIPCHK(hipMemsetAsync(b->ptr, 0, type_size(b->type) * b->n, ctx->stream));
  const Dims &d = ctx->d[0];
  with_step_type(ctx->t0, [&](auto tag) {
    using T = std::remove_pointer_t<decltype(tag)>;
    hipLaunchKernelGGL(k_init_blobs<T>, cell_grid(d), kCellBlock, 0, ctx->stream, d,
                       (const BlobDev *)ctx->blobs, nblobs, (T *)ctx->dye.ptr, (T *)ctx->v.ptr);
  });
  LAUNCHCHK();
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return MAC3D_OK;
}


// ===========================================================================
// Adjoint of the step (no reference counterpart): reverse sweep with
// checkpoints every K steps, states recomputed in between.
enum { SV_U, SV_V, SV_W, SV_P, SV_DYE };

int adj_alloc(Ctx *ctx) {
  Ctx::Adj &A = ctx->adj;
  if (A.ready) return MAC3D_OK;
  int rc;
  size_t n0 = ctx->n[0], bytes = type_size(ctx->t0) * n0;
  if ((rc = dev_alloc(ctx, &A.c1, 3 * bytes))) return rc;
  void **arrs[] = {&A.vela[0], &A.vela[1], &A.vela[2],
                   &A.lvel[0], &A.lvel[1], &A.lvel[2], &A.lvelc[0], &A.lvelc[1], &A.lvelc[2],
                   &A.lvela[0], &A.lvela[1], &A.lvela[2], &A.lnew[0], &A.lnew[1], &A.lnew[2],
                   &A.lc[0], &A.lc[1], &A.lc[2], &A.ldye, &A.ldyen, &A.lpred, &A.target};
  for (void **a : arrs)
    if ((rc = dev_alloc(ctx, a, bytes))) return rc;
  int tx = ctx->mgx[0].type;
  if ((rc = alloc(ctx, A.lx, tx, n0))) return rc;
  if ((rc = alloc(ctx, A.lxa, tx, n0))) return rc;
  if ((rc = alloc(ctx, A.lb, tx, n0))) return rc;
  A.part_n = std::max(cell_blocks(ctx->d[0]), (size_t)1024 * 5 * 16);
  if ((rc = dev_alloc(ctx, (void **)&A.part, A.part_n * sizeof(double)))) return rc;
  if ((rc = dev_alloc(ctx, (void **)&A.acc, 4 * sizeof(double)))) return rc;
  A.ready = true;
  return MAC3D_OK;
}

size_t state_bytes(const Ctx *ctx, int f) {
  return type_size(f == SV_P ? ctx->p.type : ctx->t0) * ctx->n[0];
}

int states_alloc(Ctx *ctx, std::vector<void *> &v, int &cap, int count) {
  int rc;
  while (cap < count) {
    for (int f = 0; f < 5; ++f) {
      void *p = nullptr;
      if ((rc = dev_alloc(ctx, &p, state_bytes(ctx, f)))) return rc;
      v.push_back(p);
    }
    ++cap;
  }
  return MAC3D_OK;
}

void *state_field(Ctx *ctx, int f) {
  switch (f) {
    case SV_U: return ctx->u.ptr;
    case SV_V: return ctx->v.ptr;
    case SV_W: return ctx->w.ptr;
    case SV_P: return ctx->p.ptr;
    default: return ctx->dye.ptr;
  }
}

int save_state(Ctx *ctx, void *const *dst) {
  for (int f = 0; f < 5; ++f)
    HIPCHK(hipMemcpyAsync(dst[f], state_field(ctx, f), state_bytes(ctx, f),
                          hipMemcpyDeviceToDevice, ctx->stream));
  return MAC3D_OK;
}

int load_state(Ctx *ctx, void *const *src) {
  ctx->cvel_fresh = false;
  ctx->dmm_valid = false;
  ctx->hist = 0;
  ctx->iter_n = 0;
  for (int f = 0; f < 5; ++f)
    HIPCHK(hipMemcpyAsync(state_field(ctx, f), src[f], state_bytes(ctx, f),
                          hipMemcpyDeviceToDevice, ctx->stream));
  return MAC3D_OK;
}

int zero_dev(Ctx *ctx, void *p, size_t bytes) {
  HIPCHK(hipMemsetAsync(p, 0, bytes, ctx->stream));
  return MAC3D_OK;
}

// lam_rhs = L^-1 lam_p with the step's V-cycle (L is symmetric): cycles
// V-cycles from zero on the adjoint buffers; returns the solution buffer.
int adj_pressure_solve(Ctx *ctx, Buf *sol) {
  Ctx::Adj &A = ctx->adj;
  Buf sx = ctx->mgx[0], sxa = ctx->mgxa[0], sb = ctx->mgb[0];
  ctx->adj_solving = true;
  ctx->mgx[0] = A.lx;
  ctx->mgxa[0] = A.lxa;
  ctx->mgb[0] = A.lb;
  int rc = reset_scale(ctx);
  if (!rc) rc = mg_fill_zero(ctx, ctx->mgx[0]);
  for (int c = 0; c < std::max(ctx->cfg.cycles, 1) && !rc; ++c) rc = vcycle(ctx, false);
  *sol = ctx->mgx[0];
  // keep the pair's identities for the next solve
  A.lx = ctx->mgx[0];
  A.lxa = ctx->mgxa[0];
  ctx->mgx[0] = sx;
  ctx->mgxa[0] = sxa;
  ctx->mgb[0] = sb;
  ctx->adj_solving = false;
  return rc;
}

// One reverse step.  The context holds the step's input state X_n (a copy of
// src); lvel/ldye hold the adjoint of X_{n+1} on entry and of X_n on exit.
int adj_step(Ctx *ctx, double dt, void *const *src) {
  Ctx::Adj &A = ctx->adj;
  const Dims &d = ctx->d[0];
  size_t n0 = ctx->n[0], tb = type_size(ctx->t0) * n0;
  int rc, cyc = 0, it = 0, conv = 0;
  double rel = 0;
  // forward recompute of the stage inputs
  if ((rc = load_state(ctx, src))) return rc;
  if ((rc = launch_face_to_centre(ctx))) return rc;
  HIPCHK(hipMemcpyAsync(A.c1, ctx->cvel.ptr, 3 * tb, hipMemcpyDeviceToDevice, ctx->stream));
  if ((rc = launch_advect(ctx, dt))) return rc;
  void *va[3] = {ctx->u.ptr, ctx->v.ptr, ctx->w.ptr};
  for (int a = 0; a < 3; ++a)
    HIPCHK(hipMemcpyAsync(A.vela[a], va[a], tb, hipMemcpyDeviceToDevice, ctx->stream));
  if ((rc = launch_casson(ctx, dt))) return rc;
  if ((rc = launch_divergence(ctx))) return rc;
  if (ctx->prec == MAC3D_PREC_MIXED) {
    bool g = ctx->use_graphs;
    ctx->use_graphs = false;
    // the refinement loop of one_step from the step's own p (no history)
    if ((rc = mixed_begin(ctx, 0))) return rc;
    for (;;) {
      HIPCHK(hipStreamSynchronize(ctx->stream));
      double mr = ctx->red_host[0], rr = ctx->red_host[1], bb = ctx->red_host[2];
      double r = bb > 0.0 ? std::sqrt(rr / bb) : 0.0;
      if (r <= ctx->cfg.ir_tol || mr == 0.0 || it >= ctx->cfg.ir_max_iters) break;
      if ((rc = mixed_iter(ctx))) return rc;
      ++it;
    }
    ctx->use_graphs = g;
  } else {
    for (int c = 0; c < ctx->cfg.cycles; ++c)
      if ((rc = vcycle(ctx, false))) return rc;
  }
  (void)cyc; (void)conv; (void)rel;
  if ((rc = launch_project(ctx))) return rc;
  if ((rc = launch_face_to_centre(ctx))) return rc;  // C2
  if ((rc = launch_dye_predict(ctx, dt))) return rc;
  double dom[3] = {d.nx * ctx->cfg.dx, d.ny * ctx->cfg.dx, d.nz * ctx->cfg.dx};
  StepParams P = step_params(ctx, dt);
  // reverse
  if ((rc = zero_dev(ctx, A.lpred, tb)) || (rc = zero_dev(ctx, A.ldyen, tb))) return rc;
  for (int a = 0; a < 3; ++a)
    if ((rc = zero_dev(ctx, A.lc[a], tb))) return rc;
  with_step_type(ctx->t0, [&](auto tag) {
    using T = std::remove_pointer_t<decltype(tag)>;
    T s0 = (T)(dt / dom[0]), s1 = (T)(dt / dom[1]), s2 = (T)(dt / dom[2]);
    const T *dyen = (const T *)src[SV_DYE];
    hipLaunchKernelGGL(k_adj_dye_resolve<T>, cell_grid(d), kCellBlock, 0, ctx->stream, d, s0, s1,
                       s2, (T)ctx->cfg.maccormack, (const typename cvec_of<T>::type *)ctx->cvel.ptr,
                       dyen, (const T *)ctx->dpred.ptr,
                       (const T *)A.ldye, (T *)A.ldyen, (T *)A.lpred, (T *)A.lc[0],
                       (T *)A.lc[1], (T *)A.lc[2]);
    hipLaunchKernelGGL(k_adj_dye_predict<T>, cell_grid(d), kCellBlock, 0, ctx->stream, d, s0, s1,
                       s2, (const typename cvec_of<T>::type *)ctx->cvel.ptr,
                       dyen, (const T *)A.lpred, (T *)A.ldyen, (T *)A.lc[0], (T *)A.lc[1],
                       (T *)A.lc[2]);
    hipLaunchKernelGGL(k_adj_face_to_centre<T>, cell_grid(d), kCellBlock, 0, ctx->stream, d,
                       (const T *)A.lc[0], (const T *)A.lc[1], (const T *)A.lc[2],
                       (T *)A.lvel[0], (T *)A.lvel[1], (T *)A.lvel[2]);
  });
  LAUNCHCHK();
  // lam_p = D(lam') then the adjoint pressure solve
  with_step_type(ctx->t0, [&](auto tag) {
    using T = std::remove_pointer_t<decltype(tag)>;
    with_step_type(A.lb.type, [&](auto rtag) {
      using R = std::remove_pointer_t<decltype(rtag)>;
      hipLaunchKernelGGL((k_adj_project_p<T, R>), cell_grid(d), kCellBlock, 0, ctx->stream, d,
                         (const T *)A.lvel[0], (const T *)A.lvel[1], (const T *)A.lvel[2],
                         (R *)A.lb.ptr);
    });
  });
  LAUNCHCHK();
  Buf lrhs;
  if ((rc = adj_pressure_solve(ctx, &lrhs))) return rc;
  with_step_type(ctx->t0, [&](auto tag) {
    using T = std::remove_pointer_t<decltype(tag)>;
    with_step_type(lrhs.type, [&](auto rtag) {
      using R = std::remove_pointer_t<decltype(rtag)>;
      hipLaunchKernelGGL((k_adj_div<T, R>), cell_grid(d), kCellBlock, 0, ctx->stream, d,
                         (const T *)A.lvel[0], (const T *)A.lvel[1], (const T *)A.lvel[2],
                         (const R *)lrhs.ptr, (T *)A.lvelc[0], (T *)A.lvelc[1], (T *)A.lvelc[2]);
    });
  });
  LAUNCHCHK();
  for (int a = 0; a < 3; ++a)
    if ((rc = zero_dev(ctx, A.lvela[a], tb)) || (rc = zero_dev(ctx, A.lnew[a], tb)) ||
        (rc = zero_dev(ctx, A.lc[a], tb)))
      return rc;
  with_step_type(ctx->t0, [&](auto tag) {
    using T = std::remove_pointer_t<decltype(tag)>;
    const T *dyen = (const T *)src[SV_DYE];
    hipLaunchKernelGGL(k_adj_casson<T>, cell_grid(d), kCellBlock, 0, ctx->stream, d, P,
                       (const T *)A.vela[0], (const T *)A.vela[1], (const T *)A.vela[2], dyen,
                       (const T *)A.lvelc[0], (const T *)A.lvelc[1], (const T *)A.lvelc[2],
                       (T *)A.lvela[0], (T *)A.lvela[1], (T *)A.lvela[2], (T *)A.ldyen, A.part);
    hipLaunchKernelGGL(k_sum_into, dim3(1), dim3(256), 0, ctx->stream, (const double *)A.part,
                       cell_blocks(d), A.acc);
    T s0 = (T)(dt / dom[0]), s1 = (T)(dt / dom[1]), s2 = (T)(dt / dom[2]);
    hipLaunchKernelGGL(k_adj_advect<T>, cell_grid(d), kCellBlock, 0, ctx->stream, d, s0, s1, s2,
                       (T)ctx->cfg.vel_damp, (const T *)src[SV_U], (const T *)src[SV_V],
                       (const T *)src[SV_W], (const typename cvec_of<T>::type *)A.c1,
                       (const T *)A.lvela[0], (const T *)A.lvela[1],
                       (const T *)A.lvela[2], (T *)A.lnew[0], (T *)A.lnew[1], (T *)A.lnew[2],
                       (T *)A.lc[0], (T *)A.lc[1], (T *)A.lc[2]);
    hipLaunchKernelGGL(k_adj_face_to_centre<T>, cell_grid(d), kCellBlock, 0, ctx->stream, d,
                       (const T *)A.lc[0], (const T *)A.lc[1], (const T *)A.lc[2],
                       (T *)A.lnew[0], (T *)A.lnew[1], (T *)A.lnew[2]);
  });
  LAUNCHCHK();
  for (int a = 0; a < 3; ++a) std::swap(A.lvel[a], A.lnew[a]);
  std::swap(A.ldye, A.ldyen);
  return MAC3D_OK;
}

int adjoint_impl(Ctx *ctx, const mac3d_blob *blobs, int nb, double dt, int nsteps, int K,
                 const void *target, int dtype, mac3d_grad *out) {
  int rc;
  Ctx::Adj &A = ctx->adj;
  if ((rc = adj_alloc(ctx))) return rc;
  int nck = (nsteps + K - 1) / K;
  if ((rc = states_alloc(ctx, A.ckpt, A.ckpt_cap, std::max(nck, 1)))) return rc;
  if ((rc = states_alloc(ctx, A.seg, A.seg_cap, K))) return rc;
  size_t n0 = ctx->n[0], tb = type_size(ctx->t0) * n0;
  // forward with checkpoints
  if ((rc = init_blobs_impl(ctx, blobs, nb))) return rc;
  if ((rc = save_state(ctx, &A.ckpt[0]))) return rc;
  int cyc = 0, it = 0, conv = 0;
  double rel = 0;
  for (int n = 0; n < nsteps; ++n) {
    if ((rc = one_step(ctx, dt, &cyc, &it, &rel, &conv))) return rc;
    if ((n + 1) % K == 0 && n + 1 < nsteps)
      if ((rc = save_state(ctx, &A.ckpt[5 * ((n + 1) / K)]))) return rc;
  }
  // loss and seed: lam_dye = dye_N - target
  HIPCHK(hipMemcpyAsync(ctx->stage, target, (dtype == MAC3D_F32 ? 4 : 8) * n0,
                        hipMemcpyHostToDevice, ctx->stream));
  dim3 g1((unsigned)((n0 + 255) / 256));
  with_step_type(ctx->t0, [&](auto tag) {
    using T = std::remove_pointer_t<decltype(tag)>;
    if (dtype == MAC3D_F32)
      hipLaunchKernelGGL((k_convert<float, T>), g1, dim3(256), 0, ctx->stream,
                         (const float *)ctx->stage, (T *)A.target, n0);
    else
      hipLaunchKernelGGL((k_convert<double, T>), g1, dim3(256), 0, ctx->stream,
                         (const double *)ctx->stage, (T *)A.target, n0);
  });
  LAUNCHCHK();
  double *lossp = nullptr;
  if ((rc = dev_alloc(ctx, (void **)&lossp, (g1.x + 4) * sizeof(double)))) return rc;
  with_step_type(ctx->t0, [&](auto tag) {
    using T = std::remove_pointer_t<decltype(tag)>;
    hipLaunchKernelGGL(k_sub_sq<T>, g1, dim3(256), 0, ctx->stream, n0, (const T *)ctx->dye.ptr,
                       (const T *)A.target, (T *)A.ldye, lossp);
  });
  LAUNCHCHK();
  std::vector<double> lp(g1.x);
  HIPCHK(hipMemcpyAsync(lp.data(), lossp, g1.x * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  for (int a = 0; a < 3; ++a)
    if ((rc = zero_dev(ctx, A.lvel[a], tb))) return rc;
  HIPCHK(hipMemsetAsync(A.acc, 0, 4 * sizeof(double), ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  double loss = 0.0;
  for (double x : lp) loss += x;
  loss *= 0.5;
  // reverse sweep over the segments
  for (int sgi = nck - 1; sgi >= 0; --sgi) {
    int n0s = sgi * K, len = std::min(K, nsteps - n0s);
    if ((rc = load_state(ctx, &A.ckpt[5 * sgi]))) return rc;
    for (int sgs = 0; sgs < len; ++sgs) {
      if ((rc = save_state(ctx, &A.seg[5 * sgs]))) return rc;
      if (sgs + 1 < len)
        if ((rc = one_step(ctx, dt, &cyc, &it, &rel, &conv))) return rc;
    }
    for (int sgs = len - 1; sgs >= 0; --sgs)
      if ((rc = adj_step(ctx, dt, &A.seg[5 * sgs]))) return rc;
  }
  // parameters of the initial condition
  std::vector<BlobDev> hb(nb);
  for (int q = 0; q < nb; ++q)
    hb[q] = BlobDev{blobs[q].cx, blobs[q].cy, blobs[q].cz, blobs[q].radius, blobs[q].w0};
  HIPCHK(hipMemcpyAsync(ctx->blobs, hb.data(), sizeof(BlobDev) * nb, hipMemcpyHostToDevice,
                        ctx->stream));
  const int nblk = 1024, npar = 5 * nb;
  with_step_type(ctx->t0, [&](auto tag) {
    using T = std::remove_pointer_t<decltype(tag)>;
    hipLaunchKernelGGL(k_adj_init<T>, dim3(nblk), dim3(256), 4 * npar * sizeof(double),
                       ctx->stream, ctx->d[0], (const BlobDev *)ctx->blobs, nb,
                       (const T *)A.ldye, (const T *)A.lvel[1], A.part);
  });
  LAUNCHCHK();
  std::vector<double> hp((size_t)nblk * npar);
  double acc[4];
  HIPCHK(hipMemcpyAsync(hp.data(), A.part, hp.size() * sizeof(double), hipMemcpyDeviceToHost,
                        ctx->stream));
  HIPCHK(hipMemcpyAsync(acc, A.acc, 4 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  std::memset(out, 0, sizeof(*out));
  out->loss = loss;
  out->d_rho_b = acc[0];
  out->nblobs = nb;
  out->steps = nsteps;
  for (int q = 0; q < npar; ++q) {
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += hp[(size_t)b * npar + q];
    out->d_blob[q / 5][q % 5] = s;
  }
  HIPCHK(hipFree(lossp));
  ctx->allocs.erase(std::find(ctx->allocs.begin(), ctx->allocs.end(), (void *)lossp));
  return MAC3D_OK;
}

thread_local std::string g_create_err;  // last error of a call without a context

// ---- scratch of mac3d_div_norms_noalloc: the per-block (max, sum of squares)
// pairs of its divergence pass, 131072 pairs (2 MiB) per GPU, held until the
// process ends.  The host refers to it only through the address
// hipGetSymbolAddress returns for the current device (kept per device).
constexpr size_t kNoAllocMaxParts = 131072;
static __device__ __align__(256) double scratch[2 * kNoAllocMaxParts];
constexpr int kNoAllocMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_dev[kNoAllocMaxDevices] = {};

int noalloc_fail(int code, const std::string &msg) {
  g_create_err = msg;
  return code;
}

// ---- checkpoints: the state a step evolves, in one caller-owned byte buffer
// (a header, then the sections in a fixed order; layout in mac3d.h)
struct CkptHeader {
  char magic[8];
  int32_t abi, nx, ny, nz, levels, precision, elem;
  int32_t wave_n, np, hist, iter_n;
  int32_t iter_log[4];
  int32_t conc, reserved;
  uint64_t bytes;
};
static_assert(sizeof(CkptHeader) == 88, "checkpoint header layout");
const char kCkptMagic[8] = {'M', 'A', 'C', '3', 'D', 'C', 'K', '1'};

struct CkptSec {
  void *dev;
  size_t bytes;
};
// the device ranges of a checkpoint's sections, in order
// (conc: the concentration grid's refinement, 0 without it)
std::vector<CkptSec> ckpt_sections(Ctx *ctx, int wave_n, int np, int conc) {
  const size_t n0 = ctx->n[0], es = type_size(ctx->t0);
  std::vector<CkptSec> v = {{ctx->u.ptr, es * n0}, {ctx->v.ptr, es * n0}, {ctx->w.ptr, es * n0},
                            {ctx->dye.ptr, es * n0}};
  if (ctx->prec == MAC3D_PREC_MIXED)  // the current pressure and the two before it, FP64
    for (int q = 0; q < 3; ++q) v.push_back({ctx->pr[ctx->role[q]].ptr, 8 * n0});
  else
    v.push_back({ctx->p.ptr, es * n0});
  if (wave_n > 0) v.push_back({ctx->wave.ptr, 4 * sizeof(float) * (size_t)wave_n * wave_n});
  if (np > 0) v.push_back({(void *)ctx->parcels, (size_t)np * sizeof(Parcel)});
  if (conc) {
    const size_t nc = (size_t)conc * conc * conc * n0;
    v.push_back({ctx->conc.ptr, 4 * sizeof(float) * nc});
  }
  return v;
}
size_t ckpt_bytes(Ctx *ctx, int wave_n, int np, int conc) {
  size_t b = sizeof(CkptHeader);
  for (const CkptSec &c : ckpt_sections(ctx, wave_n, np, conc)) b += c.bytes;
  return b;
}
int ckpt_wave_n(const Ctx *ctx) { return ctx->wave_on ? ctx->wp.n : 0; }
int ckpt_conc(const Ctx *ctx) { return ctx->conc_on ? ctx->conc_r : 0; }

}  // namespace

// ===========================================================================
// C ABI
extern "C" {

int mac3d_default_config(mac3d_config *c) {
  if (!c) return MAC3D_EINVAL;
  std::memset(c, 0, sizeof(*c));
  c->nx = 32; c->ny = 48; c->nz = 32;
  c->levels = 3; c->nu1 = 2; c->nu2 = 2; c->cycles = 8;
  c->precision = MAC3D_PREC_FP32;
  c->ir_max_iters = 30;
  c->device = 0;
  c->dx = 1e-3;
  c->rho_w = 1000.0; c->mu_w = 1e-3;
  c->rho_b = 1060.0; c->tau_y = 0.005; c->mu_inf = 0.0035;
  c->g = 9.81;
  c->gamma_reg = 0.40;       // fluid.ts:461
  c->shear_eps2 = 1e-6;      // fluid.ts:454
  c->alpha_max = 1.0 / 6.0;  // FTCS stability bound of the 7-point update
  c->vel_damp = 1.0;
  c->maccormack = 0.45 * 0.995;  // advect.ts:202 with concSharpness 0.995
  c->ir_tol = 1e-8;
  c->extrapolate = -1;
  c->sor_omega = 0.0;
  return MAC3D_OK;
}

int mac3d_create(const mac3d_config *cfg, void **out) {
  if (!cfg || !out) return MAC3D_EINVAL;
  *out = nullptr;
  std::string why;
  if (!valid_config(cfg, why)) {
    g_create_err = why;
    return MAC3D_EINVAL;
  }
  Ctx *ctx = new (std::nothrow) Ctx();
  if (!ctx) return MAC3D_ENOMEM;
  int rc = create_impl(ctx, cfg);
  if (rc) {
    g_create_err = ctx->err;
    destroy_impl(ctx);
    delete ctx;
    return rc;
  }
  *out = ctx;
  return MAC3D_OK;
}

int mac3d_destroy(void *h) {
  if (!h) return MAC3D_OK;
  Ctx *ctx = (Ctx *)h;
  destroy_impl(ctx);
  delete ctx;
  return MAC3D_OK;
}

int mac3d_upload(void *h, int field, int level, const void *host, int dtype, size_t count) {
  if (!h) return MAC3D_EINVAL;
  Ctx *ctx = (Ctx *)h;
  ctx->cvel_fresh = false;
  ctx->dmm_valid = false;
  Buf *b = field_buf(ctx, field, level);
  if (!b || !host) return fail(ctx, MAC3D_EINVAL, "unknown field/level or null host pointer");
  if (dtype != MAC3D_F32 && dtype != MAC3D_F64) return fail(ctx, MAC3D_EINVAL, "unknown dtype");
  int comp = (field >= MAC3D_FIELD_CU && field <= MAC3D_FIELD_CW) ? field - MAC3D_FIELD_CU : -1;
  if (count != (comp >= 0 ? ctx->n[0] : b->n))
    return fail(ctx, MAC3D_EINVAL, "element count does not match the field");
  if (field == MAC3D_FIELD_P) {  // an uploaded pressure restarts the history
    ctx->hist = 0;
    ctx->iter_n = 0;
  }
  size_t hsz = dtype == MAC3D_F32 ? 4 : 8;
  HIPCHK(hipMemcpyAsync(ctx->stage, host, hsz * count, hipMemcpyHostToDevice, ctx->stream));
  if (field == MAC3D_FIELD_CONC) {  // 4 values per cell -> the four planes
    const size_t n0 = b->n / 4;
    dim3 g((unsigned)((n0 + 255) / 256));
    if (dtype == MAC3D_F32)
      hipLaunchKernelGGL((k_rgba_to_planes<float, float>), g, dim3(256), 0, ctx->stream,
                         (const float *)ctx->stage, (float *)b->ptr, n0);
    else
      hipLaunchKernelGGL((k_rgba_to_planes<double, float>), g, dim3(256), 0, ctx->stream,
                         (const double *)ctx->stage, (float *)b->ptr, n0);
    LAUNCHCHK();
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MAC3D_OK;
  }
  if (comp >= 0) {
    with_step_type(b->type, [&](auto tag) {
      using S = std::remove_pointer_t<decltype(tag)>;
      dim3 g((unsigned)((count + 255) / 256));
      if (dtype == MAC3D_F32)
        hipLaunchKernelGGL((k_convert_strided<float, S>), g, dim3(256), 0, ctx->stream,
                           (const float *)ctx->stage, (S *)b->ptr, count, 1, 0, 3, comp);
      else
        hipLaunchKernelGGL((k_convert_strided<double, S>), g, dim3(256), 0, ctx->stream,
                           (const double *)ctx->stage, (S *)b->ptr, count, 1, 0, 3, comp);
    });
    LAUNCHCHK();
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MAC3D_OK;
  }
  with_type(b->type, [&](auto tag) {
    using S = std::remove_pointer_t<decltype(tag)>;
    dim3 g((unsigned)((count + 255) / 256));
    if (dtype == MAC3D_F32)
      hipLaunchKernelGGL((k_convert<float, S>), g, dim3(256), 0, ctx->stream,
                         (const float *)ctx->stage, (S *)b->ptr, count);
    else
      hipLaunchKernelGGL((k_convert<double, S>), g, dim3(256), 0, ctx->stream,
                         (const double *)ctx->stage, (S *)b->ptr, count);
  });
  LAUNCHCHK();
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return MAC3D_OK;
}

int mac3d_download(void *h, int field, int level, void *host, int dtype, size_t count) {
  if (!h) return MAC3D_EINVAL;
  Ctx *ctx = (Ctx *)h;
  Buf *b = field_buf(ctx, field, level);
  if (!b || !host) return fail(ctx, MAC3D_EINVAL, "unknown field/level or null host pointer");
  if (dtype != MAC3D_F32 && dtype != MAC3D_F64) return fail(ctx, MAC3D_EINVAL, "unknown dtype");
  int comp = (field >= MAC3D_FIELD_CU && field <= MAC3D_FIELD_CW) ? field - MAC3D_FIELD_CU : -1;
  if (count != (comp >= 0 ? ctx->n[0] : b->n))
    return fail(ctx, MAC3D_EINVAL, "element count does not match the field");
  size_t hsz = dtype == MAC3D_F32 ? 4 : 8;
  if (field == MAC3D_FIELD_CONC) {  // the four planes -> 4 values per cell
    const size_t n0 = b->n / 4;
    dim3 g((unsigned)((n0 + 255) / 256));
    if (dtype == MAC3D_F32)
      hipLaunchKernelGGL((k_planes_to_rgba<float, float>), g, dim3(256), 0, ctx->stream,
                         (const float *)b->ptr, (float *)ctx->stage, n0);
    else
      hipLaunchKernelGGL((k_planes_to_rgba<float, double>), g, dim3(256), 0, ctx->stream,
                         (const float *)b->ptr, (double *)ctx->stage, n0);
  } else if (comp >= 0) {
    with_step_type(b->type, [&](auto tag) {
      using S = std::remove_pointer_t<decltype(tag)>;
      dim3 g((unsigned)((count + 255) / 256));
      if (dtype == MAC3D_F32)
        hipLaunchKernelGGL((k_convert_strided<S, float>), g, dim3(256), 0, ctx->stream,
                           (const S *)b->ptr, (float *)ctx->stage, count, 3, comp, 1, 0);
      else
        hipLaunchKernelGGL((k_convert_strided<S, double>), g, dim3(256), 0, ctx->stream,
                           (const S *)b->ptr, (double *)ctx->stage, count, 3, comp, 1, 0);
    });
  } else
  with_type(b->type, [&](auto tag) {
    using S = std::remove_pointer_t<decltype(tag)>;
    dim3 g((unsigned)((count + 255) / 256));
    if (dtype == MAC3D_F32)
      hipLaunchKernelGGL((k_convert<S, float>), g, dim3(256), 0, ctx->stream,
                         (const S *)b->ptr, (float *)ctx->stage, count);
    else
      hipLaunchKernelGGL((k_convert<S, double>), g, dim3(256), 0, ctx->stream,
                         (const S *)b->ptr, (double *)ctx->stage, count);
  });
  LAUNCHCHK();
  HIPCHK(hipMemcpyAsync(host, ctx->stage, hsz * count, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return MAC3D_OK;
}

int mac3d_wave_default_params(mac3d_wave_params *p) {
  if (!p) return MAC3D_EINVAL;
  std::memset(p, 0, sizeof(*p));
  p->n = 512;               // WAVE_GRID_SIZE (common.ts:1)
  p->half_extent = 1.60;    // WAVE_HALF_EXTENT (common.ts:2)
  p->half_xz = 0.0;         // nx dx / 2
  p->depth = 0.0;           // ny dx
  p->surface_y = 0.30;      // WATER_SURFACE_Y
  p->marangoni = 0.0;
  p->vort_conf = 1.02;      // vorticityConfinement (uniformLayout.ts:32)
  p->sim_time = 0.0;
  p->schlieren = 0.032;     // schlierenStrength (uniformLayout.ts:44)
  p->hemolysis = 0.088;     // hemolysisRate (uniformLayout.ts:43)
  p->physics = 0;
  p->casson_scale = 1.0;    // cassonYield (uniformLayout.ts:42)
  p->buoyancy = 0.48;       // buoyancyStrength (uniformLayout.ts:33)
  return MAC3D_OK;
}

int mac3d_wave_enable(void *h, const mac3d_wave_params *p) {
  if (!h) return MAC3D_EINVAL;
  Ctx *ctx = (Ctx *)h;
  drop_graphs(ctx);  // the step's launch sequence changes
  if (!p) {
    ctx->wave_on = false;
    return MAC3D_OK;
  }
  const int n = p->n;
  if (n < 16 || n > 1024 || (n & (n - 1)))
    return fail(ctx, MAC3D_EINVAL, "wave n must be a power of two in [16, 1024]");
  const double ext = p->half_extent;
  const double hx = p->half_xz > 0.0 ? p->half_xz : 0.5 * ctx->cfg.nx * ctx->cfg.dx;
  const double depth = p->depth > 0.0 ? p->depth : ctx->cfg.ny * ctx->cfg.dx;
  if (!(ext > 0.0) || !std::isfinite(ext) || !std::isfinite(hx) || !std::isfinite(depth) ||
      !std::isfinite(p->surface_y) || !std::isfinite(p->marangoni) || p->half_xz < 0.0 ||
      p->depth < 0.0 || !std::isfinite(p->vort_conf) || !std::isfinite(p->sim_time) ||
      !std::isfinite(p->schlieren) || !std::isfinite(p->hemolysis) ||
      !std::isfinite(p->casson_scale) || !std::isfinite(p->buoyancy))
    return fail(ctx, MAC3D_EINVAL, "wave lengths must be finite, half_extent > 0");
  if (p->physics != 0 && p->physics != 1)
    return fail(ctx, MAC3D_EINVAL, "physics must be 0 (the workload's) or 1 (the reference's)");
  int rc;
  const size_t nn = (size_t)n * n;
  if (!ctx->wave.ptr || ctx->wp.n != n) {  // (earlier buffers stay owned until destroy)
    if ((rc = alloc(ctx, ctx->wave, TF32, 4 * nn))) return rc;
    if ((rc = alloc(ctx, ctx->wz, TF32, 2 * (size_t)n * wv_pitch(n)))) return rc;
    if ((rc = alloc(ctx, ctx->wfilm, TF32, nn))) return rc;
    if ((rc = dev_alloc(ctx, (void **)&ctx->wtw, (size_t)(n / 2) * sizeof(float2)))) return rc;
    std::vector<float2> tw(n / 2);
    for (int m = 0; m < n / 2; ++m) {
      const double a = 2.0 * M_PI * m / n;
      tw[m] = make_float2((float)std::cos(a), (float)-std::sin(a));
    }
    HIPCHK(hipMemcpyAsync(ctx->wtw, tw.data(), tw.size() * sizeof(float2), hipMemcpyHostToDevice,
                          ctx->stream));
    if (4 * nn > ctx->stage_n) {
      if ((rc = dev_alloc(ctx, (void **)&ctx->stage, 4 * nn * sizeof(double)))) return rc;
      ctx->stage_n = 4 * nn;
    }
  }
  if (!ctx->ptop.ptr && (rc = alloc(ctx, ctx->ptop, TF32, (size_t)ctx->cfg.nx * ctx->cfg.nz)))
    return rc;
  if (!ctx->drops && (rc = dev_alloc(ctx, (void **)&ctx->drops, kDropBytes))) return rc;
  HIPCHK(hipMemsetAsync(ctx->drops, 0, WV_MAX_DROPS * sizeof(WaveDrop), ctx->stream));
  if (p->scene && !ctx->curl.ptr) {
    const size_t n4 = 4 * ctx->n[0];
    if ((rc = alloc(ctx, ctx->curl, ctx->t0, n4))) return rc;  // (zeroed)
    if (n4 > ctx->stage_n) {
      if ((rc = dev_alloc(ctx, (void **)&ctx->stage, n4 * sizeof(double)))) return rc;
      ctx->stage_n = n4;
    }
  }
  const SceneClock clk{(float)p->sim_time, (float)p->schlieren, (float)p->vort_conf,
                       (float)p->hemolysis};
  ctx->clk_host = clk;
  HIPCHK(hipMemcpyAsync((void *)scene_clock(ctx), &clk, sizeof(clk), hipMemcpyHostToDevice,
                        ctx->stream));
  // a flat, film-free surface (WAVE_CLEAR, wave.ts:15-30), no drops and no lid
  // pressure
  HIPCHK(hipMemsetAsync(ctx->wave.ptr, 0, 4 * nn * sizeof(float), ctx->stream));
  HIPCHK(hipMemsetAsync(ctx->ptop.ptr, 0, ctx->ptop.n * sizeof(float), ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->wp.n = n;
  ctx->wp.ext = (float)ext;
  ctx->wp.half_xz = (float)hx;
  ctx->wp.depth = (float)depth;
  ctx->wp.surface_y = (float)p->surface_y;
  ctx->wp.marangoni = (float)p->marangoni;
  ctx->wp.dt = 0.f;
  ctx->wave_scene = p->scene != 0;
  ctx->ref_physics = p->physics == 1;
  ctx->casson_scale = p->casson_scale;
  ctx->buoyancy = p->buoyancy;
  ctx->wave_on = true;
  return MAC3D_OK;
}

int mac3d_set_drops(void *h, const mac3d_drop *drops, int ndrops) {
  if (!h) return MAC3D_EINVAL;
  Ctx *ctx = (Ctx *)h;
  if (ndrops < 0 || ndrops > MAC3D_MAX_DROPS || (ndrops > 0 && !drops))
    return fail(ctx, MAC3D_EINVAL, "need 0..MAC3D_MAX_DROPS drops");
  if (!ctx->wave_on) return fail(ctx, MAC3D_ESTATE, "the wave coupling is not enabled");
  WaveDrop hd[WV_MAX_DROPS];
  DropPhase hp[WV_MAX_DROPS];
  std::memset(hd, 0, sizeof(hd));
  std::memset(hp, 0, sizeof(hp));
  for (int q = 0; q < ndrops; ++q) {
    const mac3d_drop &a = drops[q];
    const double v[8] = {a.x, a.y, a.z, a.radius, a.vel_y, a.osc, a.drain, a.seed};
    for (double x : v)
      if (!std::isfinite(x)) return fail(ctx, MAC3D_EINVAL, "drop values must be finite");
    hd[q] = WaveDrop{(float)a.x, (float)a.y, (float)a.z, (float)a.radius,
                     (float)a.vel_y, (float)a.osc, (float)a.drain, (float)a.seed};
    const double sd = hd[q].seed, dr = hd[q].drain;
    hp[q] = DropPhase{(float)std::cos(1.73 * sd), (float)std::sin(1.73 * sd),
                      (float)std::cos(-2.31 * sd), (float)std::sin(-2.31 * sd),
                      (float)std::cos(0.89 * sd), (float)std::sin(0.89 * sd),
                      (float)std::min(std::max(std::pow(std::max(dr, 0.0), 0.66) * 0.84, 0.0), 1.24),
                      0.f};
  }
  // (the captured step sequences read the slots from device memory)
  HIPCHK(hipMemcpyAsync(ctx->drops, hd, sizeof(hd), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync((void *)drop_phases(ctx), hp, sizeof(hp), hipMemcpyHostToDevice,
                        ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return MAC3D_OK;
}

int mac3d_set_uniforms(void *h, const float *u, size_t count) {
  if (!h) return MAC3D_EINVAL;
  Ctx *ctx = (Ctx *)h;
  if (!u || count != 128) return fail(ctx, MAC3D_EINVAL, "need the 128-float SimUniforms block");
  if (!ctx->wave_on) return fail(ctx, MAC3D_ESTATE, "the wave coupling is not enabled");
  for (size_t q = 35; q < 128; ++q)
    if (!std::isfinite(u[q])) return fail(ctx, MAC3D_EINVAL, "non-finite uniform value");
  const float half_xz = u[52], depth = u[54] - u[53];
  if (!(half_xz > 0.f) || !(depth > 0.f))
    return fail(ctx, MAC3D_EINVAL, "domainBounds must give half_xz > 0 and yMax > yMin");
  // the coupling's tank top is the still surface (the reference scene packs
  // surfaceParams.x = domainBounds.z, uniformLayout.ts:97,150-153); a block
  // with the two apart would shift the world-to-grid mapping of the scene
  // terms and the parcels
  if (u[40] != u[54])
    return fail(ctx, MAC3D_EINVAL, "surfaceParams.x (surface height) must equal domainBounds.z "
                                   "(tank top)");
  mac3d_drop d[MAC3D_MAX_DROPS];
  for (int q = 0; q < MAC3D_MAX_DROPS; ++q) {
    const float *b = u + 64 + 8 * q;
    d[q] = mac3d_drop{b[0], b[1], b[2], b[3], b[4], b[5], b[6], b[7]};
  }
  int rc = mac3d_set_drops(h, d, MAC3D_MAX_DROPS);
  if (rc) return rc;
  // simTime, schlierenStr, vorticityConf (device memory: no graph rebuild)
  const SceneClock clk{u[35], u[60], u[44], u[59]};
  ctx->clk_host = clk;
  HIPCHK(hipMemcpyAsync((void *)scene_clock(ctx), &clk, sizeof(clk), hipMemcpyHostToDevice,
                        ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  bool changed = ctx->wp.surface_y != u[40] || ctx->wp.marangoni != u[41] ||
                 ctx->wp.half_xz != half_xz || ctx->wp.depth != depth;
  if (ctx->ref_physics) {
    // the reference physics' own parameters: buoyancyStrength [45],
    // velocityDamping [46], concSharpness [47] (MacCormack 0.45 sharpness,
    // advect.ts:197), cassonYield [58]
    const double vd = u[46], mc = 0.45 * (double)u[47];
    changed = changed || ctx->buoyancy != u[45] || ctx->casson_scale != u[58] ||
              ctx->cfg.vel_damp != vd || ctx->cfg.maccormack != mc;
    ctx->buoyancy = u[45];
    ctx->casson_scale = u[58];
    ctx->cfg.vel_damp = vd;
    ctx->cfg.maccormack = mc;
  }
  if (changed) drop_graphs(ctx);  // (the captured launches hold the parameters by value)
  // the parcels' emission of the next parcel stage: spawnCursor [61],
  // spawnCount [62] (particleParams.y/.z, particles.ts:65-66,95)
  if (ctx->np > 0) {
    const int cursor = (int)std::max(0.f, u[61]), count = (int)std::max(0.f, u[62]);
    if (count > kMaxEmit) return fail(ctx, MAC3D_EINVAL, "spawnCount above 8192");
    int rc2 = mac3d_parcels_emit(h, cursor, count);
    if (rc2) return rc2;
  }
  ctx->wp.surface_y = u[40];
  ctx->wp.marangoni = u[41];
  ctx->wp.half_xz = half_xz;
  ctx->wp.depth = depth;
  return MAC3D_OK;
}

int mac3d_parcels_upload(void *h, const float *parcels, int32_t n) {
  if (!h) return MAC3D_EINVAL;
  Ctx *ctx = (Ctx *)h;
  if (n < 0 || (n > 0 && !parcels)) return fail(ctx, MAC3D_EINVAL, "need n >= 0 parcels");
  if (!ctx->wave_on) return fail(ctx, MAC3D_ESTATE, "the wave coupling is not enabled");
  // the splat counters hold 65536 per unit mass in 32 bits: a cell overflows
  // at 65536 units, so the parcels' total mass (which a step never raises)
  // must stay below that
  double total = 0.0;
  for (int32_t q = 0; q < n; ++q) {
    const float m = parcels[8 * q + 3];
    if (!std::isfinite(m) || m < 0.f)
      return fail(ctx, MAC3D_EINVAL, "parcel masses must be finite and >= 0");
    total += m;
  }
  if (total >= 65536.0)
    return fail(ctx, MAC3D_EINVAL, "total parcel mass must stay below 65536 (32-bit splat "
                                   "counters at 65536 per unit mass)");
  int rc;
  if (n > ctx->np_cap) {  // (an earlier buffer stays owned until destroy)
    if ((rc = dev_alloc(ctx, (void **)&ctx->parcels, (size_t)n * sizeof(Parcel)))) return rc;
    ctx->np_cap = n;
  }
  if (!ctx->splat || ctx->splat_cap < splat_cells(ctx)) {
    drop_graphs(ctx);  // (a captured parcel stage holds the old counters)
    if ((rc = ensure_splat(ctx))) return rc;
  }
  if (n)
    HIPCHK(hipMemcpyAsync(ctx->parcels, parcels, (size_t)n * sizeof(Parcel), hipMemcpyHostToDevice,
                          ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->np = n;
  return MAC3D_OK;
}

int mac3d_conc_enable(void *h, int enable) {
  if (!h) return MAC3D_EINVAL;
  Ctx *ctx = (Ctx *)h;
  if (enable < 0 || enable > 2) return fail(ctx, MAC3D_EINVAL, "enable is 0, 1 or 2");
  drop_graphs(ctx);  // the step's launch sequence changes
  if (!enable) {
    ctx->conc_on = false;
    return MAC3D_OK;
  }
  if (!ctx->wave_on) return fail(ctx, MAC3D_ESTATE, "the wave coupling is not enabled");
  if (ctx->t0 != TF32)
    return fail(ctx, MAC3D_EINVAL, "the 4-channel concentration needs the FP32 or mixed step");
  int rc;
  const Dims &d = ctx->d[0];
  const Dims dc{d.nx * enable, d.ny * enable, d.nz * enable};
  const size_t nc = dc.cells();
  // the kernels index the concentration grid and its sampled planes with
  // 32-bit cell coordinates and size_t offsets; the reference's is 9.4 M cells
  if (dc.nz > 65535 || nc > ((size_t)1 << 32))
    return fail(ctx, MAC3D_EINVAL, "concentration grid too large");
  if (ctx->conc.ptr && ctx->conc_r != enable) {  // another grid: start from zero
    Buf *bs[] = {&ctx->conc, &ctx->conc2, &ctx->cpred, &ctx->cmac};
    for (Buf *b : bs) {
      dev_free(ctx, b->ptr);
      *b = Buf{};
    }
  }
  if (!ctx->conc.ptr) {
    if ((rc = alloc(ctx, ctx->conc, TF32, 4 * nc))) return rc;  // (zeroed)
    if ((rc = alloc(ctx, ctx->conc2, TF32, 4 * nc))) return rc;
    if ((rc = alloc(ctx, ctx->cpred, TF32, nc))) return rc;
    if (enable != 1 && (rc = alloc(ctx, ctx->cmac, TF32, 4 * d.cells()))) return rc;
    ctx->conc_home = ctx->conc.ptr;
  }
  ctx->conc_r = enable;
  ctx->dcg = dc;
  ctx->conc_on = true;
  if ((rc = ensure_splat(ctx))) return rc;
  if (4 * nc > ctx->stage_n) {
    dev_free(ctx, ctx->stage);
    ctx->stage = nullptr;
    ctx->stage_n = 0;
    if ((rc = dev_alloc(ctx, (void **)&ctx->stage, 4 * nc * sizeof(double)))) return rc;
    ctx->stage_n = 4 * nc;
  }
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return MAC3D_OK;
}

int mac3d_parcels_emit(void *h, int32_t cursor, int32_t count) {
  if (!h) return MAC3D_EINVAL;
  Ctx *ctx = (Ctx *)h;
  if (!ctx->wave_on) return fail(ctx, MAC3D_ESTATE, "the wave coupling is not enabled");
  if (cursor < 0 || count < 0 || count > kMaxEmit)
    return fail(ctx, MAC3D_EINVAL, "need cursor >= 0 and 0 <= count <= 8192");
  if (count > 0 && ctx->np == 0)
    return fail(ctx, MAC3D_ESTATE, "upload the parcel ring (mac3d_parcels_upload) first");
  int rc;
  if (!ctx->emit) {
    if ((rc = dev_alloc(ctx, (void **)&ctx->emit, sizeof(EmitCtl)))) return rc;
    drop_graphs(ctx);  // (the parcel stage now runs the emission kernel)
  }
  const EmitCtl e{ctx->np > 0 ? cursor % ctx->np : 0, count};
  HIPCHK(hipMemcpyAsync(ctx->emit, &e, sizeof(e), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return MAC3D_OK;
}

int mac3d_parcels_download(void *h, float *parcels, int32_t n) {
  if (!h) return MAC3D_EINVAL;
  Ctx *ctx = (Ctx *)h;
  if (n < 0 || n > ctx->np || (n > 0 && !parcels))
    return fail(ctx, MAC3D_EINVAL, "n must be within the uploaded parcels");
  if (n)
    HIPCHK(hipMemcpyAsync(parcels, ctx->parcels, (size_t)n * sizeof(Parcel), hipMemcpyDeviceToHost,
                          ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return MAC3D_OK;
}

int mac3d_splat_download(void *h, int which, uint32_t *out, size_t count) {
  if (!h) return MAC3D_EINVAL;
  Ctx *ctx = (Ctx *)h;
  if ((which != 0 && which != 1) || !out || count != splat_cells(ctx))
    return fail(ctx, MAC3D_EINVAL, "which is 0 or 1 and count the splat grid's cells");
  if (!ctx->splat || ctx->splat_cap < count) return fail(ctx, MAC3D_ESTATE, "no parcels uploaded");
  HIPCHK(hipMemcpyAsync(out, ctx->splat + which * count, count * sizeof(unsigned),
                        hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return MAC3D_OK;
}

int mac3d_init_blobs(void *h, const mac3d_blob *blobs, int nblobs) {
  if (!h) return MAC3D_EINVAL;
  return init_blobs_impl((Ctx *)h, blobs, nblobs);
}

int mac3d_step(void *h, double dt, int nsteps, mac3d_stats *stats) {
  if (!h) return MAC3D_EINVAL;
  Ctx *ctx = (Ctx *)h;
  if (!std::isfinite(dt) || !(dt > 0.0)) return fail(ctx, MAC3D_EINVAL, "dt must be finite and > 0");
  if (nsteps < 0) return fail(ctx, MAC3D_EINVAL, "nsteps must be >= 0");
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, ctx->stream));
  int cycles = 0, iters = 0, conv = 1, rc;
  double rel = -1.0;
  for (int s = 0; s < nsteps; ++s) {
    if ((rc = one_step(ctx, dt, &cycles, &iters, &rel, &conv))) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      return rc;
    }
    if (ctx->prof.pending() > 1024) {
      HIPCHK(hipStreamSynchronize(ctx->stream));
      ctx->prof.collect();
    }
  }
  HIPCHK(hipEventRecord(e1, ctx->stream));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  ctx->prof.collect();
  if (stats) {
    stats->steps = nsteps;
    stats->vcycles = cycles;
    stats->ir_iters = iters;
    stats->converged = conv;
    stats->rel_res = rel;
    stats->ms = ms;
    double mx = 0.0, l2 = 0.0;
    if ((rc = div_norms(ctx, &mx, &l2))) return rc;
    stats->max_div = mx;
    stats->l2_div = l2;
    stats->max_div_rate = mx / ctx->cfg.dx;
    if (!std::isfinite(mx) || !std::isfinite(l2))
      return fail(ctx, MAC3D_ENONFINITE, "the divergence norms are not finite (NaN or inf in the velocity)");
  }
  return MAC3D_OK;
}

int mac3d_checkpoint_size(void *h, size_t *bytes) {
  if (!h) return MAC3D_EINVAL;
  Ctx *ctx = (Ctx *)h;
  if (!bytes) return fail(ctx, MAC3D_EINVAL, "null size pointer");
  *bytes = ckpt_bytes(ctx, ckpt_wave_n(ctx), ctx->np, ckpt_conc(ctx));
  return MAC3D_OK;
}

int mac3d_checkpoint_save(void *h, void *host, size_t bytes) {
  if (!h) return MAC3D_EINVAL;
  Ctx *ctx = (Ctx *)h;
  const int wn = ckpt_wave_n(ctx), np = ctx->np, conc = ckpt_conc(ctx);
  const size_t need = ckpt_bytes(ctx, wn, np, conc);
  if (!host || bytes < need)
    return fail(ctx, MAC3D_EINVAL, "checkpoint buffer is null or smaller than mac3d_checkpoint_size");
  CkptHeader hd;
  std::memset(&hd, 0, sizeof(hd));
  std::memcpy(hd.magic, kCkptMagic, 8);
  hd.abi = MAC3D_ABI_VERSION;
  hd.nx = ctx->cfg.nx;
  hd.ny = ctx->cfg.ny;
  hd.nz = ctx->cfg.nz;
  hd.levels = ctx->L;
  hd.precision = ctx->prec;
  hd.elem = (int32_t)type_size(ctx->t0);
  hd.wave_n = wn;
  hd.np = np;
  hd.hist = ctx->hist;
  hd.iter_n = ctx->iter_n;
  for (int q = 0; q < 4; ++q) hd.iter_log[q] = ctx->iter_log[q];
  hd.conc = conc;
  hd.bytes = need;
  std::memcpy(host, &hd, sizeof(hd));
  char *dst = (char *)host + sizeof(hd);
  for (const CkptSec &c : ckpt_sections(ctx, wn, np, conc)) {
    HIPCHK(hipMemcpyAsync(dst, c.dev, c.bytes, hipMemcpyDeviceToHost, ctx->stream));
    dst += c.bytes;
  }
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return MAC3D_OK;
}

int mac3d_checkpoint_load(void *h, const void *host, size_t bytes) {
  if (!h) return MAC3D_EINVAL;
  Ctx *ctx = (Ctx *)h;
  CkptHeader hd;
  if (!host || bytes < sizeof(hd)) return fail(ctx, MAC3D_EINVAL, "checkpoint buffer too small");
  std::memcpy(&hd, host, sizeof(hd));
  if (std::memcmp(hd.magic, kCkptMagic, 8) != 0 || hd.abi != MAC3D_ABI_VERSION)
    return fail(ctx, MAC3D_EINVAL, "not a mac3d checkpoint of this ABI version");
  if (hd.nx != ctx->cfg.nx || hd.ny != ctx->cfg.ny || hd.nz != ctx->cfg.nz || hd.levels != ctx->L ||
      hd.precision != ctx->prec || hd.elem != (int32_t)type_size(ctx->t0))
    return fail(ctx, MAC3D_EINVAL, "checkpoint grid, levels or precision differ from the context's");
  if (hd.wave_n != ckpt_wave_n(ctx))
    return fail(ctx, MAC3D_EINVAL, "checkpoint wave grid differs from the context's coupling "
                                   "(enable it with the same n first, or disable it)");
  if (hd.np < 0 || (hd.np > 0 && !ctx->wave_on))
    return fail(ctx, MAC3D_EINVAL, "checkpoint holds parcels: enable the wave coupling first");
  if (hd.hist < 0 || hd.hist > 2 || hd.iter_n < 0 || hd.iter_n > 4 || hd.conc < 0 || hd.conc > 2)
    return fail(ctx, MAC3D_EINVAL, "checkpoint header is corrupt");
  if (hd.conc != ckpt_conc(ctx))
    return fail(ctx, MAC3D_EINVAL, "checkpoint's 4-channel concentration differs from the "
                                   "context's (enable it on the same grid, or disable it, first)");
  for (int q = 0; q < 4; ++q)
    if (hd.iter_log[q] < 0) return fail(ctx, MAC3D_EINVAL, "checkpoint header is corrupt");
  const size_t need = ckpt_bytes(ctx, hd.wave_n, hd.np, hd.conc);
  if (hd.bytes != need || bytes < need)
    return fail(ctx, MAC3D_EINVAL, "checkpoint size does not match its header");
  int rc;
  if (hd.np > ctx->np_cap) {  // (an earlier buffer stays owned until destroy)
    if ((rc = dev_alloc(ctx, (void **)&ctx->parcels, (size_t)hd.np * sizeof(Parcel)))) return rc;
    ctx->np_cap = hd.np;
  }
  if (hd.np > 0 && (!ctx->splat || ctx->splat_cap < splat_cells(ctx))) {
    drop_graphs(ctx);  // (a captured parcel stage holds the old counters)
    if ((rc = ensure_splat(ctx))) return rc;
  }
  const char *src = (const char *)host + sizeof(hd);
  for (const CkptSec &c : ckpt_sections(ctx, hd.wave_n, hd.np, hd.conc)) {
    HIPCHK(hipMemcpyAsync(c.dev, src, c.bytes, hipMemcpyHostToDevice, ctx->stream));
    src += c.bytes;
  }
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->np = hd.np;
  ctx->hist = hd.hist;
  ctx->iter_n = hd.iter_n;
  for (int q = 0; q < 4; ++q) ctx->iter_log[q] = hd.iter_log[q];
  ctx->cvel_fresh = false;
  ctx->dmm_valid = false;
  sync_p(ctx);
  return MAC3D_OK;
}

int mac3d_adjoint(void *h, const mac3d_blob *blobs, int nblobs, double dt, int nsteps,
                  int checkpoint_every, const void *target, int dtype, mac3d_grad *out) {
  if (!h) return MAC3D_EINVAL;
  Ctx *ctx = (Ctx *)h;
  if (!out || !target) return fail(ctx, MAC3D_EINVAL, "null target or result");
  if (ctx->wave_on)
    return fail(ctx, MAC3D_ESTATE, "the adjoint covers the flat lid only: disable the wave");
  if (nblobs < 1 || nblobs > MAC3D_MAX_BLOBS || !blobs)
    return fail(ctx, MAC3D_EINVAL, "need 1..MAC3D_MAX_BLOBS blobs");
  if (!std::isfinite(dt) || !(dt > 0.0)) return fail(ctx, MAC3D_EINVAL, "dt must be finite and > 0");
  if (nsteps < 1 || checkpoint_every < 1)
    return fail(ctx, MAC3D_EINVAL, "nsteps and checkpoint_every must be >= 1");
  if (dtype != MAC3D_F32 && dtype != MAC3D_F64) return fail(ctx, MAC3D_EINVAL, "unknown dtype");
  int rc = adjoint_impl(ctx, blobs, nblobs, dt, nsteps, checkpoint_every, target, dtype, out);
  ctx->cvel_fresh = false;  // the reverse sweep reuses the work fields
  ctx->dmm_valid = false;
  ctx->prof.collect();
  return rc;
}

// ---- drop releases (release.hpp): the blobs of the releases, the adjoint
// carried on to the release parameters, and a quasi-Newton fit
int mac3d_release_default_map(const mac3d_config *cfg, mac3d_release_map *m) {
  if (!cfg || !m) return noalloc_fail(MAC3D_EINVAL, "null config or map");
  if (cfg->nx < 1 || cfg->ny < 1 || cfg->nz < 1)
    return noalloc_fail(MAC3D_EINVAL, "grid sizes must be positive");
  // DOMAIN_HALF_XZ 0.48, WATER_SURFACE_Y 0.30 (uniformLayout.ts:9-11,
  // dropKinematics.ts:5); G_SLOW, MAX_DROP_OFFSET_XZ, PENDANT_MIN_RADIUS and
  // the 0.095 radius cap (dropKinematics.ts:8-16, 364)
  const double cell = 2.0 * 0.48 / cfg->nx;
  m->cell = cell;
  m->origin_x = -0.5 * cell * cfg->nx;
  m->origin_z = -0.5 * cell * cfg->nz;
  m->surface_y = 0.30;
  m->origin_y = m->surface_y - cell * cfg->ny;
  m->gravity = -0.155;
  m->vel_scale = 1.0;
  m->submerge = 1.0;
  m->max_offset = 0.26;
  m->r_min = 0.018;
  m->r_max = 0.095;
  return MAC3D_OK;
}

int mac3d_release_blobs(const mac3d_release_map *map, const mac3d_release *rel, int n,
                        mac3d_blob *blobs, double *jac) {
  if (!map || !rel || !blobs) return noalloc_fail(MAC3D_EINVAL, "null map, releases or blobs");
  if (n < 1 || n > MAC3D_MAX_BLOBS) return noalloc_fail(MAC3D_EINVAL, "need 1..MAC3D_MAX_BLOBS releases");
  if (!release_map_ok(*map)) return noalloc_fail(MAC3D_EINVAL, "release map: non-finite value, cell <= 0, gravity >= 0 or bad clamps");
  for (int q = 0; q < n; ++q)
    if (!release_ok(rel[q])) return noalloc_fail(MAC3D_EINVAL, "release parameters must be finite");
  for (int q = 0; q < n; ++q) release_blob(*map, rel[q], blobs[q], jac ? jac + 25 * q : nullptr);
  return MAC3D_OK;
}

int mac3d_adjoint_release(void *h, const mac3d_release_map *map, const mac3d_release *rel, int n,
                          double dt, int nsteps, int checkpoint_every, const void *target,
                          int dtype, mac3d_release_grad *out) {
  if (!h) return MAC3D_EINVAL;
  Ctx *ctx = (Ctx *)h;
  if (!out) return fail(ctx, MAC3D_EINVAL, "null result");
  mac3d_blob blobs[MAC3D_MAX_BLOBS];
  double jac[25 * MAC3D_MAX_BLOBS];
  int rc = mac3d_release_blobs(map, rel, n, blobs, jac);
  if (rc) return fail(ctx, rc, g_create_err);
  mac3d_grad g;
  if ((rc = mac3d_adjoint(h, blobs, n, dt, nsteps, checkpoint_every, target, dtype, &g))) return rc;
  std::memset(out, 0, sizeof *out);
  out->loss = g.loss;
  out->d_rho_b = g.d_rho_b;
  out->n = n;
  out->steps = g.steps;
  for (int q = 0; q < n; ++q)
    for (int b = 0; b < 5; ++b) {
      double a = 0.0;  // (J^T dL/dblob)_b
      for (int r = 0; r < 5; ++r) a += jac[25 * q + 5 * r + b] * g.d_blob[q][r];
      out->d_release[q][b] = a;
      out->d_blob[q][b] = g.d_blob[q][b];
    }
  return MAC3D_OK;
}

int mac3d_fit_default_opts(mac3d_fit_opts *o) {
  if (!o) return noalloc_fail(MAC3D_EINVAL, "null options");
  std::memset(o, 0, sizeof *o);
  o->max_iters = 20;
  o->checkpoint_every = 10;
  o->mask = 0;
  o->tol = 1e-6;
  o->gtol = 1e-9;
  return MAC3D_OK;
}

int mac3d_fit_release(void *h, const mac3d_release_map *map, mac3d_release *rel, int n,
                      double dt, int nsteps, const void *target, int dtype,
                      const mac3d_fit_opts *opts, mac3d_fit_report *report) {
  if (!h) return MAC3D_EINVAL;
  Ctx *ctx = (Ctx *)h;
  mac3d_fit_opts o;
  mac3d_fit_default_opts(&o);
  if (opts) o = *opts;
  if (o.max_iters < 0 || o.checkpoint_every < 1 || !(o.tol >= 0.0) || !(o.gtol >= 0.0))
    return fail(ctx, MAC3D_EINVAL, "fit options: max_iters >= 0, checkpoint_every >= 1, tol and gtol >= 0");
  // the mask has five bits per release: at most 12 releases in 64 bits
  if (n < 1 || n > 12 || !rel) return fail(ctx, MAC3D_EINVAL, "need 1..12 releases to fit");
  if (!std::isfinite(dt) || !(dt > 0.0) || nsteps < 1)
    return fail(ctx, MAC3D_EINVAL, "dt must be finite and > 0, nsteps >= 1");
  uint64_t mask = o.mask;
  if (!mask)
    for (int q = 0; q < n; ++q)
      mask |= (uint64_t)0x17 << (5 * q);  // x, z, radius, vel_y
  if (mask >> (5 * n))
    return fail(ctx, MAC3D_EINVAL, "fit mask names parameters of releases past n");
  // the fitted parameters and their scales (about one cell of blob motion)
  mac3d_blob b0[MAC3D_MAX_BLOBS];
  double jac[25 * MAC3D_MAX_BLOBS];
  int rc = mac3d_release_blobs(map, rel, n, b0, jac);
  if (rc) return fail(ctx, rc, g_create_err);
  const double run = nsteps * dt;
  std::vector<int> qi, bi;
  std::vector<double> scale, p0;
  for (int q = 0; q < n; ++q)
    for (int b = 0; b < 5; ++b) {
      if (!((mask >> (5 * q + b)) & 1)) continue;
      double sc = map->cell;
      if (b == RL_START_Y || b == RL_VEL_Y) {
        const double dw = std::fabs(jac[25 * q + 5 * RB_W0 + b]);
        const double slope = dw > 1e-12 ? dw : std::max(std::fabs(map->vel_scale), 1e-12);
        sc = ctx->cfg.dx / (run * slope);
      }
      qi.push_back(q);
      bi.push_back(b);
      scale.push_back(sc);
      p0.push_back(rel_param(rel[q], b));
    }
  std::vector<mac3d_release> cur(rel, rel + n);
  auto eval = [&](const std::vector<double> &x, double &f, std::vector<double> &g) -> int {
    for (size_t i = 0; i < x.size(); ++i) rel_param(cur[qi[i]], bi[i]) = p0[i] + scale[i] * x[i];
    mac3d_release_grad rg;
    int r = mac3d_adjoint_release(h, map, cur.data(), n, dt, nsteps, o.checkpoint_every,
                                  target, dtype, &rg);
    if (r) return r;
    f = rg.loss;
    for (size_t i = 0; i < x.size(); ++i) g[i] = rg.d_release[qi[i]][bi[i]] * scale[i];
    return MAC3D_OK;
  };
  std::vector<double> x(qi.size(), 0.0);
  FitResult R = bfgs_minimise(x, eval, o.max_iters, o.tol, o.gtol);
  if (R.status) return R.status;
  for (size_t i = 0; i < x.size(); ++i) rel_param(rel[qi[i]], bi[i]) = p0[i] + scale[i] * x[i];
  if (report) {
    std::memset(report, 0, sizeof *report);
    report->iters = R.iters;
    report->evals = R.evals;
    report->converged = R.converged;
    report->loss0 = R.f0;
    report->loss = R.f;
  }
  return MAC3D_OK;
}

int mac3d_stage(void *h, int stage, double dt) {
  if (!h) return MAC3D_EINVAL;
  Ctx *ctx = (Ctx *)h;
  if (!std::isfinite(dt) || dt < 0.0) return fail(ctx, MAC3D_EINVAL, "dt must be finite and >= 0");
  int rc;
  if (stage == MAC3D_STAGE_PARCELS) {  // (reads the velocity, changes none of it)
    if (!ctx->wave_on || !ctx->splat) return fail(ctx, MAC3D_ESTATE, "needs the coupling and parcels");
    // the parcels sample the centre velocity of the current faces: form it
    // first when the faces changed since the last face->centre pass
    if (!ctx->cvel_fresh) {
      if ((rc = launch_face_to_centre(ctx))) return rc;
      ctx->cvel_fresh = true;
    }
    if ((rc = launch_parcels(ctx, dt))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->prof.collect();
    return MAC3D_OK;
  }
  ctx->cvel_fresh = false;
  switch (stage) {
    case MAC3D_STAGE_FACE_TO_CENTRE: rc = launch_face_to_centre(ctx); break;
    case MAC3D_STAGE_ADVECT: rc = launch_advect(ctx, dt); break;
    case MAC3D_STAGE_CASSON: rc = launch_casson(ctx, dt); break;
    case MAC3D_STAGE_DIVERGENCE: rc = launch_divergence(ctx); break;
    case MAC3D_STAGE_VCYCLE: rc = reset_scale(ctx) ? MAC3D_EHIP : vcycle(ctx, false); break;
    case MAC3D_STAGE_PROJECT: rc = launch_project(ctx); break;
    case MAC3D_STAGE_DYE_PREDICT: rc = launch_dye_predict(ctx, dt); break;
    case MAC3D_STAGE_DYE_RESOLVE: rc = launch_dye_resolve(ctx, dt); break;
    case MAC3D_STAGE_CONC_PREDICT:
    case MAC3D_STAGE_CONC_RESOLVE:
      if (!ctx->conc_on) return fail(ctx, MAC3D_ESTATE, "the 4-channel concentration is off");
      rc = stage == MAC3D_STAGE_CONC_PREDICT ? launch_conc_predict(ctx, dt)
                                             : launch_conc_resolve(ctx, dt);
      break;
    case MAC3D_STAGE_WAVE:
    case MAC3D_STAGE_TOP_PRESSURE:
      if (!ctx->wave_on) return fail(ctx, MAC3D_ESTATE, "the wave coupling is not enabled");
      rc = stage == MAC3D_STAGE_WAVE ? launch_wave_step(ctx, dt) : launch_top_pressure(ctx, dt);
      break;
    default: return fail(ctx, MAC3D_EINVAL, "unknown stage");
  }
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->prof.collect();
  if (stage == MAC3D_STAGE_FACE_TO_CENTRE) ctx->cvel_fresh = true;
  return MAC3D_OK;
}

int mac3d_mg_op(void *h, int op, int level, int count) {
  if (!h) return MAC3D_EINVAL;
  Ctx *ctx = (Ctx *)h;
  ctx->cvel_fresh = false;
  if (level < 0 || level > ctx->L) return fail(ctx, MAC3D_EINVAL, "level out of range");
  if (count < 0) return fail(ctx, MAC3D_EINVAL, "count must be >= 0");
  int rc;
  switch (op) {
    case MAC3D_MG_SMOOTH:
      rc = reset_scale(ctx) ? MAC3D_EHIP : mg_smooth(ctx, level, count, false);
      break;
    case MAC3D_MG_RESIDUAL: rc = reset_scale(ctx) ? MAC3D_EHIP : mg_residual(ctx, level); break;
    case MAC3D_MG_RESTRICT:
      if (level >= ctx->L) return fail(ctx, MAC3D_EINVAL, "restrict needs level < levels");
      rc = mg_restrict(ctx, level);
      break;
    case MAC3D_MG_PROLONG:
      if (level >= ctx->L) return fail(ctx, MAC3D_EINVAL, "prolong needs level < levels");
      rc = mg_prolong(ctx, level);
      break;
    case MAC3D_MG_COARSE:
      if (level != ctx->L) return fail(ctx, MAC3D_EINVAL, "coarse solve runs on the coarsest level");
      rc = mg_coarse(ctx);
      break;
    case MAC3D_MG_ZERO: rc = mg_fill_zero(ctx, ctx->mgx[level]); break;
    case MAC3D_MG_SMOOTH_ZERO:
      rc = reset_scale(ctx) ? MAC3D_EHIP : mg_smooth(ctx, level, count, true);
      break;
    case MAC3D_MG_SMOOTH2_MARCH:
      if (ctx->mgx[level].type != TF32 && ctx->mgx[level].type != TF16)
        return fail(ctx, MAC3D_EINVAL, "the z-marching pass takes FP32 or FP16 levels");
      if (count > 1) return fail(ctx, MAC3D_EINVAL, "count: 0 (given guess) or 1 (zero guess)");
      rc = reset_scale(ctx) ? MAC3D_EHIP : mg_sweep2(ctx, level, count == 1, true);
      break;
    default: return fail(ctx, MAC3D_EINVAL, "unknown multigrid op");
  }
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->prof.collect();
  return MAC3D_OK;
}

int mac3d_div_norms(void *h, double *max_div, double *l2_div) {
  if (!h || !max_div || !l2_div) return MAC3D_EINVAL;
  return div_norms((Ctx *)h, max_div, l2_div);
}

int mac3d_profile(void *h, int enable) {
  if (!h) return MAC3D_EINVAL;
  Ctx *ctx = (Ctx *)h;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->prof.collect();
  ctx->prof.reset();
  ctx->prof.enabled = enable != 0;
  return MAC3D_OK;
}

int mac3d_profile_read(void *h, int kclass, double *ms_total, int64_t *launches) {
  if (!h || !ms_total || !launches) return MAC3D_EINVAL;
  Ctx *ctx = (Ctx *)h;
  if (kclass < 0 || kclass >= KC_COUNT) return fail(ctx, MAC3D_EINVAL, "unknown kernel class");
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->prof.collect();
  *ms_total = ctx->prof.ms[kclass];
  *launches = ctx->prof.count[kclass];
  return MAC3D_OK;
}

int mac3d_div_norms_noalloc(const void *u, const void *v, const void *w, int dtype, int nx,
                            int ny, int nz, double *out, void *stream) {
  if (!u || !v || !w || !out) return noalloc_fail(MAC3D_EINVAL, "null field or output pointer");
  if (dtype != MAC3D_F32 && dtype != MAC3D_F64)
    return noalloc_fail(MAC3D_EINVAL, "dtype must be MAC3D_F32 or MAC3D_F64");
  if (nx < 1 || ny < 1 || nz < 1) return noalloc_fail(MAC3D_EINVAL, "need nx, ny, nz >= 1");
  const Dims d{nx, ny, nz};
  const dim3 g = cell_grid(d);
  const size_t parts = (size_t)g.x * g.y * g.z;
  if (g.y > 65535 || g.z > 65535 || parts > kNoAllocMaxParts)
    return noalloc_fail(MAC3D_EINVAL,
                        "grid too large for the static scratch: ceil(nx/64) ceil(ny/4) nz = " +
                            std::to_string(parts) + " partial sums, at most " +
                            std::to_string(kNoAllocMaxParts) + " (and ceil(ny/4), nz <= 65535)");
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return noalloc_fail(MAC3D_EHIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kNoAllocMaxDevices) return noalloc_fail(MAC3D_EINVAL, "device ordinal out of range");
  double *scr = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {  // once per device, with that device current
      void *p = nullptr;
      e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p)
        return noalloc_fail(MAC3D_EHIP, std::string("hipGetSymbolAddress(scratch): ") +
                                            hipGetErrorString(e));
      g_scratch_dev[dev] = (double *)p;
    }
    scr = g_scratch_dev[dev];
  }
  // every block writes its pair, so nothing in the scratch needs clearing
  hipStream_t st = (hipStream_t)stream;
  if (dtype == MAC3D_F32)
    hipLaunchKernelGGL(k_div_partials<float>, g, kCellBlock, 0, st, d, (const float *)u,
                       (const float *)v, (const float *)w, scr);
  else
    hipLaunchKernelGGL(k_div_partials<double>, g, kCellBlock, 0, st, d, (const double *)u,
                       (const double *)v, (const double *)w, scr);
  hipLaunchKernelGGL(k_fold_div_norms, dim3(1), dim3(256), 0, st, (const double *)scr, parts,
                     (double)d.cells(), out);
  e = hipGetLastError();
  if (e != hipSuccess) return noalloc_fail(MAC3D_EHIP, std::string("kernel launch: ") + hipGetErrorString(e));
  return MAC3D_OK;
}

const char *mac3d_last_error(void *h) {
  if (!h) return g_create_err.c_str();
  return ((Ctx *)h)->err.c_str();
}

const char *mac3d_version(void) { return "mac3d 0.2.0 (gfx950, abi 2)"; }

}  // extern "C"
