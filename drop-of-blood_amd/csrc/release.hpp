// release.hpp — drop releases as blobs of the grid, and fitting them (host
// code).  The release model is MultiDropSystem.spawnAerialDrop and the aerial
// branch of MultiDropSystem.step (src/sim/dropKinematics.ts:354-391,
// 651-707): a clamped offset and radius, then free fall under G_SLOW with
// the contact time solved exactly, which makes the contact state a closed
// form of the release.  The blob gradient of the adjoint is carried to the
// release parameters through that closed form's Jacobian.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <vector>

#include "mac3d.h"

namespace mac3d {

enum { RB_CX = 0, RB_CY, RB_CZ, RB_RADIUS, RB_W0 };  // blob rows of the Jacobian
enum { RL_X = 0, RL_Z, RL_RADIUS, RL_START_Y, RL_VEL_Y };  // release columns

// release parameter b (RL_*) of r
inline double &rel_param(mac3d_release &r, int b) {
  switch (b) {
    case RL_X: return r.x;
    case RL_Z: return r.z;
    case RL_RADIUS: return r.radius;
    case RL_START_Y: return r.start_y;
    default: return r.vel_y;
  }
}

inline bool release_map_ok(const mac3d_release_map &m) {
  const double v[] = {m.origin_x, m.origin_y, m.origin_z, m.cell, m.surface_y, m.gravity,
                      m.vel_scale, m.submerge, m.max_offset, m.r_min, m.r_max};
  for (double x : v)
    if (!std::isfinite(x)) return false;
  return m.cell > 0.0 && m.gravity < 0.0 && m.max_offset > 0.0 && m.r_min > 0.0 &&
         m.r_max >= m.r_min;
}

inline bool release_ok(const mac3d_release &r) {
  return std::isfinite(r.x) && std::isfinite(r.z) && std::isfinite(r.radius) &&
         std::isfinite(r.start_y) && std::isfinite(r.vel_y);
}

// One release: the blob, and (if J) d blob[a] / d release[b] at J[5 a + b].
inline void release_blob(const mac3d_release_map &m, const mac3d_release &r, mac3d_blob &b,
                         double *J) {
  double jl[25] = {0};
  // clampDropDiskXZ (dropKinematics.ts:49-58): radial scaling onto the disk
  double xc = r.x, zc = r.z;
  double dxx = 1.0, dxz = 0.0, dzx = 0.0, dzz = 1.0;
  const double rr = std::hypot(r.x, r.z);
  if (!(rr <= m.max_offset || rr < 1e-8)) {
    const double s = m.max_offset / rr;
    xc = r.x * s;
    zc = r.z * s;
    const double r3 = rr * rr * rr;
    dxx = m.max_offset * r.z * r.z / r3;
    dxz = -m.max_offset * r.x * r.z / r3;
    dzx = dxz;
    dzz = m.max_offset * r.x * r.x / r3;
  }
  // radius clamp (spawnAerialDrop: clamp(radius, PENDANT_MIN_RADIUS, 0.095))
  const double R = std::min(std::max(r.radius, m.r_min), m.r_max);
  const double dR = (r.radius >= m.r_min && r.radius <= m.r_max) ? 1.0 : 0.0;
  // free fall until y - R = surface: the contact speed from the fall height
  // H (the exact sub-step contact solve of dropKinematics.ts:667-676 lands
  // on this value whatever the substep); already at or below the surface,
  // the drop keeps its release velocity (dtAir = 0)
  const double H = r.start_y - R - m.surface_y;
  const double disc = r.vel_y * r.vel_y - 2.0 * m.gravity * H;
  double vc, dvc_dv, dvc_dy, dvc_dR;
  if (H > 0.0 && disc > 0.0) {
    const double sq = std::sqrt(disc);
    vc = -sq;
    dvc_dv = -r.vel_y / sq;
    dvc_dy = m.gravity / sq;    // d(-sq)/dH = g/sq, dH/dstart_y = 1
    dvc_dR = -m.gravity / sq * dR;  // dH/dR = -1
  } else {
    vc = r.vel_y;
    dvc_dv = 1.0;
    dvc_dy = 0.0;
    dvc_dR = 0.0;
  }
  const double ic = 1.0 / m.cell;
  b.cx = (xc - m.origin_x) * ic;
  b.cz = (zc - m.origin_z) * ic;
  b.cy = (m.surface_y - m.submerge * R - m.origin_y) * ic;
  b.radius = R * ic;
  b.w0 = -m.vel_scale * vc;
  if (!J) return;
  jl[5 * RB_CX + RL_X] = dxx * ic;
  jl[5 * RB_CX + RL_Z] = dxz * ic;
  jl[5 * RB_CZ + RL_X] = dzx * ic;
  jl[5 * RB_CZ + RL_Z] = dzz * ic;
  jl[5 * RB_CY + RL_RADIUS] = -m.submerge * dR * ic;
  jl[5 * RB_RADIUS + RL_RADIUS] = dR * ic;
  jl[5 * RB_W0 + RL_RADIUS] = -m.vel_scale * dvc_dR;
  jl[5 * RB_W0 + RL_START_Y] = -m.vel_scale * dvc_dy;
  jl[5 * RB_W0 + RL_VEL_Y] = -m.vel_scale * dvc_dv;
  std::memcpy(J, jl, sizeof jl);
}

// BFGS with a backtracking (Armijo) line search on x (already scaled), for
// f returning the loss and writing the gradient; false from f aborts.
struct FitResult {
  int iters = 0, evals = 0, converged = 0;
  double f0 = 0.0, f = 0.0;
  int status = 0;  // the first failing evaluation's status
};

inline FitResult bfgs_minimise(std::vector<double> &x,
                               const std::function<int(const std::vector<double> &, double &,
                                                       std::vector<double> &)> &f,
                               int max_iters, double tol, double gtol) {
  FitResult R;
  const size_t n = x.size();
  std::vector<double> g(n), H(n * n, 0.0), p(n), xt(n), gt(n), s(n), y(n), Hy(n);
  double fx;
  if ((R.status = f(x, fx, g))) return R;
  R.evals = 1;
  R.f0 = R.f = fx;
  bool scaled_h = false;
  for (size_t i = 0; i < n; ++i) H[i * n + i] = 1.0;
  for (int it = 0; it < max_iters; ++it) {
    double gmax = 0.0;
    for (double v : g) gmax = std::max(gmax, std::fabs(v));
    if (gmax <= gtol * std::max(fx, 1e-300)) {
      R.converged = 1;
      break;
    }
    for (size_t i = 0; i < n; ++i) {
      double a = 0.0;
      for (size_t j = 0; j < n; ++j) a -= H[i * n + j] * g[j];
      p[i] = a;
    }
    double slope = 0.0;
    for (size_t i = 0; i < n; ++i) slope += p[i] * g[i];
    if (!(slope < 0.0)) {  // not a descent direction: restart from steepest descent
      std::fill(H.begin(), H.end(), 0.0);
      for (size_t i = 0; i < n; ++i) H[i * n + i] = 1.0, p[i] = -g[i];
      slope = 0.0;
      for (size_t i = 0; i < n; ++i) slope += p[i] * g[i];
      scaled_h = false;
    }
    // before the first curvature pair, a step of at most one scaled unit
    double pmax = 0.0;
    for (double v : p) pmax = std::max(pmax, std::fabs(v));
    double t = (!scaled_h && pmax > 1.0) ? 1.0 / pmax : 1.0;
    double ft = 0.0;
    bool ok = false;
    for (int ls = 0; ls < 12; ++ls) {
      for (size_t i = 0; i < n; ++i) xt[i] = x[i] + t * p[i];
      if ((R.status = f(xt, ft, gt))) return R;
      ++R.evals;
      if (std::isfinite(ft) && ft <= fx + 1e-4 * t * slope) {
        ok = true;
        break;
      }
      t *= 0.5;
    }
    if (!ok) break;  // no decrease along p: a stationary point to this precision
    ++R.iters;
    double sy = 0.0;
    for (size_t i = 0; i < n; ++i) {
      s[i] = xt[i] - x[i];
      y[i] = gt[i] - g[i];
      sy += s[i] * y[i];
    }
    const double drop = fx - ft;
    x = xt;
    g = gt;
    fx = ft;
    R.f = fx;
    if (sy > 1e-300) {
      if (!scaled_h) {  // Nocedal-Wright (6.20): H0 = (s.y / y.y) I before the first update
        double yy = 0.0;
        for (double v : y) yy += v * v;
        std::fill(H.begin(), H.end(), 0.0);
        for (size_t i = 0; i < n; ++i) H[i * n + i] = sy / yy;
        scaled_h = true;
      }
      // H <- (I - rho s y^T) H (I - rho y s^T) + rho s s^T
      const double rho = 1.0 / sy;
      double yHy = 0.0;
      for (size_t i = 0; i < n; ++i) {
        double a = 0.0;
        for (size_t j = 0; j < n; ++j) a += H[i * n + j] * y[j];
        Hy[i] = a;
        yHy += y[i] * a;
      }
      for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j)
          H[i * n + j] += -rho * (Hy[i] * s[j] + s[i] * Hy[j]) +
                          (rho * rho * yHy + rho) * s[i] * s[j];
    }
    if (drop <= tol * std::max(fx, 1e-300)) {
      R.converged = 1;
      break;
    }
  }
  return R;
}

}  // namespace mac3d
