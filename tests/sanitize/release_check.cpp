// Host code of the library under AddressSanitizer and
// UndefinedBehaviorSanitizer: the release map and its Jacobian
// (release.hpp) on releases that reach every branch (inside and outside
// the offset disk, both radius clamps, a drop thrown upward, one released at
// the surface), checked against central differences, and the BFGS fit on
// the Rosenbrock function.  Built by tests/test_sanitizers.py with g++.
#include <cmath>
#include <cstdio>
#include <vector>

#include "../../drop-of-blood_amd/csrc/release.hpp"

using namespace mac3d;

static int fails = 0;
#define CHECK(c, ...)                \
  do {                               \
    if (!(c)) {                      \
      std::printf(__VA_ARGS__);      \
      std::printf("\n");             \
      ++fails;                       \
    }                                \
  } while (0)

int main() {
  mac3d_release_map m;
  m.cell = 0.02;
  m.origin_x = -0.48;
  m.origin_z = -0.48;
  m.surface_y = 0.30;
  m.origin_y = 0.30 - 64 * 0.02;
  m.gravity = -0.155;
  m.vel_scale = 2.5;
  m.submerge = 1.3;
  m.max_offset = 0.26;
  m.r_min = 0.018;
  m.r_max = 0.095;
  CHECK(release_map_ok(m), "map rejected");
  const mac3d_release rel[] = {{0.05, -0.03, 0.058, 0.495, -0.075}, {-0.12, 0.2, 0.03, 0.55, 0.0},
                               {0.0, 0.0, 0.07, 0.45, 0.2},        {0.3, 0.2, 0.05, 0.5, -0.1},
                               {0.01, 0.02, 0.12, 0.6, -0.05},     {0.01, 0.02, 0.01, 0.6, -0.05},
                               {0.02, 0.0, 0.04, 0.32, -0.3}};
  for (const mac3d_release &r0 : rel) {
    mac3d_blob b;
    double J[25];
    release_blob(m, r0, b, J);
    const double bv[5] = {b.cx, b.cy, b.cz, b.radius, b.w0};
    for (int k = 0; k < 5; ++k) CHECK(std::isfinite(bv[k]), "non-finite blob");
    for (int p = 0; p < 5; ++p) {
      mac3d_release lo = r0, hi = r0;
      const double h = 1e-7;
      rel_param(lo, p) -= h;
      rel_param(hi, p) += h;
      mac3d_blob bl, bh;
      release_blob(m, lo, bl, nullptr);
      release_blob(m, hi, bh, nullptr);
      const double l[5] = {bl.cx, bl.cy, bl.cz, bl.radius, bl.w0};
      const double u[5] = {bh.cx, bh.cy, bh.cz, bh.radius, bh.w0};
      for (int a = 0; a < 5; ++a) {
        const double fd = (u[a] - l[a]) / (2 * h), an = J[5 * a + p];
        CHECK(std::fabs(fd - an) <= 1e-5 * (1.0 + std::fabs(an)), "jacobian (%d,%d): %g vs %g", a,
              p, an, fd);
      }
    }
  }
  // BFGS on the Rosenbrock function from (-1.2, 1)
  std::vector<double> x = {-1.2, 1.0};
  auto f = [](const std::vector<double> &v, double &fx, std::vector<double> &g) {
    const double a = 1.0 - v[0], b = v[1] - v[0] * v[0];
    fx = a * a + 100.0 * b * b;
    g[0] = -2.0 * a - 400.0 * v[0] * b;
    g[1] = 200.0 * b;
    return 0;
  };
  FitResult R = bfgs_minimise(x, f, 200, 0.0, 1e-12);
  CHECK(R.status == 0 && R.f < 1e-12 && std::fabs(x[0] - 1) < 1e-5 && std::fabs(x[1] - 1) < 1e-5,
        "bfgs: f %g at (%g, %g) after %d iterations", R.f, x[0], x[1], R.iters);
  std::printf("release_check: %d failures, bfgs %d iterations %d evaluations\n", fails, R.iters,
              R.evals);
  return fails ? 1 : 0;
}
