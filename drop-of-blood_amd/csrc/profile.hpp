// profile.hpp — per-kernel-class timing with HIP events on the context's
// stream (enabled by mac3d_profile; off by default and then free).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/mac3d.h"

namespace mac3d {

enum {
  KC_F2C = MAC3D_KC_FACE_TO_CENTRE,
  KC_ADVECT = MAC3D_KC_ADVECT,
  KC_CASSON = MAC3D_KC_CASSON,
  KC_DIV = MAC3D_KC_DIVERGENCE,
  KC_SMOOTH0 = MAC3D_KC_SMOOTH_L0,
  KC_RESID0 = MAC3D_KC_RESTRICT_L0,
  KC_PROLONG0 = MAC3D_KC_PROLONG_L0,
  KC_MG = MAC3D_KC_MG_COARSE_LEVELS,
  KC_COARSE = MAC3D_KC_COARSE_SOLVE,
  KC_PROJECT = MAC3D_KC_PROJECT,
  KC_DYE = MAC3D_KC_DYE,
  KC_RESID64 = MAC3D_KC_REFINE64,
  KC_SMOOTH0X2 = MAC3D_KC_SMOOTH2_L0,
  KC_WAVE = MAC3D_KC_WAVE,
  KC_PARCELS = MAC3D_KC_PARCELS,
  KC_COUNT = MAC3D_KC_COUNT
};

struct Profiler {
  bool enabled = false;
  std::vector<hipEvent_t> pool;
  std::vector<int> cls;
  size_t used = 0;
  double ms[KC_COUNT];
  long long count[KC_COUNT];

  int init(std::string &) {
    reset();
    return 0;
  }
  void reset() {
    used = 0;
    cls.clear();
    for (int k = 0; k < KC_COUNT; ++k) { ms[k] = 0.0; count[k] = 0; }
  }
  size_t pending() const { return used / 2; }
  void ensure_pool() {
    if (pool.empty()) {
      pool.resize(4096);
      for (auto &e : pool) (void)hipEventCreate(&e);
    }
  }
  void start(hipStream_t s, int kc) {
    ensure_pool();
    if (used + 2 > pool.size()) {
      (void)hipStreamSynchronize(s);
      collect();
    }
    cls.push_back(kc);
    (void)hipEventRecord(pool[used], s);
  }
  void stop(hipStream_t s) {
    (void)hipEventRecord(pool[used + 1], s);
    used += 2;
  }
  // caller has synchronised the stream
  void collect() {
    for (size_t q = 0; q < used / 2; ++q) {
      float t = 0.f;
      if (hipEventElapsedTime(&t, pool[2 * q], pool[2 * q + 1]) == hipSuccess) {
        ms[cls[q]] += t;
        count[cls[q]] += 1;
      }
    }
    used = 0;
    cls.clear();
  }
  void destroy() {
    for (auto &e : pool) (void)hipEventDestroy(e);
    pool.clear();
  }
};

struct ProfScope {
  Profiler &p;
  hipStream_t s;
  bool on;
  ProfScope(Profiler &p_, hipStream_t s_, int kc) : p(p_), s(s_), on(p_.enabled) {
    if (on) p.start(s, kc);
  }
  ~ProfScope() {
    if (on) p.stop(s);
  }
};

}  // namespace mac3d
