// mac3d_napi.cpp — Node N-API addon: a thin shim from JavaScript onto the
// C ABI of include/mac3d.h.  No compute here: every call forwards to
// libmac3d.so; a non-zero status becomes a thrown JS Error carrying
// mac3d_last_error (the reference throws from create and on device errors,
// src/engine/BloodDropEngine.ts:337-345,1465-1476).
#include <node_api.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "mac3d.h"

namespace {

#define NAPI_CALL(env, call)                                   \
  do {                                                         \
    if ((call) != napi_ok) {                                   \
      napi_throw_error((env), nullptr, "N-API call failed: " #call); \
      return nullptr;                                          \
    }                                                          \
  } while (0)

napi_value throw_status(napi_env env, void *ctx, int rc) {
  std::string msg = std::string("mac3d status ") + std::to_string(rc) + ": " + mac3d_last_error(ctx);
  napi_throw_error(env, nullptr, msg.c_str());
  return nullptr;
}

bool get_args(napi_env env, napi_callback_info info, size_t want, napi_value *argv) {
  size_t argc = want;
  if (napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr) != napi_ok || argc < want) {
    napi_throw_type_error(env, nullptr, "missing arguments");
    return false;
  }
  return true;
}

void *get_handle(napi_env env, napi_value v) {
  void *p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, nullptr, "expected a live solver handle");
    return nullptr;
  }
  return p;
}

double get_double(napi_env env, napi_value v) {
  double d = NAN;
  napi_get_value_double(env, v, &d);
  return d;
}

int get_int(napi_env env, napi_value v) {
  int32_t i = 0;
  napi_get_value_int32(env, v, &i);
  return i;
}

bool prop_num(napi_env env, napi_value obj, const char *name, double *out) {
  bool has = false;
  if (napi_has_named_property(env, obj, name, &has) != napi_ok || !has) return false;
  napi_value v;
  napi_get_named_property(env, obj, name, &v);
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t != napi_number) return false;
  napi_get_value_double(env, v, out);
  return true;
}

void set_num(napi_env env, napi_value obj, const char *name, double v) {
  napi_value n;
  napi_create_double(env, v, &n);
  napi_set_named_property(env, obj, name, n);
}

#define CONFIG_INT_FIELDS(X) X(nx) X(ny) X(nz) X(levels) X(nu1) X(nu2) X(cycles) \
  X(precision) X(ir_max_iters) X(device) X(extrapolate)
#define CONFIG_DBL_FIELDS(X) X(dx) X(rho_w) X(mu_w) X(rho_b) X(tau_y) X(mu_inf) X(g) \
  X(gamma_reg) X(shear_eps2) X(alpha_max) X(vel_damp) X(maccormack) X(ir_tol) X(sor_omega)

napi_value config_to_js(napi_env env, const mac3d_config &c) {
  napi_value o;
  napi_create_object(env, &o);
#define X(f) set_num(env, o, #f, (double)c.f);
  CONFIG_INT_FIELDS(X)
  CONFIG_DBL_FIELDS(X)
#undef X
  return o;
}

void config_from_js(napi_env env, napi_value o, mac3d_config &c) {
  double v;
#define X(f) if (prop_num(env, o, #f, &v)) c.f = (int32_t)v;
  CONFIG_INT_FIELDS(X)
#undef X
#define X(f) if (prop_num(env, o, #f, &v)) c.f = v;
  CONFIG_DBL_FIELDS(X)
#undef X
}

napi_value DefaultConfig(napi_env env, napi_callback_info) {
  mac3d_config c;
  mac3d_default_config(&c);
  return config_to_js(env, c);
}

napi_value Create(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  mac3d_config c;
  mac3d_default_config(&c);
  config_from_js(env, argv[0], c);
  void *ctx = nullptr;
  int rc = mac3d_create(&c, &ctx);
  if (rc) return throw_status(env, nullptr, rc);
  napi_value h;
  NAPI_CALL(env, napi_create_external(env, ctx, nullptr, nullptr, &h));
  return h;
}

napi_value Destroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  void *ctx = get_handle(env, argv[0]);
  if (!ctx) return nullptr;
  mac3d_destroy(ctx);
  return nullptr;
}

// (handle, field, level, typedArray) for upload and download
napi_value Transfer(napi_env env, napi_callback_info info, bool up) {
  napi_value argv[4];
  if (!get_args(env, info, 4, argv)) return nullptr;
  void *ctx = get_handle(env, argv[0]);
  if (!ctx) return nullptr;
  int field = get_int(env, argv[1]), level = get_int(env, argv[2]);
  bool is_ta = false;
  napi_is_typedarray(env, argv[3], &is_ta);
  if (!is_ta) {
    napi_throw_type_error(env, nullptr, "expected a Float32Array or Float64Array");
    return nullptr;
  }
  napi_typedarray_type tt;
  size_t len = 0, off = 0;
  void *data = nullptr;
  napi_value ab;
  NAPI_CALL(env, napi_get_typedarray_info(env, argv[3], &tt, &len, &data, &ab, &off));
  int dtype;
  if (tt == napi_float32_array) dtype = MAC3D_F32;
  else if (tt == napi_float64_array) dtype = MAC3D_F64;
  else {
    napi_throw_type_error(env, nullptr, "expected a Float32Array or Float64Array");
    return nullptr;
  }
  int rc = up ? mac3d_upload(ctx, field, level, data, dtype, len)
              : mac3d_download(ctx, field, level, data, dtype, len);
  if (rc) return throw_status(env, ctx, rc);
  return nullptr;
}
napi_value Upload(napi_env env, napi_callback_info info) { return Transfer(env, info, true); }
napi_value Download(napi_env env, napi_callback_info info) { return Transfer(env, info, false); }

napi_value InitBlobs(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  void *ctx = get_handle(env, argv[0]);
  if (!ctx) return nullptr;
  uint32_t n = 0;
  NAPI_CALL(env, napi_get_array_length(env, argv[1], &n));
  std::vector<mac3d_blob> blobs(n);
  for (uint32_t q = 0; q < n; ++q) {
    napi_value o;
    NAPI_CALL(env, napi_get_element(env, argv[1], q, &o));
    mac3d_blob &b = blobs[q];
    b.cx = b.cy = b.cz = b.radius = b.w0 = NAN;
    prop_num(env, o, "cx", &b.cx);
    prop_num(env, o, "cy", &b.cy);
    prop_num(env, o, "cz", &b.cz);
    prop_num(env, o, "radius", &b.radius);
    prop_num(env, o, "w0", &b.w0);
  }
  int rc = mac3d_init_blobs(ctx, blobs.data(), (int)n);
  if (rc) return throw_status(env, ctx, rc);
  return nullptr;
}

// waveEnable(h, params | null): mac3d_wave_enable; missing fields keep the
// defaults of mac3d_wave_default_params
napi_value WaveEnable(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  void *ctx = get_handle(env, argv[0]);
  if (!ctx) return nullptr;
  napi_valuetype t;
  NAPI_CALL(env, napi_typeof(env, argv[1], &t));
  int rc;
  if (t == napi_null || t == napi_undefined) {
    rc = mac3d_wave_enable(ctx, nullptr);
  } else {
    mac3d_wave_params p;
    mac3d_wave_default_params(&p);
    double n = p.n;
    prop_num(env, argv[1], "n", &n);
    p.n = (int32_t)n;
    prop_num(env, argv[1], "halfExtent", &p.half_extent);
    prop_num(env, argv[1], "halfXZ", &p.half_xz);
    prop_num(env, argv[1], "depth", &p.depth);
    prop_num(env, argv[1], "surfaceY", &p.surface_y);
    prop_num(env, argv[1], "marangoni", &p.marangoni);
    prop_num(env, argv[1], "vortConf", &p.vort_conf);
    prop_num(env, argv[1], "simTime", &p.sim_time);
    prop_num(env, argv[1], "schlieren", &p.schlieren);
    prop_num(env, argv[1], "hemolysis", &p.hemolysis);
    prop_num(env, argv[1], "cassonScale", &p.casson_scale);
    prop_num(env, argv[1], "buoyancy", &p.buoyancy);
    double scene = 0.0, physics = 0.0;
    prop_num(env, argv[1], "scene", &scene);
    prop_num(env, argv[1], "physics", &physics);
    p.scene = scene != 0.0;
    p.physics = (int32_t)physics;
    rc = mac3d_wave_enable(ctx, &p);
  }
  if (rc) return throw_status(env, ctx, rc);
  return nullptr;
}

// setDrops(h, [{x, y, z, radius, velY, osc, drain, seed}, ...])
napi_value SetDrops(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  void *ctx = get_handle(env, argv[0]);
  if (!ctx) return nullptr;
  uint32_t n = 0;
  NAPI_CALL(env, napi_get_array_length(env, argv[1], &n));
  std::vector<mac3d_drop> drops(n ? n : 1);
  for (uint32_t q = 0; q < n; ++q) {
    napi_value o;
    NAPI_CALL(env, napi_get_element(env, argv[1], q, &o));
    mac3d_drop &d = drops[q];
    d.x = d.y = d.z = d.radius = d.vel_y = d.drain = d.seed = NAN;
    d.osc = 0.0;
    prop_num(env, o, "x", &d.x);
    prop_num(env, o, "y", &d.y);
    prop_num(env, o, "z", &d.z);
    prop_num(env, o, "radius", &d.radius);
    prop_num(env, o, "velY", &d.vel_y);
    prop_num(env, o, "osc", &d.osc);
    prop_num(env, o, "drain", &d.drain);
    prop_num(env, o, "seed", &d.seed);
  }
  int rc = mac3d_set_drops(ctx, drops.data(), (int)n);
  if (rc) return throw_status(env, ctx, rc);
  return nullptr;
}

// setUniforms(h, Float32Array(128)): the packed SimUniforms block
napi_value SetUniforms(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  void *ctx = get_handle(env, argv[0]);
  if (!ctx) return nullptr;
  napi_typedarray_type tt;
  size_t len = 0, off = 0;
  void *data = nullptr;
  napi_value ab;
  NAPI_CALL(env, napi_get_typedarray_info(env, argv[1], &tt, &len, &data, &ab, &off));
  if (tt != napi_float32_array) {
    napi_throw_type_error(env, nullptr, "expected a Float32Array");
    return nullptr;
  }
  int rc = mac3d_set_uniforms(ctx, (const float *)data, len);
  if (rc) return throw_status(env, ctx, rc);
  return nullptr;
}

// the typed array argv[i] of type want: its data and element count
static bool typed_arg(napi_env env, napi_value v, napi_typedarray_type want, const char *what,
                      void **data, size_t *len) {
  napi_typedarray_type tt;
  size_t off = 0;
  napi_value ab;
  if (napi_get_typedarray_info(env, v, &tt, len, data, &ab, &off) != napi_ok || tt != want) {
    napi_throw_type_error(env, nullptr, what);
    return false;
  }
  return true;
}

// parcelsUpload(h, Float32Array(8 n)) / parcelsDownload(h, Float32Array(8 n)):
// the parcel storage buffer (RbcParcel, 8 floats each)
napi_value ParcelsUpload(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  void *ctx = get_handle(env, argv[0]);
  if (!ctx) return nullptr;
  void *data = nullptr;
  size_t len = 0;
  if (!typed_arg(env, argv[1], napi_float32_array, "expected a Float32Array", &data, &len))
    return nullptr;
  if (len % 8) {
    napi_throw_range_error(env, nullptr, "parcels are 8 floats each");
    return nullptr;
  }
  int rc = mac3d_parcels_upload(ctx, (const float *)data, (int32_t)(len / 8));
  if (rc) return throw_status(env, ctx, rc);
  return nullptr;
}

napi_value ParcelsDownload(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  void *ctx = get_handle(env, argv[0]);
  if (!ctx) return nullptr;
  void *data = nullptr;
  size_t len = 0;
  if (!typed_arg(env, argv[1], napi_float32_array, "expected a Float32Array", &data, &len))
    return nullptr;
  int rc = mac3d_parcels_download(ctx, (float *)data, (int32_t)(len / 8));
  if (rc) return throw_status(env, ctx, rc);
  return nullptr;
}

// splatDownload(h, which, Uint32Array(splat grid cells)): 0 intact, 1 lysed
napi_value SplatDownload(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  void *ctx = get_handle(env, argv[0]);
  if (!ctx) return nullptr;
  int32_t which = 0;
  NAPI_CALL(env, napi_get_value_int32(env, argv[1], &which));
  void *data = nullptr;
  size_t len = 0;
  if (!typed_arg(env, argv[2], napi_uint32_array, "expected a Uint32Array", &data, &len))
    return nullptr;
  int rc = mac3d_splat_download(ctx, which, (uint32_t *)data, len);
  if (rc) return throw_status(env, ctx, rc);
  return nullptr;
}

napi_value Step(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  void *ctx = get_handle(env, argv[0]);
  if (!ctx) return nullptr;
  mac3d_stats st;
  std::memset(&st, 0, sizeof(st));
  int rc = mac3d_step(ctx, get_double(env, argv[1]), get_int(env, argv[2]), &st);
  if (rc) return throw_status(env, ctx, rc);
  napi_value o;
  NAPI_CALL(env, napi_create_object(env, &o));
  set_num(env, o, "steps", st.steps);
  set_num(env, o, "vcycles", st.vcycles);
  set_num(env, o, "irIters", st.ir_iters);
  set_num(env, o, "converged", st.converged);
  set_num(env, o, "maxDiv", st.max_div);
  set_num(env, o, "l2Div", st.l2_div);
  set_num(env, o, "maxDivRate", st.max_div_rate);
  set_num(env, o, "relRes", st.rel_res);
  set_num(env, o, "ms", st.ms);
  return o;
}

napi_value Stage(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  void *ctx = get_handle(env, argv[0]);
  if (!ctx) return nullptr;
  int rc = mac3d_stage(ctx, get_int(env, argv[1]), get_double(env, argv[2]));
  if (rc) return throw_status(env, ctx, rc);
  return nullptr;
}

napi_value DivNorms(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  void *ctx = get_handle(env, argv[0]);
  if (!ctx) return nullptr;
  double mx = 0, l2 = 0;
  int rc = mac3d_div_norms(ctx, &mx, &l2);
  if (rc) return throw_status(env, ctx, rc);
  napi_value o;
  NAPI_CALL(env, napi_create_object(env, &o));
  set_num(env, o, "maxDiv", mx);
  set_num(env, o, "l2Div", l2);
  return o;
}

// (handle, blobs[], dt, steps, checkpointEvery, target Float32Array|Float64Array)
napi_value Adjoint(napi_env env, napi_callback_info info) {
  napi_value argv[6];
  if (!get_args(env, info, 6, argv)) return nullptr;
  void *ctx = get_handle(env, argv[0]);
  if (!ctx) return nullptr;
  uint32_t n = 0;
  NAPI_CALL(env, napi_get_array_length(env, argv[1], &n));
  std::vector<mac3d_blob> blobs(n);
  for (uint32_t q = 0; q < n; ++q) {
    napi_value o;
    NAPI_CALL(env, napi_get_element(env, argv[1], q, &o));
    mac3d_blob &b = blobs[q];
    b.cx = b.cy = b.cz = b.radius = b.w0 = NAN;
    prop_num(env, o, "cx", &b.cx);
    prop_num(env, o, "cy", &b.cy);
    prop_num(env, o, "cz", &b.cz);
    prop_num(env, o, "radius", &b.radius);
    prop_num(env, o, "w0", &b.w0);
  }
  napi_typedarray_type tt;
  size_t len = 0, off = 0;
  void *data = nullptr;
  napi_value ab;
  bool is_ta = false;
  napi_is_typedarray(env, argv[5], &is_ta);
  if (!is_ta) {
    napi_throw_type_error(env, nullptr, "target must be a Float32Array or Float64Array");
    return nullptr;
  }
  NAPI_CALL(env, napi_get_typedarray_info(env, argv[5], &tt, &len, &data, &ab, &off));
  if (tt != napi_float32_array && tt != napi_float64_array) {
    napi_throw_type_error(env, nullptr, "target must be a Float32Array or Float64Array");
    return nullptr;
  }
  mac3d_grad g;
  std::memset(&g, 0, sizeof(g));
  int rc = mac3d_adjoint(ctx, blobs.data(), (int)n, get_double(env, argv[2]), get_int(env, argv[3]),
                         get_int(env, argv[4]), data,
                         tt == napi_float32_array ? MAC3D_F32 : MAC3D_F64, &g);
  if (rc) return throw_status(env, ctx, rc);
  napi_value o, arr;
  NAPI_CALL(env, napi_create_object(env, &o));
  set_num(env, o, "loss", g.loss);
  set_num(env, o, "dRhoB", g.d_rho_b);
  NAPI_CALL(env, napi_create_array_with_length(env, n, &arr));
  const char *names[5] = {"cx", "cy", "cz", "radius", "w0"};
  for (uint32_t q = 0; q < n; ++q) {
    napi_value e;
    NAPI_CALL(env, napi_create_object(env, &e));
    for (int p = 0; p < 5; ++p) set_num(env, e, names[p], g.d_blob[q][p]);
    NAPI_CALL(env, napi_set_element(env, arr, q, e));
  }
  NAPI_CALL(env, napi_set_named_property(env, o, "dBlobs", arr));
  return o;
}

// ---- drop releases: {x, z, radius, startY, velY} objects, the release map
// as {originX, originY, originZ, cell, surfaceY, gravity, velScale, submerge,
// maxOffset, rMin, rMax}
#define MAP_FIELDS(X) X(origin_x, "originX") X(origin_y, "originY") X(origin_z, "originZ") \
  X(cell, "cell") X(surface_y, "surfaceY") X(gravity, "gravity") X(vel_scale, "velScale") \
  X(submerge, "submerge") X(max_offset, "maxOffset") X(r_min, "rMin") X(r_max, "rMax")
const char *kReleaseNames[5] = {"x", "z", "radius", "startY", "velY"};
const char *kBlobNames[5] = {"cx", "cy", "cz", "radius", "w0"};

napi_value map_to_js(napi_env env, const mac3d_release_map &m) {
  napi_value o;
  napi_create_object(env, &o);
#define X(f, js) set_num(env, o, js, m.f);
  MAP_FIELDS(X)
#undef X
  return o;
}

// every field must be given (a missing one is NaN, which the library rejects)
void map_from_js(napi_env env, napi_value o, mac3d_release_map &m) {
#define X(f, js) m.f = NAN; prop_num(env, o, js, &m.f);
  MAP_FIELDS(X)
#undef X
}

bool releases_from_js(napi_env env, napi_value arr, std::vector<mac3d_release> &rel) {
  uint32_t n = 0;
  if (napi_get_array_length(env, arr, &n) != napi_ok) {
    napi_throw_type_error(env, nullptr, "releases must be an array");
    return false;
  }
  rel.assign(n, mac3d_release{});
  for (uint32_t q = 0; q < n; ++q) {
    napi_value o;
    napi_get_element(env, arr, q, &o);
    double *f[5] = {&rel[q].x, &rel[q].z, &rel[q].radius, &rel[q].start_y, &rel[q].vel_y};
    for (int b = 0; b < 5; ++b) {
      *f[b] = NAN;
      prop_num(env, o, kReleaseNames[b], f[b]);
    }
  }
  return true;
}

napi_value releases_to_js(napi_env env, const mac3d_release *rel, size_t n) {
  napi_value arr;
  napi_create_array_with_length(env, n, &arr);
  for (size_t q = 0; q < n; ++q) {
    napi_value e;
    napi_create_object(env, &e);
    const double f[5] = {rel[q].x, rel[q].z, rel[q].radius, rel[q].start_y, rel[q].vel_y};
    for (int b = 0; b < 5; ++b) set_num(env, e, kReleaseNames[b], f[b]);
    napi_set_element(env, arr, (uint32_t)q, e);
  }
  return arr;
}

// a dye target: Float32Array or Float64Array
bool target_arg(napi_env env, napi_value v, void **data, int *dtype) {
  bool is_ta = false;
  napi_is_typedarray(env, v, &is_ta);
  napi_typedarray_type tt = napi_int8_array;
  size_t len = 0, off = 0;
  napi_value ab;
  if (is_ta) napi_get_typedarray_info(env, v, &tt, &len, data, &ab, &off);
  if (!is_ta || (tt != napi_float32_array && tt != napi_float64_array)) {
    napi_throw_type_error(env, nullptr, "target must be a Float32Array or Float64Array");
    return false;
  }
  *dtype = tt == napi_float32_array ? MAC3D_F32 : MAC3D_F64;
  return true;
}

// releaseDefaultMap(config) -> map
napi_value ReleaseDefaultMap(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  mac3d_config c;
  mac3d_default_config(&c);
  config_from_js(env, argv[0], c);
  mac3d_release_map m;
  int rc = mac3d_release_default_map(&c, &m);
  if (rc) return throw_status(env, nullptr, rc);
  return map_to_js(env, m);
}

// releaseBlobs(map, releases) -> [{cx, cy, cz, radius, w0}]
napi_value ReleaseBlobs(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  mac3d_release_map m;
  map_from_js(env, argv[0], m);
  std::vector<mac3d_release> rel;
  if (!releases_from_js(env, argv[1], rel)) return nullptr;
  std::vector<mac3d_blob> blobs(rel.size() ? rel.size() : 1);
  int rc = mac3d_release_blobs(&m, rel.data(), (int)rel.size(), blobs.data(), nullptr);
  if (rc) return throw_status(env, nullptr, rc);
  napi_value arr;
  NAPI_CALL(env, napi_create_array_with_length(env, rel.size(), &arr));
  for (size_t q = 0; q < rel.size(); ++q) {
    napi_value e;
    NAPI_CALL(env, napi_create_object(env, &e));
    const double f[5] = {blobs[q].cx, blobs[q].cy, blobs[q].cz, blobs[q].radius, blobs[q].w0};
    for (int b = 0; b < 5; ++b) set_num(env, e, kBlobNames[b], f[b]);
    NAPI_CALL(env, napi_set_element(env, arr, (uint32_t)q, e));
  }
  return arr;
}

// adjointRelease(h, map, releases, dt, steps, checkpointEvery, target)
//   -> {loss, dRhoB, dReleases: [{x, z, radius, startY, velY}], dBlobs: [...]}
napi_value AdjointRelease(napi_env env, napi_callback_info info) {
  napi_value argv[7];
  if (!get_args(env, info, 7, argv)) return nullptr;
  void *ctx = get_handle(env, argv[0]);
  if (!ctx) return nullptr;
  mac3d_release_map m;
  map_from_js(env, argv[1], m);
  std::vector<mac3d_release> rel;
  if (!releases_from_js(env, argv[2], rel)) return nullptr;
  void *data = nullptr;
  int dtype = 0;
  if (!target_arg(env, argv[6], &data, &dtype)) return nullptr;
  mac3d_release_grad g;
  std::memset(&g, 0, sizeof(g));
  int rc = mac3d_adjoint_release(ctx, &m, rel.data(), (int)rel.size(), get_double(env, argv[3]),
                                 get_int(env, argv[4]), get_int(env, argv[5]), data, dtype, &g);
  if (rc) return throw_status(env, ctx, rc);
  napi_value o, dr, db;
  NAPI_CALL(env, napi_create_object(env, &o));
  set_num(env, o, "loss", g.loss);
  set_num(env, o, "dRhoB", g.d_rho_b);
  NAPI_CALL(env, napi_create_array_with_length(env, g.n, &dr));
  NAPI_CALL(env, napi_create_array_with_length(env, g.n, &db));
  for (int q = 0; q < g.n; ++q) {
    napi_value a, b;
    NAPI_CALL(env, napi_create_object(env, &a));
    NAPI_CALL(env, napi_create_object(env, &b));
    for (int p = 0; p < 5; ++p) {
      set_num(env, a, kReleaseNames[p], g.d_release[q][p]);
      set_num(env, b, kBlobNames[p], g.d_blob[q][p]);
    }
    NAPI_CALL(env, napi_set_element(env, dr, q, a));
    NAPI_CALL(env, napi_set_element(env, db, q, b));
  }
  NAPI_CALL(env, napi_set_named_property(env, o, "dReleases", dr));
  NAPI_CALL(env, napi_set_named_property(env, o, "dBlobs", db));
  return o;
}

// fitRelease(h, map, releases, dt, steps, target, {maxIters, checkpointEvery, mask,
//   tol, gtol}) -> {releases, iters, evals, converged, loss0, loss}; mask is a
//   Number (bits 5 q + b, q < 10)
napi_value FitRelease(napi_env env, napi_callback_info info) {
  napi_value argv[7];
  if (!get_args(env, info, 7, argv)) return nullptr;
  void *ctx = get_handle(env, argv[0]);
  if (!ctx) return nullptr;
  mac3d_release_map m;
  map_from_js(env, argv[1], m);
  std::vector<mac3d_release> rel;
  if (!releases_from_js(env, argv[2], rel)) return nullptr;
  void *data = nullptr;
  int dtype = 0;
  if (!target_arg(env, argv[5], &data, &dtype)) return nullptr;
  mac3d_fit_opts o;
  mac3d_fit_default_opts(&o);
  napi_valuetype t;
  NAPI_CALL(env, napi_typeof(env, argv[6], &t));
  if (t == napi_object) {
    double v;
    if (prop_num(env, argv[6], "maxIters", &v)) o.max_iters = (int32_t)v;
    if (prop_num(env, argv[6], "checkpointEvery", &v)) o.checkpoint_every = (int32_t)v;
    if (prop_num(env, argv[6], "mask", &v)) {
      if (!(v >= 0.0 && v < 9007199254740992.0) || v != std::floor(v)) {
        napi_throw_range_error(env, nullptr, "mask must be a non-negative integer below 2^53");
        return nullptr;
      }
      o.mask = (uint64_t)v;
    }
    if (prop_num(env, argv[6], "tol", &v)) o.tol = v;
    if (prop_num(env, argv[6], "gtol", &v)) o.gtol = v;
  }
  mac3d_fit_report r;
  std::memset(&r, 0, sizeof(r));
  int rc = mac3d_fit_release(ctx, &m, rel.data(), (int)rel.size(), get_double(env, argv[3]),
                             get_int(env, argv[4]), data, dtype, &o, &r);
  if (rc) return throw_status(env, ctx, rc);
  napi_value out;
  NAPI_CALL(env, napi_create_object(env, &out));
  NAPI_CALL(env, napi_set_named_property(env, out, "releases",
                                         releases_to_js(env, rel.data(), rel.size())));
  set_num(env, out, "iters", r.iters);
  set_num(env, out, "evals", r.evals);
  set_num(env, out, "converged", r.converged);
  set_num(env, out, "loss0", r.loss0);
  set_num(env, out, "loss", r.loss);
  return out;
}

// concEnable(h, on): false/0 off, true/1 on the MAC grid, 2 on the grid
// twice as fine per axis
napi_value ConcEnable(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  void *ctx = get_handle(env, argv[0]);
  if (!ctx) return nullptr;
  int e = 0;
  napi_valuetype t;
  NAPI_CALL(env, napi_typeof(env, argv[1], &t));
  if (t == napi_boolean) {
    bool on = false;
    napi_get_value_bool(env, argv[1], &on);
    e = on ? 1 : 0;
  } else {
    e = get_int(env, argv[1]);
  }
  int rc = mac3d_conc_enable(ctx, e);
  if (rc) return throw_status(env, ctx, rc);
  return nullptr;
}

// parcelsEmit(h, cursor, count)
napi_value ParcelsEmit(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  void *ctx = get_handle(env, argv[0]);
  if (!ctx) return nullptr;
  int rc = mac3d_parcels_emit(ctx, get_int(env, argv[1]), get_int(env, argv[2]));
  if (rc) return throw_status(env, ctx, rc);
  return nullptr;
}

// checkpointSave(h) -> ArrayBuffer holding mac3d_checkpoint_save's bytes
napi_value CheckpointSave(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  void *ctx = get_handle(env, argv[0]);
  if (!ctx) return nullptr;
  size_t n = 0;
  int rc = mac3d_checkpoint_size(ctx, &n);
  if (rc) return throw_status(env, ctx, rc);
  void *data = nullptr;
  napi_value ab;
  NAPI_CALL(env, napi_create_arraybuffer(env, n, &data, &ab));
  rc = mac3d_checkpoint_save(ctx, data, n);
  if (rc) return throw_status(env, ctx, rc);
  return ab;
}

// checkpointLoad(h, Uint8Array | ArrayBuffer)
napi_value CheckpointLoad(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  void *ctx = get_handle(env, argv[0]);
  if (!ctx) return nullptr;
  void *data = nullptr;
  size_t n = 0;
  bool is_ab = false, is_ta = false;
  napi_is_arraybuffer(env, argv[1], &is_ab);
  napi_is_typedarray(env, argv[1], &is_ta);
  if (is_ab) {
    NAPI_CALL(env, napi_get_arraybuffer_info(env, argv[1], &data, &n));
  } else if (is_ta) {
    napi_typedarray_type t;
    size_t len = 0, off = 0;
    napi_value ab;
    NAPI_CALL(env, napi_get_typedarray_info(env, argv[1], &t, &len, &data, &ab, &off));
    if (t != napi_uint8_array) {
      napi_throw_type_error(env, nullptr, "checkpoint must be a Uint8Array or an ArrayBuffer");
      return nullptr;
    }
    n = len;
  } else {
    napi_throw_type_error(env, nullptr, "checkpoint must be a Uint8Array or an ArrayBuffer");
    return nullptr;
  }
  int rc = mac3d_checkpoint_load(ctx, data, n);
  if (rc) return throw_status(env, ctx, rc);
  return nullptr;
}

napi_value Version(napi_env env, napi_callback_info) {
  napi_value s;
  NAPI_CALL(env, napi_create_string_utf8(env, mac3d_version(), NAPI_AUTO_LENGTH, &s));
  return s;
}

napi_value Init(napi_env env, napi_value exports) {
  napi_property_descriptor props[] = {
      {"defaultConfig", nullptr, DefaultConfig, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"create", nullptr, Create, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"destroy", nullptr, Destroy, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"upload", nullptr, Upload, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"download", nullptr, Download, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"initBlobs", nullptr, InitBlobs, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"waveEnable", nullptr, WaveEnable, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"setDrops", nullptr, SetDrops, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"setUniforms", nullptr, SetUniforms, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"parcelsUpload", nullptr, ParcelsUpload, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"parcelsDownload", nullptr, ParcelsDownload, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"splatDownload", nullptr, SplatDownload, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"step", nullptr, Step, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"stage", nullptr, Stage, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"divNorms", nullptr, DivNorms, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"adjoint", nullptr, Adjoint, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"checkpointSave", nullptr, CheckpointSave, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"concEnable", nullptr, ConcEnable, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"parcelsEmit", nullptr, ParcelsEmit, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"checkpointLoad", nullptr, CheckpointLoad, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"releaseDefaultMap", nullptr, ReleaseDefaultMap, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"releaseBlobs", nullptr, ReleaseBlobs, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"adjointRelease", nullptr, AdjointRelease, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"fitRelease", nullptr, FitRelease, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"version", nullptr, Version, nullptr, nullptr, nullptr, napi_default, nullptr},
  };
  napi_define_properties(env, exports, sizeof(props) / sizeof(props[0]), props);
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
