// fenix_napi.cpp — Node-API addon over libfenix (include/fenix.h) for the JS host
// (fenix.js). Each export is a thin call of one fx_* entry point: plain JS values
// and typed arrays in, status checked, the library's message thrown as an Error,
// as the reference's TypeScript API throws. Built by fenix_amd/build.py with the
// system Node headers; links libfenix.so from the directory above.
#include <node_api.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fenix.h"

namespace {

#define NAPI_OK(call)                                                   \
  do {                                                                  \
    if ((call) != napi_ok) {                                            \
      napi_throw_error(env, nullptr, "fenix: Node-API call failed");    \
      return nullptr;                                                   \
    }                                                                   \
  } while (0)

napi_value undefined_value(napi_env env) {
  napi_value u;
  napi_get_undefined(env, &u);
  return u;
}

// Throws the simulation's last error (the reference throws Error(message)).
napi_value throw_status(napi_env env, fx_sim* sim, int rc, const char* what) {
  const char* msg = fx_last_error(sim);
  std::string text = std::string(what) + ": " + (msg && *msg ? msg : "error") + " (status " + std::to_string(rc) + ")";
  napi_throw_error(env, nullptr, text.c_str());
  return nullptr;
}

bool get_args(napi_env env, napi_callback_info info, size_t want, napi_value* argv) {
  size_t argc = want;
  if (napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr) != napi_ok || argc < want) {
    napi_throw_type_error(env, nullptr, "fenix: missing arguments");
    return false;
  }
  return true;
}

fx_sim* sim_of(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, nullptr, "fenix: not a simulation handle (disposed?)");
    return nullptr;
  }
  return static_cast<fx_sim*>(p);
}

bool has(napi_env env, napi_value obj, const char* key) {
  bool r = false;
  napi_valuetype t;
  if (napi_typeof(env, obj, &t) != napi_ok || t != napi_object) return false;
  if (napi_has_named_property(env, obj, key, &r) != napi_ok || !r) return false;
  napi_value v;
  napi_get_named_property(env, obj, key, &v);
  napi_typeof(env, v, &t);
  return t != napi_undefined && t != napi_null;
}
double num(napi_env env, napi_value obj, const char* key, double dflt) {
  if (!has(env, obj, key)) return dflt;
  napi_value v;
  double d = dflt;
  napi_get_named_property(env, obj, key, &v);
  napi_get_value_double(env, v, &d);
  return d;
}
napi_value child(napi_env env, napi_value obj, const char* key) {
  napi_value v = nullptr;
  if (has(env, obj, key)) napi_get_named_property(env, obj, key, &v);
  return v;
}
void vec3(napi_env env, napi_value obj, const char* key, float* out) {
  napi_value a = child(env, obj, key);
  if (!a) return;
  for (uint32_t i = 0; i < 3; ++i) {
    napi_value e;
    double d = out[i];
    if (napi_get_element(env, a, i, &e) == napi_ok) napi_get_value_double(env, e, &d);
    out[i] = float(d);
  }
}
napi_value make_num(napi_env env, double d) {
  napi_value v;
  napi_create_double(env, d, &v);
  return v;
}
void set_num(napi_env env, napi_value obj, const char* key, double d) { napi_set_named_property(env, obj, key, make_num(env, d)); }
void set_vec3(napi_env env, napi_value obj, const char* key, const float* v) {
  napi_value a;
  napi_create_array_with_length(env, 3, &a);
  for (uint32_t i = 0; i < 3; ++i) napi_set_element(env, a, i, make_num(env, v[i]));
  napi_set_named_property(env, obj, key, a);
}

// SimulationRuntimeParams keys (CVS/types.ts:46-58) -> fx_runtime_params, present-mask bit
void runtime_from(napi_env env, napi_value o, fx_runtime_params* r, uint32_t* mask) {
  uint32_t m = 0;
  if (has(env, o, "wind")) { vec3(env, o, "wind", r->wind); m |= FX_RP_WIND; }
  if (has(env, o, "gravity")) { vec3(env, o, "gravity", r->gravity); m |= FX_RP_GRAVITY; }
  struct K { const char* key; float* dst; uint32_t bit; } keys[] = {
      {"windStrength", &r->wind_strength, FX_RP_WIND_STRENGTH},
      {"gravityStrength", &r->gravity_strength, FX_RP_GRAVITY_STRENGTH},
      {"buoyancy", &r->buoyancy, FX_RP_BUOYANCY},
      {"vorticityStrength", &r->vorticity_strength, FX_RP_VORTICITY_STRENGTH},
      {"vorticityConstantMask", &r->vorticity_constant_mask, FX_RP_VORTICITY_CONSTANT_MASK},
      {"vorticityVelocityMask", &r->vorticity_velocity_mask, FX_RP_VORTICITY_VELOCITY_MASK},
      {"vorticityHeatMask", &r->vorticity_heat_mask, FX_RP_VORTICITY_HEAT_MASK},
      {"vorticityDensityMask", &r->vorticity_density_mask, FX_RP_VORTICITY_DENSITY_MASK},
      {"worldSize", &r->world_size, FX_RP_WORLD_SIZE},
  };
  for (const K& k : keys)
    if (has(env, o, k.key)) { *k.dst = float(num(env, o, k.key, 0)); m |= k.bit; }
  *mask = m;
}
// SimulationQualitySettings keys (CVS/types.ts:6-14)
void quality_from(napi_env env, napi_value o, fx_quality* q) {
  q->pressure_interval = int32_t(num(env, o, "pressureInterval", q->pressure_interval));
  q->vorticity_interval = int32_t(num(env, o, "vorticityInterval", q->vorticity_interval));
  q->fine_pre_iterations = int32_t(num(env, o, "finePreIterations", q->fine_pre_iterations));
  q->fine_post_iterations = int32_t(num(env, o, "finePostIterations", q->fine_post_iterations));
  q->mid_pre_iterations = int32_t(num(env, o, "midPreIterations", q->mid_pre_iterations));
  q->mid_post_iterations = int32_t(num(env, o, "midPostIterations", q->mid_post_iterations));
  q->coarse_iterations = int32_t(num(env, o, "coarseIterations", q->coarse_iterations));
}

// create(desc) -> handle. desc: resolution fields plus the library knobs; absent
// keys keep fx_default_create_desc (the reference defaults, ORCH:54, 194-204).
napi_value js_create(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!get_args(env, info, 1, a)) return nullptr;
  fx_create_desc d;
  fx_default_create_desc(&d);
  d.width = uint32_t(num(env, a[0], "width", d.width));
  d.height = uint32_t(num(env, a[0], "height", d.height));
  d.depth = uint32_t(num(env, a[0], "depth", d.depth));
  d.advection_mode = int32_t(num(env, a[0], "advectionMode", d.advection_mode));
  if (napi_value rt = child(env, a[0], "runtime")) {
    uint32_t mask = 0;
    runtime_from(env, rt, &d.runtime, &mask);
  }
  if (napi_value q = child(env, a[0], "quality")) quality_from(env, q, &d.quality);
  d.domain_mode = int32_t(num(env, a[0], "domainMode", d.domain_mode));
  d.activation_rule = int32_t(num(env, a[0], "activationRule", d.activation_rule));
  d.solver = int32_t(num(env, a[0], "solver", d.solver));
  d.precision = int32_t(num(env, a[0], "precision", d.precision));
  if (napi_value v = child(env, a[0], "vcycle")) {
    d.vcycle.levels = int32_t(num(env, v, "levels", d.vcycle.levels));
    d.vcycle.pre_sweeps = int32_t(num(env, v, "preSweeps", d.vcycle.pre_sweeps));
    d.vcycle.post_sweeps = int32_t(num(env, v, "postSweeps", d.vcycle.post_sweeps));
    d.vcycle.cycles = int32_t(num(env, v, "cycles", d.vcycle.cycles));
    d.vcycle.coarse_sweeps = int32_t(num(env, v, "coarseSweeps", d.vcycle.coarse_sweeps));
    d.vcycle.omega = float(num(env, v, "omega", d.vcycle.omega));
  }
  d.ignition_lo = float(num(env, a[0], "ignitionLo", d.ignition_lo));
  d.ignition_hi = float(num(env, a[0], "ignitionHi", d.ignition_hi));
  d.burn_rate = float(num(env, a[0], "burnRate", d.burn_rate));
  d.activation_density = float(num(env, a[0], "activationDensity", d.activation_density));
  d.activation_speed = float(num(env, a[0], "activationSpeed", d.activation_speed));
  d.debug_fields = int32_t(num(env, a[0], "debugFields", d.debug_fields));
  d.device = int32_t(num(env, a[0], "device", d.device));
  fx_sim* sim = nullptr;
  const int rc = fx_create(&d, &sim);
  if (rc != FX_OK) return throw_status(env, nullptr, rc, "createCombustionVolumeSimulation");
  napi_value h;
  NAPI_OK(napi_create_external(env, sim, nullptr, nullptr, &h));
  return h;
}

napi_value js_destroy(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!get_args(env, info, 1, a)) return nullptr;
  fx_sim* s = sim_of(env, a[0]);
  if (!s) return nullptr;
  fx_destroy(s);
  return undefined_value(env);
}

#define SIM_CALL(name, expr)                                   \
  napi_value name(napi_env env, napi_callback_info info) {    \
    napi_value a[1];                                           \
    if (!get_args(env, info, 1, a)) return nullptr;            \
    fx_sim* s = sim_of(env, a[0]);                             \
    if (!s) return nullptr;                                    \
    const int rc = (expr);                                     \
    if (rc != FX_OK) return throw_status(env, s, rc, #name);   \
    return undefined_value(env);                               \
  }
SIM_CALL(js_reset, fx_reset(s))
SIM_CALL(js_synchronize, fx_synchronize(s))
SIM_CALL(js_request_debug, fx_request_debug_fields(s))

napi_value js_step(napi_env env, napi_callback_info info) {
  napi_value a[3];
  if (!get_args(env, info, 3, a)) return nullptr;
  fx_sim* s = sim_of(env, a[0]);
  if (!s) return nullptr;
  double e = 0, dt = 0;
  NAPI_OK(napi_get_value_double(env, a[1], &e));
  NAPI_OK(napi_get_value_double(env, a[2], &dt));
  const int rc = fx_step(s, e, dt);
  if (rc != FX_OK) return throw_status(env, s, rc, "step");
  return undefined_value(env);
}

// updateSources(h, Float32Array packed, count)
napi_value js_update_sources(napi_env env, napi_callback_info info) {
  napi_value a[3];
  if (!get_args(env, info, 3, a)) return nullptr;
  fx_sim* s = sim_of(env, a[0]);
  if (!s) return nullptr;
  napi_typedarray_type t;
  size_t len = 0;
  void* data = nullptr;
  NAPI_OK(napi_get_typedarray_info(env, a[1], &t, &len, &data, nullptr, nullptr));
  uint32_t count = 0;
  NAPI_OK(napi_get_value_uint32(env, a[2], &count));
  if (t != napi_float32_array || len < size_t(count) * FX_FLOATS_PER_EMITTER) {
    napi_throw_type_error(env, nullptr, "updateSources: packed sources must be a Float32Array of 32 floats per source");
    return nullptr;
  }
  const int rc = fx_update_sources(s, static_cast<const float*>(data), count);
  if (rc != FX_OK) return throw_status(env, s, rc, "updateSources");
  return undefined_value(env);
}

napi_value js_get_runtime(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!get_args(env, info, 1, a)) return nullptr;
  fx_sim* s = sim_of(env, a[0]);
  if (!s) return nullptr;
  fx_runtime_params r;
  const int rc = fx_get_runtime_params(s, &r);
  if (rc != FX_OK) return throw_status(env, s, rc, "getRuntimeParams");
  napi_value o;
  NAPI_OK(napi_create_object(env, &o));
  set_vec3(env, o, "wind", r.wind);
  set_num(env, o, "windStrength", r.wind_strength);
  set_vec3(env, o, "gravity", r.gravity);
  set_num(env, o, "gravityStrength", r.gravity_strength);
  set_num(env, o, "buoyancy", r.buoyancy);
  set_num(env, o, "vorticityStrength", r.vorticity_strength);
  set_num(env, o, "vorticityConstantMask", r.vorticity_constant_mask);
  set_num(env, o, "vorticityVelocityMask", r.vorticity_velocity_mask);
  set_num(env, o, "vorticityHeatMask", r.vorticity_heat_mask);
  set_num(env, o, "vorticityDensityMask", r.vorticity_density_mask);
  set_num(env, o, "worldSize", r.world_size);
  return o;
}

napi_value js_set_runtime(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!get_args(env, info, 2, a)) return nullptr;
  fx_sim* s = sim_of(env, a[0]);
  if (!s) return nullptr;
  fx_runtime_params r;
  std::memset(&r, 0, sizeof r);
  uint32_t mask = 0;
  runtime_from(env, a[1], &r, &mask);
  const int rc = fx_set_runtime_params(s, &r, mask);
  if (rc != FX_OK) return throw_status(env, s, rc, "setRuntimeParams");
  return undefined_value(env);
}

napi_value js_get_quality(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!get_args(env, info, 1, a)) return nullptr;
  fx_sim* s = sim_of(env, a[0]);
  if (!s) return nullptr;
  fx_quality q;
  const int rc = fx_get_quality(s, &q);
  if (rc != FX_OK) return throw_status(env, s, rc, "getQualitySettings");
  napi_value o;
  NAPI_OK(napi_create_object(env, &o));
  set_num(env, o, "pressureInterval", q.pressure_interval);
  set_num(env, o, "vorticityInterval", q.vorticity_interval);
  set_num(env, o, "finePreIterations", q.fine_pre_iterations);
  set_num(env, o, "finePostIterations", q.fine_post_iterations);
  set_num(env, o, "midPreIterations", q.mid_pre_iterations);
  set_num(env, o, "midPostIterations", q.mid_post_iterations);
  set_num(env, o, "coarseIterations", q.coarse_iterations);
  return o;
}

// setQualitySettings(h, partial): merged over the current settings; the library clamps.
napi_value js_set_quality(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!get_args(env, info, 2, a)) return nullptr;
  fx_sim* s = sim_of(env, a[0]);
  if (!s) return nullptr;
  fx_quality q;
  int rc = fx_get_quality(s, &q);
  if (rc == FX_OK) {
    quality_from(env, a[1], &q);
    rc = fx_set_quality(s, &q);
  }
  if (rc != FX_OK) return throw_status(env, s, rc, "setQualitySettings");
  return undefined_value(env);
}

napi_value js_get_mode(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!get_args(env, info, 1, a)) return nullptr;
  fx_sim* s = sim_of(env, a[0]);
  if (!s) return nullptr;
  int32_t m = 0;
  const int rc = fx_get_advection_mode(s, &m);
  if (rc != FX_OK) return throw_status(env, s, rc, "getScalarAdvectionMode");
  return make_num(env, m);
}

napi_value js_set_mode(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!get_args(env, info, 2, a)) return nullptr;
  fx_sim* s = sim_of(env, a[0]);
  if (!s) return nullptr;
  int32_t m = 0;
  NAPI_OK(napi_get_value_int32(env, a[1], &m));
  const int rc = fx_set_advection_mode(s, m);
  if (rc != FX_OK) return throw_status(env, s, rc, "setScalarAdvectionMode");
  return undefined_value(env);
}

napi_value typed(napi_env env, napi_typedarray_type t, size_t count, size_t esz, void** data) {
  napi_value buf, arr;
  if (napi_create_arraybuffer(env, count * esz, data, &buf) != napi_ok) return nullptr;
  if (napi_create_typedarray(env, t, count, buf, 0, &arr) != napi_ok) return nullptr;
  return arr;
}

// readField(h, fieldId, voxelCount) -> Float32Array (row-major x + W (y + H z))
napi_value js_read_field(napi_env env, napi_callback_info info) {
  napi_value a[3];
  if (!get_args(env, info, 3, a)) return nullptr;
  fx_sim* s = sim_of(env, a[0]);
  if (!s) return nullptr;
  int32_t id = 0;
  double n = 0;
  NAPI_OK(napi_get_value_int32(env, a[1], &id));
  NAPI_OK(napi_get_value_double(env, a[2], &n));
  void* data = nullptr;
  napi_value arr = typed(env, napi_float32_array, size_t(n), 4, &data);
  if (!arr) return nullptr;
  const int rc = fx_read_field(s, id, static_cast<float*>(data), size_t(n) * 4);
  if (rc != FX_OK) return throw_status(env, s, rc, "readField");
  return arr;
}

// readFlags(h, which, brickCount) -> Uint32Array
napi_value js_read_flags(napi_env env, napi_callback_info info) {
  napi_value a[3];
  if (!get_args(env, info, 3, a)) return nullptr;
  fx_sim* s = sim_of(env, a[0]);
  if (!s) return nullptr;
  int32_t which = 0;
  uint32_t n = 0;
  NAPI_OK(napi_get_value_int32(env, a[1], &which));
  NAPI_OK(napi_get_value_uint32(env, a[2], &n));
  void* data = nullptr;
  napi_value arr = typed(env, napi_uint32_array, n, 4, &data);
  if (!arr) return nullptr;
  const int rc = fx_read_flags(s, which, static_cast<uint32_t*>(data), n);
  if (rc != FX_OK) return throw_status(env, s, rc, "readFlags");
  return arr;
}

napi_value js_active_bricks(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!get_args(env, info, 1, a)) return nullptr;
  fx_sim* s = sim_of(env, a[0]);
  if (!s) return nullptr;
  uint32_t n = 0;
  const int rc = fx_active_brick_count(s, &n);
  if (rc != FX_OK) return throw_status(env, s, rc, "activeBrickCount");
  return make_num(env, n);
}

// raymarch(h, camera, params, width, height) -> Uint16Array of RGBA binary16 (row 0 = top)
napi_value js_raymarch(napi_env env, napi_callback_info info) {
  napi_value a[5];
  if (!get_args(env, info, 5, a)) return nullptr;
  fx_sim* s = sim_of(env, a[0]);
  if (!s) return nullptr;
  fx_camera c;
  std::memset(&c, 0, sizeof c);
  vec3(env, a[1], "position", c.position);
  vec3(env, a[1], "target", c.target);
  vec3(env, a[1], "forward", c.forward);
  vec3(env, a[1], "right", c.right);
  vec3(env, a[1], "up", c.up);
  c.aspect = float(num(env, a[1], "aspect", 1.0));
  c.tan_half_fov_y = float(num(env, a[1], "tanHalfFovY", 0.4));
  fx_render_params p;
  std::memset(&p, 0, sizeof p);
  p.step_count = int32_t(num(env, a[2], "stepCount", 400));
  p.scattering_forward = float(num(env, a[2], "scatteringForward", 0.32));
  p.scattering_back = float(num(env, a[2], "scatteringBack", -0.18));
  p.display_mode = int32_t(num(env, a[2], "displayMode", FX_DISPLAY_TEMPERATURE));
  p.elapsed_s = float(num(env, a[2], "elapsedSeconds", 0.0));
  if (napi_value lights = child(env, a[2], "lights")) {
    uint32_t nl = 0;
    napi_get_array_length(env, lights, &nl);
    p.light_count = int32_t(nl < FX_MAX_RENDER_LIGHTS ? nl : FX_MAX_RENDER_LIGHTS);
    for (int32_t i = 0; i < p.light_count; ++i) {
      napi_value l;
      napi_get_element(env, lights, uint32_t(i), &l);
      p.lights[i].type = int32_t(num(env, l, "type", 0));
      vec3(env, l, "direction", p.lights[i].direction);
      vec3(env, l, "position", p.lights[i].position);
      vec3(env, l, "color", p.lights[i].color);
      p.lights[i].intensity = float(num(env, l, "intensity", 0));
    }
  }
  uint32_t w = 0, h = 0;
  NAPI_OK(napi_get_value_uint32(env, a[3], &w));
  NAPI_OK(napi_get_value_uint32(env, a[4], &h));
  void* data = nullptr;
  napi_value arr = typed(env, napi_uint16_array, size_t(w) * h * 4, 2, &data);
  if (!arr) return nullptr;
  const int rc = fx_raymarch(s, &c, &p, w, h, static_cast<uint16_t*>(data));
  if (rc != FX_OK) return throw_status(env, s, rc, "render");
  return arr;
}

napi_value js_last_error(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  napi_get_cb_info(env, info, &argc, a, nullptr, nullptr);
  fx_sim* s = nullptr;
  if (argc >= 1) {
    void* p = nullptr;
    if (napi_get_value_external(env, a[0], &p) == napi_ok) s = static_cast<fx_sim*>(p);
  }
  const char* msg = fx_last_error(s);
  napi_value v;
  NAPI_OK(napi_create_string_utf8(env, msg ? msg : "", NAPI_AUTO_LENGTH, &v));
  return v;
}

napi_value init(napi_env env, napi_value exports) {
  struct F { const char* name; napi_callback cb; } fns[] = {
      {"create", js_create}, {"destroy", js_destroy}, {"reset", js_reset}, {"synchronize", js_synchronize},
      {"step", js_step}, {"updateSources", js_update_sources}, {"getRuntimeParams", js_get_runtime},
      {"setRuntimeParams", js_set_runtime}, {"getQualitySettings", js_get_quality}, {"setQualitySettings", js_set_quality},
      {"getAdvectionMode", js_get_mode}, {"setAdvectionMode", js_set_mode}, {"readField", js_read_field},
      {"readFlags", js_read_flags}, {"activeBrickCount", js_active_bricks}, {"requestDebugFields", js_request_debug},
      {"raymarch", js_raymarch}, {"lastError", js_last_error},
  };
  for (const F& f : fns) {
    napi_value fn;
    NAPI_OK(napi_create_function(env, f.name, NAPI_AUTO_LENGTH, f.cb, nullptr, &fn));
    NAPI_OK(napi_set_named_property(env, exports, f.name, fn));
  }
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
