// N-API addon: binds the C ABI of include/cfd_mi355x.h for Node.js.
//
// Thin on purpose: argument unpacking, typed-array views, and rc < 0 turned
// into a thrown Error carrying cfd_last_error().  The JavaScript class in
// solver.js keeps the reference's method names on top of these functions.
#include <node_api.h>

#include <cstdint>
#include <cstdio>
#include <cstddef>
#include <cstring>
#include <string>

#include "cfd_mi355x.h"

namespace {

struct Handle {
  cfd_solver* s = nullptr;
};

#define NAPI_OK(call)                                               \
  do {                                                              \
    if ((call) != napi_ok) {                                        \
      napi_throw_error(env, nullptr, "cfdmi addon: N-API call failed"); \
      return nullptr;                                               \
    }                                                               \
  } while (0)

napi_value throw_rc(napi_env env, int rc) {
  std::string msg = cfd_last_error();
  if (msg.empty()) msg = "cfd error " + std::to_string(rc);
  napi_throw_error(env, nullptr, msg.c_str());
  return nullptr;
}

#define CFD_OK_OR_THROW(expr)            \
  do {                                   \
    int rc_ = (expr);                    \
    if (rc_ < 0) return throw_rc(env, rc_); \
  } while (0)

bool args(napi_env env, napi_callback_info info, size_t want, napi_value* argv) {
  size_t argc = want;
  if (napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr) != napi_ok) return false;
  if (argc < want) {
    napi_throw_type_error(env, nullptr, "cfdmi addon: missing arguments");
    return false;
  }
  return true;
}

Handle* handle(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, nullptr, "cfdmi addon: not a solver handle");
    return nullptr;
  }
  Handle* h = static_cast<Handle*>(p);
  if (!h->s) {
    napi_throw_error(env, nullptr, "cfdmi addon: solver already destroyed");
    return nullptr;
  }
  return h;
}

double num(napi_env env, napi_value v) {
  double d = 0;
  napi_get_value_double(env, v, &d);
  return d;
}

double prop_num(napi_env env, napi_value obj, const char* name, double dflt) {
  bool has = false;
  if (napi_has_named_property(env, obj, name, &has) != napi_ok || !has) return dflt;
  napi_value v;
  napi_get_named_property(env, obj, name, &v);
  napi_valuetype t;
  napi_typeof(env, v, &t);
  return t == napi_number ? num(env, v) : dflt;
}

void prop_vec3(napi_env env, napi_value obj, const char* name, double out[3]) {
  out[0] = out[1] = out[2] = 0.0;
  bool has = false;
  if (napi_has_named_property(env, obj, name, &has) != napi_ok || !has) return;
  napi_value arr;
  napi_get_named_property(env, obj, name, &arr);
  for (uint32_t k = 0; k < 3; ++k) {
    napi_value e;
    if (napi_get_element(env, arr, k, &e) == napi_ok) out[k] = num(env, e);
  }
}

// Typed-array view (data pointer already offset); returns false and throws
// when the value is not a typed array of the expected kind.
template <typename T>
bool view(napi_env env, napi_value v, napi_typedarray_type want, T** data, size_t* len) {
  napi_typedarray_type t;
  size_t n = 0, off = 0;
  void* d = nullptr;
  napi_value ab;
  bool is = false;
  napi_is_typedarray(env, v, &is);
  if (!is || napi_get_typedarray_info(env, v, &t, &n, &d, &ab, &off) != napi_ok || t != want) {
    napi_throw_type_error(env, nullptr, "cfdmi addon: wrong typed-array type");
    return false;
  }
  *data = static_cast<T*>(d);
  *len = n;
  return true;
}

template <typename T>
napi_value make_typed(napi_env env, napi_typedarray_type type, size_t n, T** data) {
  napi_value ab, out;
  void* d = nullptr;
  if (napi_create_arraybuffer(env, n * sizeof(T), &d, &ab) != napi_ok) return nullptr;
  if (napi_create_typedarray(env, type, n, ab, 0, &out) != napi_ok) return nullptr;
  *data = static_cast<T*>(d);
  return out;
}

napi_value number(napi_env env, double d) {
  napi_value v;
  napi_create_double(env, d, &v);
  return v;
}

void finalize(napi_env, void* data, void*) {
  Handle* h = static_cast<Handle*>(data);
  if (h->s) cfd_destroy(h->s);
  delete h;
}

// create(nx, ny, nz, fp16) -> handle
napi_value Create(napi_env env, napi_callback_info info) {
  napi_value a[4];
  if (!args(env, info, 4, a)) return nullptr;
  // a boolean (float16 or not) or the cfd_create flags as a number
  uint32_t flags = CFD_STORE_FP32;
  napi_valuetype t;
  NAPI_OK(napi_typeof(env, a[3], &t));
  if (t == napi_boolean) {
    bool fp16 = false;
    napi_get_value_bool(env, a[3], &fp16);
    flags = fp16 ? CFD_STORE_FP16 : CFD_STORE_FP32;
  } else {
    flags = (uint32_t)num(env, a[3]);
  }
  cfd_solver* s = nullptr;
  CFD_OK_OR_THROW(cfd_create((int)num(env, a[0]), (int)num(env, a[1]), (int)num(env, a[2]),
                             flags, &s));
  Handle* h = new Handle{s};
  napi_value ext;
  NAPI_OK(napi_create_external(env, h, finalize, nullptr, &ext));
  return ext;
}

napi_value Destroy(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!args(env, info, 1, a)) return nullptr;
  void* p = nullptr;
  napi_get_value_external(env, a[0], &p);
  Handle* h = static_cast<Handle*>(p);
  if (h && h->s) {
    cfd_destroy(h->s);
    h->s = nullptr;
  }
  return nullptr;
}

// configure(h, {omegaPlus, magic, lesCs, beltU, inletU, spongeStart, spongeLen, blThickness})
napi_value Configure(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!args(env, info, 2, a)) return nullptr;
  Handle* h = handle(env, a[0]);
  if (!h) return nullptr;
  cfd_params p{};
  p.omega_plus = prop_num(env, a[1], "omegaPlus", 0.0);
  p.magic = prop_num(env, a[1], "magic", 3.0 / 16.0);
  p.les_cs = prop_num(env, a[1], "lesCs", 0.1);
  prop_vec3(env, a[1], "beltU", p.belt);
  prop_vec3(env, a[1], "inletU", p.inlet);
  p.sponge_start = prop_num(env, a[1], "spongeStart", 0.0);
  p.sponge_len = prop_num(env, a[1], "spongeLen", 1.0);
  p.bl_thickness = prop_num(env, a[1], "blThickness", 0.0);
  CFD_OK_OR_THROW(cfd_configure(h->s, &p));
  return nullptr;
}

napi_value SetTypes(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!args(env, info, 2, a)) return nullptr;
  Handle* h = handle(env, a[0]);
  if (!h) return nullptr;
  uint32_t* t = nullptr;
  size_t n = 0;
  if (!view(env, a[1], napi_uint32_array, &t, &n)) return nullptr;
  CFD_OK_OR_THROW(cfd_set_types(h->s, t, n));
  return nullptr;
}

napi_value InitEquilibrium(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!args(env, info, 1, a)) return nullptr;
  Handle* h = handle(env, a[0]);
  if (!h) return nullptr;
  CFD_OK_OR_THROW(cfd_init_equilibrium(h->s));
  return nullptr;
}

napi_value SetPopulations(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!args(env, info, 2, a)) return nullptr;
  Handle* h = handle(env, a[0]);
  if (!h) return nullptr;
  float* d = nullptr;
  size_t n = 0;
  if (!view(env, a[1], napi_float32_array, &d, &n)) return nullptr;
  CFD_OK_OR_THROW(cfd_set_populations(h->s, d, n));
  return nullptr;
}

// getPopulations(h, n) -> Float32Array(n) with n = 19 N
napi_value GetPopulations(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!args(env, info, 2, a)) return nullptr;
  Handle* h = handle(env, a[0]);
  if (!h) return nullptr;
  const size_t n = (size_t)num(env, a[1]);
  float* d = nullptr;
  napi_value out = make_typed(env, napi_float32_array, n, &d);
  if (!out) return nullptr;
  CFD_OK_OR_THROW(cfd_get_populations(h->s, d, n));
  return out;
}

napi_value SetFront(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!args(env, info, 2, a)) return nullptr;
  Handle* h = handle(env, a[0]);
  if (!h) return nullptr;
  CFD_OK_OR_THROW(cfd_set_front(h->s, (int)num(env, a[1])));
  return nullptr;
}

napi_value GetFront(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!args(env, info, 1, a)) return nullptr;
  Handle* h = handle(env, a[0]);
  if (!h) return nullptr;
  int f = 0;
  CFD_OK_OR_THROW(cfd_get_front(h->s, &f));
  return number(env, f);
}

napi_value Step(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!args(env, info, 2, a)) return nullptr;
  Handle* h = handle(env, a[0]);
  if (!h) return nullptr;
  CFD_OK_OR_THROW(cfd_step(h->s, (int)num(env, a[1])));
  return nullptr;
}

napi_value StepTimed(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!args(env, info, 2, a)) return nullptr;
  Handle* h = handle(env, a[0]);
  if (!h) return nullptr;
  double ms = 0;
  CFD_OK_OR_THROW(cfd_step_timed(h->s, (int)num(env, a[1]), &ms));
  return number(env, ms);
}

napi_value Forces(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!args(env, info, 1, a)) return nullptr;
  Handle* h = handle(env, a[0]);
  if (!h) return nullptr;
  double* d = nullptr;
  napi_value out = make_typed(env, napi_float64_array, 3, &d);
  if (!out) return nullptr;
  CFD_OK_OR_THROW(cfd_forces(h->s, d));
  return out;
}

napi_value UpdateMacros(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!args(env, info, 1, a)) return nullptr;
  Handle* h = handle(env, a[0]);
  if (!h) return nullptr;
  CFD_OK_OR_THROW(cfd_update_macros(h->s));
  return nullptr;
}

// readMacros(h, n) -> Float32Array(n) with n = 4 N
napi_value ReadMacros(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!args(env, info, 2, a)) return nullptr;
  Handle* h = handle(env, a[0]);
  if (!h) return nullptr;
  const size_t n = (size_t)num(env, a[1]);
  float* d = nullptr;
  napi_value out = make_typed(env, napi_float32_array, n, &d);
  if (!out) return nullptr;
  CFD_OK_OR_THROW(cfd_read_macros(h->s, d, n));
  return out;
}

napi_value Mass(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!args(env, info, 1, a)) return nullptr;
  Handle* h = handle(env, a[0]);
  if (!h) return nullptr;
  double m = 0;
  int64_t bad = 0;
  CFD_OK_OR_THROW(cfd_mass(h->s, &m, &bad));
  napi_value out;
  NAPI_OK(napi_create_array_with_length(env, 2, &out));
  napi_set_element(env, out, 0, number(env, m));
  napi_set_element(env, out, 1, number(env, (double)bad));
  return out;
}

napi_value Synchronize(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!args(env, info, 1, a)) return nullptr;
  Handle* h = handle(env, a[0]);
  if (!h) return nullptr;
  CFD_OK_OR_THROW(cfd_synchronize(h->s));
  return nullptr;
}

napi_value Bytes(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!args(env, info, 1, a)) return nullptr;
  Handle* h = handle(env, a[0]);
  if (!h) return nullptr;
  int64_t b = 0;
  CFD_OK_OR_THROW(cfd_bytes(h->s, &b));
  return number(env, (double)b);
}

// voxelize(pos Float32Array, idx Uint32Array, nx, ny, nz, band, seal, withSdf)
napi_value Voxelize(napi_env env, napi_callback_info info) {
  napi_value a[8];
  if (!args(env, info, 8, a)) return nullptr;
  float* pos = nullptr;
  uint32_t* idx = nullptr;
  size_t np = 0, ni = 0;
  if (!view(env, a[0], napi_float32_array, &pos, &np)) return nullptr;
  if (!view(env, a[1], napi_uint32_array, &idx, &ni)) return nullptr;
  const int nx = (int)num(env, a[2]), ny = (int)num(env, a[3]), nz = (int)num(env, a[4]);
  const double band = num(env, a[5]);
  bool seal = true, with_sdf = true;
  napi_get_value_bool(env, a[6], &seal);
  napi_get_value_bool(env, a[7], &with_sdf);
  if (nx < 1 || ny < 1 || nz < 1) {
    napi_throw_range_error(env, nullptr, "voxelize: grid must be positive");
    return nullptr;
  }
  const size_t N = (size_t)nx * ny * nz;
  uint8_t* tipo = nullptr;
  float* sdf = nullptr;
  napi_value jt = make_typed(env, napi_uint8_array, N, &tipo);
  napi_value js = nullptr;
  if (with_sdf) js = make_typed(env, napi_float32_array, N, &sdf);
  if (!jt || (with_sdf && !js)) return nullptr;
  cfd_vox_stats st{};
  CFD_OK_OR_THROW(cfd_voxelize(pos, np / 3, idx, ni / 3, nx, ny, nz, band, seal ? 1 : 0, tipo,
                               sdf, &st));
  napi_value out, stats, nul;
  NAPI_OK(napi_create_object(env, &out));
  NAPI_OK(napi_create_object(env, &stats));
  napi_get_null(env, &nul);
  napi_set_named_property(env, out, "tipo", jt);
  napi_set_named_property(env, out, "sdf", with_sdf ? js : nul);
  const struct {
    const char* k;
    double v;
  } f[] = {{"nCelulas", (double)st.n_cells},      {"nCasca", (double)st.n_casca},
           {"nDentro", (double)st.n_dentro},      {"nSolido", (double)st.n_solido},
           {"trianguloFora", (double)st.triangles_outside},
           {"iteracoesInundacao", (double)st.flood_iterations},
           {"msRasterizacao", st.ms_raster},      {"msInundacao", st.ms_flood},
           {"msTotal", st.ms_total}};
  for (const auto& e : f) napi_set_named_property(env, stats, e.k, number(env, e.v));
  napi_value oco, frac;
  napi_get_boolean(env, st.hollow_suspect != 0, &oco);
  napi_set_named_property(env, stats, "ocoSuspeito", oco);
  frac = number(env, (double)st.n_solido / (double)N);
  napi_set_named_property(env, stats, "fracaoSolida", frac);
  napi_set_named_property(env, out, "stats", stats);
  return out;
}

napi_value SetTuning(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!args(env, info, 2, a)) return nullptr;
  Handle* h = handle(env, a[0]);
  if (!h) return nullptr;
  cfd_tuning t{};
  CFD_OK_OR_THROW(cfd_get_tuning(h->s, &t));
  t.block = (int32_t)prop_num(env, a[1], "block", t.block);
  t.remap = (int32_t)prop_num(env, a[1], "remap", t.remap);
  t.layout = (int32_t)prop_num(env, a[1], "layout", t.layout);
  t.slice_pad = (int32_t)prop_num(env, a[1], "slicePad", t.slice_pad);
  t.sweep = (int32_t)prop_num(env, a[1], "sweep", t.sweep);
  t.rotate = (int32_t)prop_num(env, a[1], "rotate", t.rotate);
  CFD_OK_OR_THROW(cfd_set_tuning(h->s, &t));
  return nullptr;
}

// getTuning(h) -> {block, remap, layout, slicePad, sweep, rotate}
napi_value GetTuning(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!args(env, info, 1, a)) return nullptr;
  Handle* h = handle(env, a[0]);
  if (!h) return nullptr;
  cfd_tuning t{};
  CFD_OK_OR_THROW(cfd_get_tuning(h->s, &t));
  napi_value out;
  NAPI_OK(napi_create_object(env, &out));
  const struct { const char* k; int32_t v; } f[] = {
      {"block", t.block}, {"remap", t.remap}, {"layout", t.layout},
      {"slicePad", t.slice_pad}, {"sweep", t.sweep}, {"rotate", t.rotate}};
  for (const auto& e : f) napi_set_named_property(env, out, e.k, number(env, e.v));
  return out;
}

// setClosure(h, Float32Array(1411) | null): the cfd_closure fields flat, in
// their order (the 1377 MLP parameters, the activation code as a number,
// in_shift, in_scale, in_lo, in_hi, out_scale); null returns to Smagorinsky.
napi_value SetClosure(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!args(env, info, 2, a)) return nullptr;
  Handle* h = handle(env, a[0]);
  if (!h) return nullptr;
  napi_valuetype t;
  NAPI_OK(napi_typeof(env, a[1], &t));
  if (t == napi_null || t == napi_undefined) {
    CFD_OK_OR_THROW(cfd_set_closure(h->s, nullptr));
    return nullptr;
  }
  float* d = nullptr;
  size_t n = 0;
  if (!view(env, a[1], napi_float32_array, &d, &n)) return nullptr;
  constexpr size_t kParams = offsetof(cfd_mlp, activation) / sizeof(float);
  static_assert(kParams == 1377, "cfd_mlp: 1377 parameters before the activation");
  if (n != kParams + 1 + 33) {
    napi_throw_range_error(env, nullptr, "cfdmi addon: a closure is 1411 float32 values");
    return nullptr;
  }
  if (d[kParams] != 0.0f && d[kParams] != 1.0f) {
    napi_throw_range_error(env, nullptr, "cfdmi addon: the activation code is 0 or 1");
    return nullptr;
  }
  cfd_closure c;
  std::memcpy(&c.mlp, d, kParams * sizeof(float));
  c.mlp.activation = (int32_t)d[kParams];
  const float* r = d + kParams + 1;
  std::memcpy(c.in_shift, r, sizeof c.in_shift);
  std::memcpy(c.in_scale, r + 8, sizeof c.in_scale);
  std::memcpy(c.in_lo, r + 16, sizeof c.in_lo);
  std::memcpy(c.in_hi, r + 24, sizeof c.in_hi);
  c.out_scale = r[32];
  CFD_OK_OR_THROW(cfd_set_closure(h->s, &c));
  return nullptr;
}

napi_value SetWallDistance(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!args(env, info, 2, a)) return nullptr;
  Handle* h = handle(env, a[0]);
  if (!h) return nullptr;
  float* d = nullptr;
  size_t n = 0;
  if (!view(env, a[1], napi_float32_array, &d, &n)) return nullptr;
  CFD_OK_OR_THROW(cfd_set_wall_distance(h->s, d, n));
  return nullptr;
}

napi_value ComputeWallDistance(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!args(env, info, 1, a)) return nullptr;
  Handle* h = handle(env, a[0]);
  if (!h) return nullptr;
  CFD_OK_OR_THROW(cfd_compute_wall_distance(h->s));
  return nullptr;
}

napi_value ClosureFallbacks(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!args(env, info, 2, a)) return nullptr;
  Handle* h = handle(env, a[0]);
  if (!h) return nullptr;
  bool reset = false;
  NAPI_OK(napi_get_value_bool(env, a[1], &reset));
  int64_t n = 0;
  CFD_OK_OR_THROW(cfd_closure_fallbacks(h->s, &n, reset ? 1 : 0));
  return number(env, (double)n);
}

napi_value Version(napi_env env, napi_callback_info) {
  napi_value v;
  napi_create_string_utf8(env, cfd_version(), NAPI_AUTO_LENGTH, &v);
  return v;
}

}  // namespace

NAPI_MODULE_INIT() {
  // the structs bound here follow this header's layout version
  if (cfd_abi_version() != CFD_ABI_VERSION) {
    char msg[160];
    snprintf(msg, sizeof msg,
             "libcfdmi has struct layout version %d; this addon was built for %d: rebuild",
             cfd_abi_version(), CFD_ABI_VERSION);
    napi_throw_error(env, nullptr, msg);
    return nullptr;
  }
  const struct {
    const char* name;
    napi_callback cb;
  } fns[] = {
      {"create", Create},           {"destroy", Destroy},
      {"configure", Configure},     {"setTypes", SetTypes},
      {"initEquilibrium", InitEquilibrium},
      {"setPopulations", SetPopulations},
      {"getPopulations", GetPopulations},
      {"setFront", SetFront},       {"getFront", GetFront},
      {"step", Step},               {"stepTimed", StepTimed},
      {"forces", Forces},           {"updateMacros", UpdateMacros},
      {"readMacros", ReadMacros},   {"mass", Mass},
      {"synchronize", Synchronize}, {"bytes", Bytes},
      {"voxelize", Voxelize},       {"setTuning", SetTuning},
      {"getTuning", GetTuning},
      {"setClosure", SetClosure},   {"setWallDistance", SetWallDistance},
      {"computeWallDistance", ComputeWallDistance},
      {"closureFallbacks", ClosureFallbacks},
      {"version", Version},
  };
  for (const auto& f : fns) {
    napi_value fn;
    if (napi_create_function(env, f.name, NAPI_AUTO_LENGTH, f.cb, nullptr, &fn) != napi_ok)
      return nullptr;
    napi_set_named_property(env, exports, f.name, fn);
  }
  return exports;
}
