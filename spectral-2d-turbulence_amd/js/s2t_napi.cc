// N-API addon over libs2t.so (include/s2t.h). Thin: every function converts JS values, calls one
// C entry point and throws a JS Error carrying s2t_last_error() on a negative status (the
// reference throws Error for programming errors, src/gpu/fft.js:28,84). Long device work that
// returns data (diagnostics) is also offered as a Promise on a worker thread, like the reference's
// async readDiagnostics (src/gpu/solver.js:570-608). The serving loop reports each request's final
// status from its native worker thread through a thread-safe function.
#include <node_api.h>

#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "s2t.h"

namespace {

#define NAPI_CALL(env, call)                                        \
  do {                                                              \
    if ((call) != napi_ok) {                                        \
      napi_throw_error((env), "S2T_NAPI", "N-API call failed: " #call); \
      return nullptr;                                               \
    }                                                               \
  } while (0)

napi_value throw_status(napi_env env, int rc) {
  const std::string code = "S2T_" + std::to_string(rc);
  napi_throw_error(env, code.c_str(), s2t_last_error());
  return nullptr;
}

#define S2T_CALL(env, call)             \
  do {                                  \
    int rc_ = (call);                   \
    if (rc_ != S2T_OK) return throw_status((env), rc_); \
  } while (0)

struct Args {
  size_t argc = 8;
  napi_value argv[8];
  napi_value self;
};

bool get_args(napi_env env, napi_callback_info info, Args& a, size_t need) {
  if (napi_get_cb_info(env, info, &a.argc, a.argv, &a.self, nullptr) != napi_ok) return false;
  if (a.argc < need) {
    napi_throw_type_error(env, "S2T_ARGS", "missing arguments");
    return false;
  }
  return true;
}

s2t_solver* handle(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, "S2T_HANDLE", "not a solver handle (destroyed?)");
    return nullptr;
  }
  return static_cast<s2t_solver*>(p);
}

double num(napi_env env, napi_value v, double dflt) {
  napi_valuetype t;
  if (napi_typeof(env, v, &t) != napi_ok || t != napi_number) return dflt;
  double d = dflt;
  napi_get_value_double(env, v, &d);
  return d;
}

double prop(napi_env env, napi_value obj, const char* key, double dflt) {
  bool has = false;
  if (napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return dflt;
  napi_value v;
  if (napi_get_named_property(env, obj, key, &v) != napi_ok) return dflt;
  return num(env, v, dflt);
}

// typed-array view (data pointer and element count) with a type check
template <class T>
T* typed(napi_env env, napi_value v, napi_typedarray_type want, size_t* count) {
  bool is = false;
  napi_is_typedarray(env, v, &is);
  if (!is) {
    napi_throw_type_error(env, "S2T_ARGS", "expected a typed array");
    return nullptr;
  }
  napi_typedarray_type t;
  size_t len = 0;
  void* data = nullptr;
  napi_value ab;
  size_t off = 0;
  napi_get_typedarray_info(env, v, &t, &len, &data, &ab, &off);
  if (t != want) {
    napi_throw_type_error(env, "S2T_ARGS", "typed array of the wrong element type");
    return nullptr;
  }
  *count = len;
  return static_cast<T*>(data);
}

napi_value new_typed(napi_env env, napi_typedarray_type t, size_t count, size_t elem, void** data) {
  napi_value ab, arr;
  if (napi_create_arraybuffer(env, count * elem, data, &ab) != napi_ok) return nullptr;
  if (napi_create_typedarray(env, t, count, ab, 0, &arr) != napi_ok) return nullptr;
  return arr;
}

napi_value make_num(napi_env env, double d) {
  napi_value v;
  napi_create_double(env, d, &v);
  return v;
}

void finalize_solver(napi_env, void* data, void*) { s2t_destroy(static_cast<s2t_solver*>(data)); }

// ---------------------------------------------------------------------------------------------

napi_value Create(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  const int n = (int)num(env, a.argv[0], 256), dye = (int)num(env, a.argv[1], 2);
  s2t_solver* h = nullptr;
  S2T_CALL(env, s2t_create(n, dye, &h));
  napi_value ext;
  NAPI_CALL(env, napi_create_external(env, h, finalize_solver, nullptr, &ext));
  return ext;
}

napi_value RadixSchedule(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 1)) return nullptr;
  int st = 0, r2 = 0;
  S2T_CALL(env, s2t_radix_schedule((int)num(env, a.argv[0], 0), &st, &r2));
  napi_value obj, b;
  NAPI_CALL(env, napi_create_object(env, &obj));
  NAPI_CALL(env, napi_set_named_property(env, obj, "stages", make_num(env, st)));
  NAPI_CALL(env, napi_get_boolean(env, r2 != 0, &b));
  NAPI_CALL(env, napi_set_named_property(env, obj, "hasRadix2", b));
  return obj;
}

// the reference config keys (src/main.js:9-36, solver.js:306-324) with their defaults
s2t_params params_of(napi_env env, napi_value c) {
  s2t_params p{};
  p.beta = prop(env, c, "beta", 0);
  p.drag = prop(env, c, "drag", 0.01);
  p.viscosity = prop(env, c, "viscosity", 0.67);
  p.hyper_p = prop(env, c, "hyperP", 4);
  p.force_amp = prop(env, c, "forceAmp", 0);
  p.force_k = prop(env, c, "forceK", 22);
  p.force_width = prop(env, c, "forceWidth", 2.5);
  p.cfl = prop(env, c, "cfl", 0.5);
  p.dt_min = prop(env, c, "dtMin", 1e-5);
  p.dt_max = prop(env, c, "dtMax", 0.02);
  p.dye_fade = prop(env, c, "dyeFade", 0.999);
  p.fixed_dt = prop(env, c, "fixedDt", 0);
  return p;
}

// configure(h, cfg): the reference config keys (src/main.js:9-36, solver.js:306-324)
napi_value Configure(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  s2t_solver* h = handle(env, a.argv[0]);
  if (!h) return nullptr;
  const s2t_params p = params_of(env, a.argv[1]);
  S2T_CALL(env, s2t_configure(h, &p));
  return nullptr;
}

napi_value NuH(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 1)) return nullptr;
  s2t_solver* h = handle(env, a.argv[0]);
  if (!h) return nullptr;
  double v = 0;
  S2T_CALL(env, s2t_nu_h(h, &v));
  return make_num(env, v);
}

// setSplats(h, list): items {x, y, dx, dy, color: [r, g, b], radius, amp, dyeAmp}
napi_value SetSplats(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  s2t_solver* h = handle(env, a.argv[0]);
  if (!h) return nullptr;
  uint32_t len = 0;
  NAPI_CALL(env, napi_get_array_length(env, a.argv[1], &len));
  std::vector<s2t_splat> sp(len ? len : 1);
  for (uint32_t i = 0; i < len; ++i) {
    napi_value o;
    NAPI_CALL(env, napi_get_element(env, a.argv[1], i, &o));
    s2t_splat& s = sp[i];
    s.x = (float)prop(env, o, "x", 0);
    s.y = (float)prop(env, o, "y", 0);
    s.dx = (float)prop(env, o, "dx", 0);
    s.dy = (float)prop(env, o, "dy", 0);
    s.radius = (float)prop(env, o, "radius", 0.045);
    s.amp = (float)prop(env, o, "amp", 0);
    s.dye_amp = (float)prop(env, o, "dyeAmp", 0);
    bool has = false;
    napi_has_named_property(env, o, "color", &has);
    if (has) {
      napi_value col;
      NAPI_CALL(env, napi_get_named_property(env, o, "color", &col));
      for (uint32_t k = 0; k < 3; ++k) {
        napi_value e;
        if (napi_get_element(env, col, k, &e) == napi_ok) s.color[k] = (float)num(env, e, 0);
      }
    }
  }
  S2T_CALL(env, s2t_set_splats(h, sp.data(), (int)len));
  return nullptr;
}

napi_value SeedSpectrum(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 4)) return nullptr;
  s2t_solver* h = handle(env, a.argv[0]);
  if (!h) return nullptr;
  const double seed = num(env, a.argv[3], 1);
  S2T_CALL(env, s2t_seed_spectrum(h, num(env, a.argv[1], 14), num(env, a.argv[2], 6),
                                  (uint32_t)(int64_t)seed));
  return nullptr;
}

napi_value SeedPhysical(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 3)) return nullptr;
  s2t_solver* h = handle(env, a.argv[0]);
  if (!h) return nullptr;
  const int n = (int)num(env, a.argv[2], 0);
  size_t cnt = 0;
  float* f = typed<float>(env, a.argv[1], napi_float32_array, &cnt);
  if (!f) return nullptr;
  if (cnt != (size_t)n * n) {
    napi_throw_range_error(env, "S2T_ARGS", "field must hold n*n values");
    return nullptr;
  }
  S2T_CALL(env, s2t_seed_physical(h, f));
  return nullptr;
}

napi_value WriteState(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 3)) return nullptr;
  s2t_solver* h = handle(env, a.argv[0]);
  if (!h) return nullptr;
  const int n = (int)num(env, a.argv[2], 0);
  size_t cnt = 0;
  float* f = typed<float>(env, a.argv[1], napi_float32_array, &cnt);
  if (!f) return nullptr;
  if (cnt != 2 * (size_t)n * n) {
    napi_throw_range_error(env, "S2T_ARGS", "spectrum must hold 2*n*n values");
    return nullptr;
  }
  S2T_CALL(env, s2t_write_state(h, f));
  return nullptr;
}

napi_value ReadState(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  s2t_solver* h = handle(env, a.argv[0]);
  if (!h) return nullptr;
  const int n = (int)num(env, a.argv[1], 0);
  void* data = nullptr;
  napi_value arr = new_typed(env, napi_float32_array, 2 * (size_t)n * n, 4, &data);
  if (!arr) return nullptr;
  S2T_CALL(env, s2t_read_state(h, static_cast<float*>(data)));
  return arr;
}

napi_value ResetClock(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 1)) return nullptr;
  s2t_solver* h = handle(env, a.argv[0]);
  if (!h) return nullptr;
  S2T_CALL(env, s2t_reset_clock(h));
  return nullptr;
}

napi_value Step(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 3)) return nullptr;
  s2t_solver* h = handle(env, a.argv[0]);
  if (!h) return nullptr;
  bool dye = false;
  napi_get_value_bool(env, a.argv[2], &dye);
  S2T_CALL(env, s2t_step(h, (int)num(env, a.argv[1], 1), dye ? S2T_STEP_DYE : 0u));
  return nullptr;
}

napi_value Sync(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 1)) return nullptr;
  s2t_solver* h = handle(env, a.argv[0]);
  if (!h) return nullptr;
  S2T_CALL(env, s2t_sync(h));
  return nullptr;
}

// diagnostics record (solver.js:591-602) in FP64
struct DiagJob {
  s2t_solver* h = nullptr;
  int n = 0;
  int rc = 0;
  std::string err;
  s2t_diag d{};
  std::vector<double> spectrum, zonal;
  napi_deferred deferred = nullptr;
  napi_async_work work = nullptr;
};

napi_value diag_object(napi_env env, DiagJob& j) {
  napi_value obj;
  napi_create_object(env, &obj);
  void* data = nullptr;
  napi_value spec = new_typed(env, napi_float64_array, j.spectrum.size(), 8, &data);
  memcpy(data, j.spectrum.data(), j.spectrum.size() * 8);
  napi_value zon = new_typed(env, napi_float64_array, j.zonal.size(), 8, &data);
  memcpy(data, j.zonal.data(), j.zonal.size() * 8);
  napi_set_named_property(env, obj, "spectrum", spec);
  napi_set_named_property(env, obj, "zonal", zon);
  napi_set_named_property(env, obj, "energy", make_num(env, j.d.energy));
  napi_set_named_property(env, obj, "enstrophy", make_num(env, j.d.enstrophy));
  napi_set_named_property(env, obj, "dt", make_num(env, j.d.dt));
  napi_set_named_property(env, obj, "maxSpeed", make_num(env, j.d.max_speed));
  napi_set_named_property(env, obj, "time", make_num(env, j.d.time));
  napi_set_named_property(env, obj, "stepIndex", make_num(env, j.d.step_index));
  napi_value nf;
  napi_get_boolean(env, j.d.nonfinite != 0, &nf);
  napi_set_named_property(env, obj, "nonfinite", nf);
  return obj;
}

void run_diag(DiagJob& j) {
  j.spectrum.assign(j.n / 2 + 1, 0.0);
  j.zonal.assign(j.n, 0.0);
  j.rc = s2t_diagnostics(j.h, &j.d, j.spectrum.data(), j.zonal.data());
  if (j.rc != S2T_OK) j.err = s2t_last_error();
}

napi_value Diagnostics(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  DiagJob j;
  j.h = handle(env, a.argv[0]);
  if (!j.h) return nullptr;
  j.n = (int)num(env, a.argv[1], 0);
  run_diag(j);
  if (j.rc != S2T_OK) return throw_status(env, j.rc);
  return diag_object(env, j);
}

void diag_execute(napi_env, void* data) { run_diag(*static_cast<DiagJob*>(data)); }

void diag_complete(napi_env env, napi_status, void* data) {
  DiagJob* j = static_cast<DiagJob*>(data);
  if (j->rc == S2T_OK) {
    napi_resolve_deferred(env, j->deferred, diag_object(env, *j));
  } else {
    napi_value msg, err;
    napi_create_string_utf8(env, j->err.c_str(), NAPI_AUTO_LENGTH, &msg);
    napi_create_error(env, nullptr, msg, &err);
    napi_reject_deferred(env, j->deferred, err);
  }
  napi_delete_async_work(env, j->work);
  delete j;
}

// diagnosticsAsync(h, n, keepAlive) -> Promise; keepAlive is held by the caller for the duration
napi_value DiagnosticsAsync(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  DiagJob* j = new DiagJob();
  j->h = handle(env, a.argv[0]);
  if (!j->h) {
    delete j;
    return nullptr;
  }
  j->n = (int)num(env, a.argv[1], 0);
  napi_value promise, name;
  NAPI_CALL(env, napi_create_promise(env, &j->deferred, &promise));
  NAPI_CALL(env, napi_create_string_utf8(env, "s2t_diagnostics", NAPI_AUTO_LENGTH, &name));
  NAPI_CALL(env, napi_create_async_work(env, nullptr, name, diag_execute, diag_complete, j,
                                        &j->work));
  NAPI_CALL(env, napi_queue_async_work(env, j->work));
  return promise;
}

napi_value PaintDye(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  s2t_solver* h = handle(env, a.argv[0]);
  if (!h) return nullptr;
  S2T_CALL(env, s2t_paint_dye(h, (int)num(env, a.argv[1], 1)));
  return nullptr;
}

napi_value ReadDye(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 1)) return nullptr;
  s2t_solver* h = handle(env, a.argv[0]);
  if (!h) return nullptr;
  int m = 0;
  S2T_CALL(env, s2t_dye_size(h, &m));
  void* data = nullptr;
  napi_value arr = new_typed(env, napi_uint16_array, (size_t)m * m * 4, 2, &data);
  if (!arr) return nullptr;
  S2T_CALL(env, s2t_read_dye(h, static_cast<uint16_t*>(data)));
  return arr;
}

// hooks on host fields: fft2(h, Float32Array, n, inverse)
napi_value Fft2(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 4)) return nullptr;
  s2t_solver* h = handle(env, a.argv[0]);
  if (!h) return nullptr;
  size_t cnt = 0;
  float* in = typed<float>(env, a.argv[1], napi_float32_array, &cnt);
  if (!in) return nullptr;
  const int n = (int)num(env, a.argv[2], 0);
  if (cnt != 2 * (size_t)n * n) {
    napi_throw_range_error(env, "S2T_ARGS", "field must hold 2*n*n values");
    return nullptr;
  }
  bool inv = false;
  napi_get_value_bool(env, a.argv[3], &inv);
  void* data = nullptr;
  napi_value arr = new_typed(env, napi_float32_array, cnt, 4, &data);
  if (!arr) return nullptr;
  S2T_CALL(env, s2t_fft2(h, in, static_cast<float*>(data), inv ? 1 : 0));
  return arr;
}

napi_value VelocityPack(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 3)) return nullptr;
  s2t_solver* h = handle(env, a.argv[0]);
  if (!h) return nullptr;
  size_t cnt = 0;
  float* w = typed<float>(env, a.argv[1], napi_float32_array, &cnt);
  if (!w) return nullptr;
  const int n = (int)num(env, a.argv[2], 0);
  if (cnt != 2 * (size_t)n * n) {
    napi_throw_range_error(env, "S2T_ARGS", "spectrum must hold 2*n*n values");
    return nullptr;
  }
  void *da = nullptr, *db = nullptr;
  napi_value pa = new_typed(env, napi_float32_array, cnt, 4, &da);
  napi_value pb = new_typed(env, napi_float32_array, cnt, 4, &db);
  if (!pa || !pb) return nullptr;
  S2T_CALL(env, s2t_velocity_pack(h, w, static_cast<float*>(da), static_cast<float*>(db)));
  napi_value out;
  NAPI_CALL(env, napi_create_array_with_length(env, 2, &out));
  NAPI_CALL(env, napi_set_element(env, out, 0, pa));
  NAPI_CALL(env, napi_set_element(env, out, 1, pb));
  return out;
}

// assemble(h, z, fhat | null, n, stepIndex, dt)
napi_value Assemble(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 6)) return nullptr;
  s2t_solver* h = handle(env, a.argv[0]);
  if (!h) return nullptr;
  size_t cnt = 0, cf = 0;
  float* z = typed<float>(env, a.argv[1], napi_float32_array, &cnt);
  if (!z) return nullptr;
  napi_valuetype t;
  napi_typeof(env, a.argv[2], &t);
  float* f = nullptr;
  if (t != napi_null && t != napi_undefined) {
    f = typed<float>(env, a.argv[2], napi_float32_array, &cf);
    if (!f) return nullptr;
  }
  const int n = (int)num(env, a.argv[3], 0);
  if (cnt != 2 * (size_t)n * n || (f && cf != cnt)) {
    napi_throw_range_error(env, "S2T_ARGS", "fields must hold 2*n*n values");
    return nullptr;
  }
  void* data = nullptr;
  napi_value arr = new_typed(env, napi_float32_array, cnt, 4, &data);
  if (!arr) return nullptr;
  S2T_CALL(env, s2t_assemble(h, z, f, (uint32_t)num(env, a.argv[4], 0), num(env, a.argv[5], 0.01),
                             static_cast<float*>(data)));
  return arr;
}

napi_value Rhs(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 4)) return nullptr;
  s2t_solver* h = handle(env, a.argv[0]);
  if (!h) return nullptr;
  const int n = (int)num(env, a.argv[1], 0);
  void* data = nullptr;
  napi_value arr = new_typed(env, napi_float32_array, 2 * (size_t)n * n, 4, &data);
  if (!arr) return nullptr;
  S2T_CALL(env, s2t_rhs(h, (uint32_t)num(env, a.argv[2], 0), num(env, a.argv[3], 0.01),
                        static_cast<float*>(data)));
  return arr;
}

// ---------------------------------------------------------------------------------------------
// serving loop (s2t_server_*)
// ---------------------------------------------------------------------------------------------
struct ServerBox {
  s2t_server* sv = nullptr;
  napi_threadsafe_function tsfn = nullptr;
  // grid size of each outstanding request submitted with keepState; recursive because a
  // rejection calls back on the submitting thread inside s2t_server_submit
  std::recursive_mutex mu;
  std::unordered_map<uint64_t, int> keep;
};

// a request's final status with its result, read on the thread that reports it, so that the
// result outlives a destroy() that comes before the JS thread runs the callback
struct Final {
  uint64_t id;
  int status;
  s2t_diag diag{};
  double latency = -1;
  std::vector<float> state;
};

// s2t_server_callback: runs on the server's worker thread (or on the submitting / cancelling
// thread when the status is final at that call) without the server's lock, so it may poll
void server_final(uint64_t id, int status, void* user) {
  ServerBox* b = static_cast<ServerBox*>(user);
  Final* f = new Final{id, status};
  int n = 0;
  {
    std::lock_guard<std::recursive_mutex> lk(b->mu);
    auto it = b->keep.find(id);
    if (it != b->keep.end()) {
      n = it->second;
      b->keep.erase(it);
    }
  }
  if (status == S2T_REQ_DONE) {
    int st = 0;
    s2t_server_poll(b->sv, id, &st, &f->diag, &f->latency);
    if (n > 0) {
      f->state.resize((size_t)2 * n * n);
      if (s2t_server_read_state(b->sv, id, f->state.data()) != S2T_OK) f->state.clear();
    }
  }
  if (napi_call_threadsafe_function(b->tsfn, f, napi_tsfn_nonblocking) != napi_ok) delete f;
}

napi_value diag_object(napi_env env, int status, const s2t_diag& d, double latency) {
  napi_value obj;
  napi_create_object(env, &obj);
  napi_set_named_property(env, obj, "status", make_num(env, status));
  napi_set_named_property(env, obj, "latency", make_num(env, latency));
  napi_set_named_property(env, obj, "energy", make_num(env, d.energy));
  napi_set_named_property(env, obj, "enstrophy", make_num(env, d.enstrophy));
  napi_set_named_property(env, obj, "dt", make_num(env, d.dt));
  napi_set_named_property(env, obj, "maxSpeed", make_num(env, d.max_speed));
  napi_set_named_property(env, obj, "time", make_num(env, d.time));
  napi_set_named_property(env, obj, "stepIndex", make_num(env, d.step_index));
  return obj;
}

// calls onFinal(id, status, result): result is the diagnostics object (with `state`, a
// Float32Array in fixture layout, for keepState) when done, else undefined
void final_to_js(napi_env env, napi_value js_cb, void*, void* data) {
  Final* f = static_cast<Final*>(data);
  if (env && js_cb) {
    napi_value args[3], undef;
    napi_get_undefined(env, &undef);
    napi_create_double(env, (double)f->id, &args[0]);
    napi_create_int32(env, f->status, &args[1]);
    args[2] = undef;
    if (f->status == S2T_REQ_DONE) {
      args[2] = diag_object(env, f->status, f->diag, f->latency);
      if (!f->state.empty()) {
        void* p = nullptr;
        napi_value arr = new_typed(env, napi_float32_array, f->state.size(), 4, &p);
        if (arr) {
          memcpy(p, f->state.data(), f->state.size() * sizeof(float));
          napi_set_named_property(env, args[2], "state", arr);
        }
      }
    }
    napi_call_function(env, undef, js_cb, 3, args, nullptr);
  }
  delete f;
}

void server_close(ServerBox* b) {
  if (b->sv) {
    s2t_server_destroy(b->sv);  // cancels what is left; its callbacks are queued first
    b->sv = nullptr;
  }
  if (b->tsfn) {
    napi_release_threadsafe_function(b->tsfn, napi_tsfn_release);
    b->tsfn = nullptr;
  }
}

void finalize_server(napi_env, void* data, void*) {
  ServerBox* b = static_cast<ServerBox*>(data);
  server_close(b);
  delete b;
}

ServerBox* server_box(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p ||
      !static_cast<ServerBox*>(p)->sv) {
    napi_throw_type_error(env, "S2T_HANDLE", "not a server handle (destroyed?)");
    return nullptr;
  }
  return static_cast<ServerBox*>(p);
}

bool int_array(napi_env env, napi_value v, std::vector<int>& out) {
  bool is = false;
  napi_is_array(env, v, &is);
  if (!is) {
    napi_throw_type_error(env, "S2T_ARGS", "expected an array of numbers");
    return false;
  }
  uint32_t len = 0;
  napi_get_array_length(env, v, &len);
  out.resize(len);
  for (uint32_t i = 0; i < len; ++i) {
    napi_value e;
    napi_get_element(env, v, i, &e);
    out[i] = (int)num(env, e, 0);
  }
  return true;
}

// serverCreate(sizes[], slots[], budgetBytes, onFinal(id, status)) -> handle
napi_value ServerCreate(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 4)) return nullptr;
  std::vector<int> sizes, slots;
  if (!int_array(env, a.argv[0], sizes) || !int_array(env, a.argv[1], slots)) return nullptr;
  if (sizes.size() != slots.size()) {
    napi_throw_type_error(env, "S2T_ARGS", "sizes and slots differ in length");
    return nullptr;
  }
  ServerBox* b = new ServerBox();
  const int rc = s2t_server_create(sizes.data(), slots.data(), (int)sizes.size(),
                                   num(env, a.argv[2], 0), &b->sv);
  if (rc != S2T_OK) {
    delete b;
    return throw_status(env, rc);
  }
  napi_value name;
  napi_create_string_utf8(env, "s2t_server_final", NAPI_AUTO_LENGTH, &name);
  if (napi_create_threadsafe_function(env, a.argv[3], nullptr, name, 0, 1, nullptr, nullptr,
                                      nullptr, final_to_js, &b->tsfn) != napi_ok) {
    server_close(b);
    delete b;
    napi_throw_error(env, "S2T_NAPI", "napi_create_threadsafe_function failed");
    return nullptr;
  }
  s2t_server_set_callback(b->sv, server_final, b);
  napi_value ext;
  NAPI_CALL(env, napi_create_external(env, b, finalize_server, nullptr, &ext));
  return ext;
}

// serverSubmit(h, {n, steps, scenario, scenarioParam, k0, amplitude, seed, keepState}, cfg) -> id
napi_value ServerSubmit(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 3)) return nullptr;
  ServerBox* b = server_box(env, a.argv[0]);
  if (!b) return nullptr;
  napi_value r = a.argv[1];
  s2t_request q{};
  q.n = (int)prop(env, r, "n", 0);
  q.steps = (int)prop(env, r, "steps", 0);
  q.scenario = (int)prop(env, r, "scenario", S2T_SCENARIO_TURBULENCE);
  q.scenario_param = prop(env, r, "scenarioParam", 0);
  q.k0 = prop(env, r, "k0", 14);
  q.amplitude = prop(env, r, "amplitude", 6);
  q.seed = (uint32_t)prop(env, r, "seed", 1);
  q.keep_state = prop(env, r, "keepState", 0) != 0;
  q.params = params_of(env, a.argv[2]);
  uint64_t id = 0;
  int rc;
  {
    std::lock_guard<std::recursive_mutex> lk(b->mu);
    rc = s2t_server_submit(b->sv, &q, &id);
    // a request rejected inside the call has already called back; one admitted waits on mu
    if (rc == S2T_OK && q.keep_state) {
      int st = S2T_REQ_QUEUED;
      s2t_server_poll(b->sv, id, &st, nullptr, nullptr);
      if (st != S2T_REQ_REJECTED) b->keep[id] = q.n;
    }
  }
  S2T_CALL(env, rc);
  return make_num(env, (double)id);
}

// serverPoll(h, id) -> {status, latency, diagnostics fields}
napi_value ServerPoll(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  ServerBox* b = server_box(env, a.argv[0]);
  if (!b) return nullptr;
  int status = 0;
  s2t_diag d{};
  double lat = -1;
  S2T_CALL(env, s2t_server_poll(b->sv, (uint64_t)num(env, a.argv[1], 0), &status, &d, &lat));
  return diag_object(env, status, d, lat);
}

napi_value ServerCancel(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  ServerBox* b = server_box(env, a.argv[0]);
  if (!b) return nullptr;
  S2T_CALL(env, s2t_server_cancel(b->sv, (uint64_t)num(env, a.argv[1], 0)));
  return nullptr;
}

// serverReadState(h, id, n) -> Float32Array(2 n^2), fixture layout
napi_value ServerReadState(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 3)) return nullptr;
  ServerBox* b = server_box(env, a.argv[0]);
  if (!b) return nullptr;
  const size_t n = (size_t)num(env, a.argv[2], 0);
  void* data = nullptr;
  napi_value arr = new_typed(env, napi_float32_array, 2 * n * n, 4, &data);
  if (!arr) return nullptr;
  S2T_CALL(env, s2t_server_read_state(b->sv, (uint64_t)num(env, a.argv[1], 0),
                                      static_cast<float*>(data)));
  return arr;
}

napi_value ServerStats(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 1)) return nullptr;
  ServerBox* b = server_box(env, a.argv[0]);
  if (!b) return nullptr;
  s2t_server_stats st{};
  S2T_CALL(env, s2t_server_get_stats(b->sv, &st));
  napi_value obj;
  napi_create_object(env, &obj);
  napi_set_named_property(env, obj, "submitted", make_num(env, (double)st.submitted));
  napi_set_named_property(env, obj, "completed", make_num(env, (double)st.completed));
  napi_set_named_property(env, obj, "rejected", make_num(env, (double)st.rejected));
  napi_set_named_property(env, obj, "cancelled", make_num(env, (double)st.cancelled));
  napi_set_named_property(env, obj, "p50Latency", make_num(env, st.p50_latency_s));
  napi_set_named_property(env, obj, "p99Latency", make_num(env, st.p99_latency_s));
  napi_set_named_property(env, obj, "meanLatency", make_num(env, st.mean_latency_s));
  napi_set_named_property(env, obj, "mptsStepsPerS", make_num(env, st.mpts_steps_per_s));
  return obj;
}

// serverRef(h, on): whether the pending completion callbacks keep the event loop alive (the JS
// class holds it while requests are outstanding)
napi_value ServerRef(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  ServerBox* b = server_box(env, a.argv[0]);
  if (!b) return nullptr;
  bool on = false;
  napi_get_value_bool(env, a.argv[1], &on);
  if (on)
    NAPI_CALL(env, napi_ref_threadsafe_function(env, b->tsfn));
  else
    NAPI_CALL(env, napi_unref_threadsafe_function(env, b->tsfn));
  return nullptr;
}

// serverDestroy(h): cancels what is left (those callbacks still arrive), stops the worker, and
// releases the thread-safe function so the event loop may end
napi_value ServerDestroy(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 1)) return nullptr;
  void* p = nullptr;
  if (napi_get_value_external(env, a.argv[0], &p) == napi_ok && p)
    server_close(static_cast<ServerBox*>(p));
  return nullptr;
}

napi_value Init(napi_env env, napi_value exports) {
  struct Fn {
    const char* name;
    napi_callback cb;
  } fns[] = {{"create", Create},         {"radixSchedule", RadixSchedule},
             {"configure", Configure},   {"nuH", NuH},
             {"setSplats", SetSplats},   {"seedSpectrum", SeedSpectrum},
             {"seedPhysical", SeedPhysical}, {"writeState", WriteState},
             {"readState", ReadState},   {"resetClock", ResetClock},
             {"step", Step},             {"sync", Sync},
             {"diagnostics", Diagnostics}, {"diagnosticsAsync", DiagnosticsAsync},
             {"paintDye", PaintDye},     {"readDye", ReadDye},
             {"fft2", Fft2},             {"velocityPack", VelocityPack},
             {"assemble", Assemble},     {"rhs", Rhs},
             {"serverCreate", ServerCreate}, {"serverSubmit", ServerSubmit},
             {"serverPoll", ServerPoll}, {"serverCancel", ServerCancel},
             {"serverReadState", ServerReadState}, {"serverStats", ServerStats},
             {"serverRef", ServerRef}, {"serverDestroy", ServerDestroy}};
  for (const Fn& f : fns) {
    napi_value fn;
    NAPI_CALL(env, napi_create_function(env, f.name, NAPI_AUTO_LENGTH, f.cb, nullptr, &fn));
    NAPI_CALL(env, napi_set_named_property(env, exports, f.name, fn));
  }
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
