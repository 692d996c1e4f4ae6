// N-API addon over libs2t.so (include/s2t.h). Thin: every function converts JS values, calls one
// C entry point and throws a JS Error carrying s2t_last_error() on a negative status (the
// reference throws Error for programming errors, src/gpu/fft.js:28,84). Long device work that
// returns data (diagnostics), and a run of steps, are also offered as Promises on a worker thread,
// like the reference's async readDiagnostics (src/gpu/solver.js:570-608). The serving loop reports each request's final
// status from its native worker thread through a thread-safe function.
#include <node_api.h>

#include <condition_variable>
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

// A solver handle. Calls on one solver must not overlap (the C library's contract), and the
// Promise-returning calls run on libuv worker threads. Every call takes a ticket on the JS thread
// when it is made and runs when its turn comes, so calls take effect in the order the program
// made them (as work encoded into the reference's GPU queue does): a synchronous call made while
// a Promise-returning one is pending waits for it.
struct SolverBox {
  s2t_solver* h = nullptr;
  std::mutex mu;
  std::condition_variable cv;
  uint64_t issued = 0, served = 0;
  uint64_t ticket() {
    std::lock_guard<std::mutex> lk(mu);
    return issued++;
  }
  void begin(uint64_t t) {
    std::unique_lock<std::mutex> lk(mu);
    cv.wait(lk, [&] { return served == t; });
  }
  void end() {
    {
      std::lock_guard<std::mutex> lk(mu);
      served += 1;
    }
    cv.notify_all();
  }
};

// the solver of a handle for the rest of a call, in its turn (empty on a bad handle, with a JS
// TypeError pending)
class Held {
 public:
  Held() = default;
  explicit Held(SolverBox* b) : b_(b) { b_->begin(b_->ticket()); }
  Held(Held&& o) noexcept : b_(o.b_) { o.b_ = nullptr; }
  Held(const Held&) = delete;
  Held& operator=(const Held&) = delete;
  ~Held() {
    if (b_) b_->end();
  }
  explicit operator bool() const { return b_ != nullptr; }
  operator s2t_solver*() const { return b_ ? b_->h : nullptr; }

 private:
  SolverBox* b_ = nullptr;
};

SolverBox* solver_box(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, "S2T_HANDLE", "not a solver handle (destroyed?)");
    return nullptr;
  }
  return static_cast<SolverBox*>(p);
}

Held handle(napi_env env, napi_value v) {
  SolverBox* b = solver_box(env, v);
  return b ? Held(b) : Held();
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

// runs when no JS value and no pending job refers to the handle any more
void finalize_solver(napi_env, void* data, void*) {
  SolverBox* b = static_cast<SolverBox*>(data);
  s2t_destroy(b->h);
  delete b;
}

// ---------------------------------------------------------------------------------------------

napi_value Create(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  const int n = (int)num(env, a.argv[0], 256), dye = (int)num(env, a.argv[1], 2);
  s2t_solver* h = nullptr;
  S2T_CALL(env, s2t_create(n, dye, &h));
  SolverBox* b = new SolverBox();
  b->h = h;
  napi_value ext;
  if (napi_create_external(env, b, finalize_solver, nullptr, &ext) != napi_ok) {
    finalize_solver(env, b, nullptr);
    napi_throw_error(env, "S2T_NAPI", "napi_create_external failed");
    return nullptr;
  }
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
  Held h = handle(env, a.argv[0]);
  if (!h) return nullptr;
  const s2t_params p = params_of(env, a.argv[1]);
  S2T_CALL(env, s2t_configure(h, &p));
  return nullptr;
}

napi_value NuH(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 1)) return nullptr;
  Held h = handle(env, a.argv[0]);
  if (!h) return nullptr;
  double v = 0;
  S2T_CALL(env, s2t_nu_h(h, &v));
  return make_num(env, v);
}

// setSplats(h, list): items {x, y, dx, dy, color: [r, g, b], radius, amp, dyeAmp}
napi_value SetSplats(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  Held h = handle(env, a.argv[0]);
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
  Held h = handle(env, a.argv[0]);
  if (!h) return nullptr;
  const double seed = num(env, a.argv[3], 1);
  S2T_CALL(env, s2t_seed_spectrum(h, num(env, a.argv[1], 14), num(env, a.argv[2], 6),
                                  (uint32_t)(int64_t)seed));
  return nullptr;
}

napi_value SeedPhysical(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 3)) return nullptr;
  Held h = handle(env, a.argv[0]);
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
  Held h = handle(env, a.argv[0]);
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
  Held h = handle(env, a.argv[0]);
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
  Held h = handle(env, a.argv[0]);
  if (!h) return nullptr;
  S2T_CALL(env, s2t_reset_clock(h));
  return nullptr;
}

// setClock(h, time, stepIndex) (s2t_set_clock)
napi_value SetClock(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 3)) return nullptr;
  Held h = handle(env, a.argv[0]);
  if (!h) return nullptr;
  S2T_CALL(env, s2t_set_clock(h, num(env, a.argv[1], 0),
                              (uint32_t)(int64_t)num(env, a.argv[2], 0)));
  return nullptr;
}

napi_value Step(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 3)) return nullptr;
  Held h = handle(env, a.argv[0]);
  if (!h) return nullptr;
  bool dye = false;
  napi_get_value_bool(env, a.argv[2], &dye);
  S2T_CALL(env, s2t_step(h, (int)num(env, a.argv[1], 1), dye ? S2T_STEP_DYE : 0u));
  return nullptr;
}

napi_value Sync(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 1)) return nullptr;
  Held h = handle(env, a.argv[0]);
  if (!h) return nullptr;
  S2T_CALL(env, s2t_sync(h));
  return nullptr;
}

// Work that runs on a libuv worker thread and settles a Promise: it holds a reference to the
// solver's handle (so the solver outlives it) and runs in the turn of the call that queued it.
struct Job {
  SolverBox* box = nullptr;
  napi_ref ref = nullptr;
  bool in_turn = true;  // false: runs beside the solver's calls (it touches nothing they use)
  uint64_t turn = 0;
  int rc = 0;
  std::string err;
  napi_deferred deferred = nullptr;
  napi_async_work work = nullptr;
  virtual ~Job() = default;
  virtual int run(s2t_solver* h) = 0;
  virtual napi_value result(napi_env env) = 0;
  void execute() {
    if (in_turn) box->begin(turn);
    rc = run(box->h);
    if (rc != S2T_OK) err = s2t_last_error();  // thread-local: read on the same thread
    if (in_turn) box->end();
  }
};

void job_execute(napi_env, void* data) { static_cast<Job*>(data)->execute(); }

void job_complete(napi_env env, napi_status, void* data) {
  Job* j = static_cast<Job*>(data);
  if (j->rc == S2T_OK) {
    napi_resolve_deferred(env, j->deferred, j->result(env));
  } else {
    napi_value msg, code, err;
    const std::string c = "S2T_" + std::to_string(j->rc);
    napi_create_string_utf8(env, c.c_str(), NAPI_AUTO_LENGTH, &code);
    napi_create_string_utf8(env, j->err.c_str(), NAPI_AUTO_LENGTH, &msg);
    napi_create_error(env, code, msg, &err);
    napi_reject_deferred(env, j->deferred, err);
  }
  napi_delete_reference(env, j->ref);
  napi_delete_async_work(env, j->work);
  delete j;
}

// queues j for the solver handle v; returns its Promise (j is owned by the queue from here on)
napi_value queue_job(napi_env env, napi_value v, Job* j, const char* what) {
  j->box = solver_box(env, v);
  napi_value promise = nullptr, name;
  if (!j->box || napi_create_reference(env, v, 1, &j->ref) != napi_ok) {
    delete j;
    return nullptr;
  }
  bool ok = napi_create_promise(env, &j->deferred, &promise) == napi_ok &&
            napi_create_string_utf8(env, what, NAPI_AUTO_LENGTH, &name) == napi_ok &&
            napi_create_async_work(env, nullptr, name, job_execute, job_complete, j, &j->work) ==
                napi_ok;
  if (ok) {
    if (j->in_turn) j->turn = j->box->ticket();
    ok = napi_queue_async_work(env, j->work) == napi_ok;
    if (!ok && j->in_turn) {
      // give the turn back in order
      j->box->begin(j->turn);
      j->box->end();
    }
  }
  if (!ok) {
    napi_delete_reference(env, j->ref);
    if (j->work) napi_delete_async_work(env, j->work);
    delete j;
    napi_throw_error(env, "S2T_NAPI", "could not queue the work");
    return nullptr;
  }
  return promise;
}

// diagnostics record (solver.js:591-602) in FP64
struct DiagJob : Job {
  int n = 0;
  s2t_diag d{};
  std::vector<double> spectrum, zonal;
  int run(s2t_solver* h) override {
    spectrum.assign(n / 2 + 1, 0.0);
    zonal.assign(n, 0.0);
    return s2t_diagnostics(h, &d, spectrum.data(), zonal.data());
  }
  napi_value result(napi_env env) override;
};

napi_value DiagJob::result(napi_env env) {
  napi_value obj;
  napi_create_object(env, &obj);
  void* data = nullptr;
  napi_value spec = new_typed(env, napi_float64_array, spectrum.size(), 8, &data);
  memcpy(data, spectrum.data(), spectrum.size() * 8);
  napi_value zon = new_typed(env, napi_float64_array, zonal.size(), 8, &data);
  memcpy(data, zonal.data(), zonal.size() * 8);
  napi_set_named_property(env, obj, "spectrum", spec);
  napi_set_named_property(env, obj, "zonal", zon);
  napi_set_named_property(env, obj, "energy", make_num(env, d.energy));
  napi_set_named_property(env, obj, "enstrophy", make_num(env, d.enstrophy));
  napi_set_named_property(env, obj, "dt", make_num(env, d.dt));
  napi_set_named_property(env, obj, "maxSpeed", make_num(env, d.max_speed));
  napi_set_named_property(env, obj, "time", make_num(env, d.time));
  napi_set_named_property(env, obj, "stepIndex", make_num(env, d.step_index));
  napi_value nf;
  napi_get_boolean(env, d.nonfinite != 0, &nf);
  napi_set_named_property(env, obj, "nonfinite", nf);
  return obj;
}

napi_value Diagnostics(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  Held h = handle(env, a.argv[0]);
  if (!h) return nullptr;
  DiagJob j;
  j.n = (int)num(env, a.argv[1], 0);
  S2T_CALL(env, j.run(h));
  return j.result(env);
}

// diagnosticsAsync(h, n) -> Promise of the diagnostics record
napi_value DiagnosticsAsync(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  DiagJob* j = new DiagJob();
  j->n = (int)num(env, a.argv[1], 0);
  return queue_job(env, a.argv[0], j, "s2t_diagnostics");
}

// readDiagnostics of the copied record (s2t_read_copied_diagnostics): waits for that copy only,
// outside the solver's turn, so the frames encoded after it are not held up (the JS Solver
// allows no new copy until this read has finished, solver.js:560-608)
struct CopiedDiagJob : DiagJob {
  int run(s2t_solver* h) override {
    spectrum.assign(n / 2 + 1, 0.0);
    zonal.assign(n, 0.0);
    return s2t_read_copied_diagnostics(h, &d, spectrum.data(), zonal.data());
  }
};

napi_value ReadCopiedDiagnosticsAsync(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  CopiedDiagJob* j = new CopiedDiagJob();
  j->n = (int)num(env, a.argv[1], 0);
  j->in_turn = false;
  return queue_job(env, a.argv[0], j, "s2t_read_copied_diagnostics");
}

#define SOLVER_CALL(NAME, EXPR)                                  \
  napi_value NAME(napi_env env, napi_callback_info info) {      \
    Args a;                                                      \
    if (!get_args(env, info, a, 1)) return nullptr;              \
    Held h = handle(env, a.argv[0]);                             \
    if (!h) return nullptr;                                      \
    S2T_CALL(env, EXPR);                                         \
    return nullptr;                                              \
  }

SOLVER_CALL(Prime, s2t_prime(h))
SOLVER_CALL(ResetDye, s2t_reset_dye(h))
SOLVER_CALL(EncodeDiagnostics, s2t_encode_diagnostics(h))
SOLVER_CALL(CopyDiagnostics, s2t_copy_diagnostics(h))

napi_value Primed(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 1)) return nullptr;
  Held h = handle(env, a.argv[0]);
  if (!h) return nullptr;
  int p = 0;
  S2T_CALL(env, s2t_primed(h, &p));
  napi_value v;
  napi_get_boolean(env, p != 0, &v);
  return v;
}

// dyeStep(h, dt): dt <= 0 takes the last step's dt on the device (encodeDye)
napi_value DyeStep(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  Held h = handle(env, a.argv[0]);
  if (!h) return nullptr;
  S2T_CALL(env, s2t_dye_step(h, num(env, a.argv[1], 0.0)));
  return nullptr;
}

// ---------------------------------------------------------------------------------------------
// device buffers (s2t_buffer_*) for the GPUBuffer-like objects of the JS device
// ---------------------------------------------------------------------------------------------
struct BufBox {
  void* p = nullptr;
  size_t bytes = 0;
};

void finalize_buffer(napi_env, void* data, void*) {
  BufBox* b = static_cast<BufBox*>(data);
  s2t_buffer_destroy(b->p);
  delete b;
}

BufBox* buffer_box(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p || !static_cast<BufBox*>(p)->p) {
    napi_throw_type_error(env, "S2T_HANDLE", "not a device buffer (destroyed?)");
    return nullptr;
  }
  return static_cast<BufBox*>(p);
}

bool in_range(napi_env env, const BufBox* b, double off, double bytes) {
  if (off < 0 || bytes < 0 || off + bytes > (double)b->bytes) {
    napi_throw_range_error(env, "S2T_ARGS", "range outside the buffer");
    return false;
  }
  return true;
}

napi_value BufferCreate(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 1)) return nullptr;
  const double bytes = num(env, a.argv[0], 0);
  BufBox* b = new BufBox();
  b->bytes = (size_t)bytes;
  const int rc = s2t_buffer_create(b->bytes, &b->p);
  if (rc != S2T_OK) {
    delete b;
    return throw_status(env, rc);
  }
  napi_value ext;
  if (napi_create_external(env, b, finalize_buffer, nullptr, &ext) != napi_ok) {
    s2t_buffer_destroy(b->p);
    delete b;
    napi_throw_error(env, "S2T_NAPI", "could not wrap the buffer");
    return nullptr;
  }
  return ext;
}

napi_value BufferDestroy(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 1)) return nullptr;
  void* p = nullptr;
  if (napi_get_value_external(env, a.argv[0], &p) == napi_ok && p) {
    BufBox* b = static_cast<BufBox*>(p);
    s2t_buffer_destroy(b->p);
    b->p = nullptr;
  }
  return nullptr;
}

// bufferWrite(buf, offset, typedArray)
napi_value BufferWrite(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 3)) return nullptr;
  BufBox* b = buffer_box(env, a.argv[0]);
  if (!b) return nullptr;
  bool is = false;
  napi_is_typedarray(env, a.argv[2], &is);
  if (!is) {
    napi_throw_type_error(env, "S2T_ARGS", "expected a typed array");
    return nullptr;
  }
  napi_typedarray_type t;
  size_t len = 0, aoff = 0;
  void* data = nullptr;
  napi_value ab;
  napi_get_typedarray_info(env, a.argv[2], &t, &len, &data, &ab, &aoff);
  size_t elem = 1;
  switch (t) {
    case napi_int16_array: case napi_uint16_array: elem = 2; break;
    case napi_int32_array: case napi_uint32_array: case napi_float32_array: elem = 4; break;
    case napi_float64_array: case napi_bigint64_array: case napi_biguint64_array: elem = 8; break;
    default: elem = 1;
  }
  const double off = num(env, a.argv[1], 0);
  if (!in_range(env, b, off, (double)(len * elem))) return nullptr;
  S2T_CALL(env, s2t_buffer_write(b->p, (size_t)off, data, len * elem));
  return nullptr;
}

// bufferRead(buf, offset, bytes) -> ArrayBuffer
napi_value BufferRead(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 3)) return nullptr;
  BufBox* b = buffer_box(env, a.argv[0]);
  if (!b) return nullptr;
  const double off = num(env, a.argv[1], 0), bytes = num(env, a.argv[2], 0);
  if (!in_range(env, b, off, bytes)) return nullptr;
  void* data = nullptr;
  napi_value ab;
  NAPI_CALL(env, napi_create_arraybuffer(env, (size_t)bytes, &data, &ab));
  S2T_CALL(env, s2t_buffer_read(b->p, (size_t)off, data, (size_t)bytes));
  return ab;
}

// bufferCopy(src, srcOffset, dst, dstOffset, bytes)
napi_value BufferCopy(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 5)) return nullptr;
  BufBox* src = buffer_box(env, a.argv[0]);
  if (!src) return nullptr;
  BufBox* dst = buffer_box(env, a.argv[2]);
  if (!dst) return nullptr;
  const double so = num(env, a.argv[1], 0), dof = num(env, a.argv[3], 0),
               bytes = num(env, a.argv[4], 0);
  if (!in_range(env, src, so, bytes) || !in_range(env, dst, dof, bytes)) return nullptr;
  S2T_CALL(env, s2t_buffer_copy(dst->p, (size_t)dof, src->p, (size_t)so, (size_t)bytes));
  return nullptr;
}

// fft2Device(h, buf, n, inverse): in place on the solver's stream
napi_value Fft2Device(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 4)) return nullptr;
  BufBox* b = buffer_box(env, a.argv[1]);
  if (!b) return nullptr;
  const double n = num(env, a.argv[2], 0);
  if (!(n > 0) || (double)b->bytes < 8.0 * n * n) {
    napi_throw_range_error(env, "S2T_ARGS", "buffer smaller than an n x n complex field");
    return nullptr;
  }
  Held h = handle(env, a.argv[0]);
  if (!h) return nullptr;
  bool inv = false;
  napi_get_value_bool(env, a.argv[3], &inv);
  S2T_CALL(env, s2t_fft2_device(h, static_cast<float*>(b->p), inv ? 1 : 0));
  return nullptr;
}

// `count` steps (+ dye), then a wait for the device: the Promise settles when they are done
struct StepJob : Job {
  int count = 1;
  uint32_t flags = 0;
  int run(s2t_solver* h) override {
    const int rc = s2t_step(h, count, flags);
    return rc != S2T_OK ? rc : s2t_sync(h);
  }
  napi_value result(napi_env env) override {
    napi_value u;
    napi_get_undefined(env, &u);
    return u;
  }
};

// stepAsync(h, count, dye) -> Promise
napi_value StepAsync(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 3)) return nullptr;
  StepJob* j = new StepJob();
  j->count = (int)num(env, a.argv[1], 1);
  bool dye = false;
  napi_get_value_bool(env, a.argv[2], &dye);
  j->flags = dye ? S2T_STEP_DYE : 0u;
  return queue_job(env, a.argv[0], j, "s2t_step");
}

// autotune(h, n, steps) -> {width, ms: {width: ms per step}} (s2t_autotune)
napi_value Autotune(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 3)) return nullptr;
  Held h = handle(env, a.argv[0]);
  if (!h) return nullptr;
  int widths[8], count = 0;
  S2T_CALL(env, s2t_column_widths((int)num(env, a.argv[1], 0), widths, 8, &count));
  int w = 0;
  double ms[8] = {0};
  S2T_CALL(env, s2t_autotune(h, (int)num(env, a.argv[2], 5), &w, ms, 8));
  napi_value obj, per;
  napi_create_object(env, &obj);
  napi_create_object(env, &per);
  for (int i = 0; i < count && i < 8; ++i)
    napi_set_named_property(env, per, std::to_string(widths[i]).c_str(), make_num(env, ms[i]));
  napi_set_named_property(env, obj, "width", make_num(env, w));
  napi_set_named_property(env, obj, "ms", per);
  return obj;
}

napi_value PaintDye(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  Held h = handle(env, a.argv[0]);
  if (!h) return nullptr;
  S2T_CALL(env, s2t_paint_dye(h, (int)num(env, a.argv[1], 1)));
  return nullptr;
}

napi_value ReadDye(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 1)) return nullptr;
  Held h = handle(env, a.argv[0]);
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
  Held h = handle(env, a.argv[0]);
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
  Held h = handle(env, a.argv[0]);
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
  Held h = handle(env, a.argv[0]);
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
  Held h = handle(env, a.argv[0]);
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
// ensembles (s2t_batch_*)
// ---------------------------------------------------------------------------------------------
void finalize_batch(napi_env, void* data, void*) { s2t_batch_destroy(static_cast<s2t_batch*>(data)); }

s2t_batch* batch(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, "S2T_HANDLE", "not a batch handle");
    return nullptr;
  }
  return static_cast<s2t_batch*>(p);
}

bool int_array(napi_env env, napi_value v, std::vector<int>& out);

// batchCreate(sizes[]) -> handle
napi_value BatchCreate(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 1)) return nullptr;
  std::vector<int> sizes;
  if (!int_array(env, a.argv[0], sizes)) return nullptr;
  s2t_batch* b = nullptr;
  S2T_CALL(env, s2t_batch_create(sizes.data(), (int)sizes.size(), &b));
  napi_value ext;
  if (napi_create_external(env, b, finalize_batch, nullptr, &ext) != napi_ok) {
    s2t_batch_destroy(b);
    napi_throw_error(env, "S2T_NAPI", "napi_create_external failed");
    return nullptr;
  }
  return ext;
}

// the member's grid size, or 0 with a JS error pending
int member_size(napi_env env, s2t_batch* b, int i) {
  int n = 0;
  if (s2t_batch_size(b, i, &n) != S2T_OK) {
    throw_status(env, S2T_EINVAL);
    return 0;
  }
  return n;
}

#define BATCH_MEMBER(argc)                          \
  Args a;                                           \
  if (!get_args(env, info, a, (argc))) return nullptr; \
  s2t_batch* b = batch(env, a.argv[0]);             \
  if (!b) return nullptr;                           \
  const int i = (int)num(env, a.argv[1], -1)

napi_value BatchSize(napi_env env, napi_callback_info info) {
  BATCH_MEMBER(2);
  const int n = member_size(env, b, i);
  return n ? make_num(env, n) : nullptr;
}

napi_value BatchConfigure(napi_env env, napi_callback_info info) {
  BATCH_MEMBER(3);
  const s2t_params p = params_of(env, a.argv[2]);
  S2T_CALL(env, s2t_batch_configure(b, i, &p));
  return nullptr;
}

napi_value BatchSeedSpectrum(napi_env env, napi_callback_info info) {
  BATCH_MEMBER(5);
  S2T_CALL(env, s2t_batch_seed_spectrum(b, i, num(env, a.argv[2], 14), num(env, a.argv[3], 6),
                                        (uint32_t)(int64_t)num(env, a.argv[4], 1)));
  return nullptr;
}

// batchSeedPhysical(h, i, Float32Array n*n) / batchWriteState(h, i, Float32Array 2*n*n)
napi_value BatchSeedPhysical(napi_env env, napi_callback_info info) {
  BATCH_MEMBER(3);
  const int n = member_size(env, b, i);
  if (!n) return nullptr;
  size_t cnt = 0;
  float* f = typed<float>(env, a.argv[2], napi_float32_array, &cnt);
  if (!f) return nullptr;
  if (cnt != (size_t)n * n) {
    napi_throw_range_error(env, "S2T_ARGS", "field must hold n*n values");
    return nullptr;
  }
  S2T_CALL(env, s2t_batch_seed_physical(b, i, f));
  return nullptr;
}

napi_value BatchWriteState(napi_env env, napi_callback_info info) {
  BATCH_MEMBER(3);
  const int n = member_size(env, b, i);
  if (!n) return nullptr;
  size_t cnt = 0;
  float* f = typed<float>(env, a.argv[2], napi_float32_array, &cnt);
  if (!f) return nullptr;
  if (cnt != 2 * (size_t)n * n) {
    napi_throw_range_error(env, "S2T_ARGS", "spectrum must hold 2*n*n values");
    return nullptr;
  }
  S2T_CALL(env, s2t_batch_write_state(b, i, f));
  return nullptr;
}

napi_value BatchReadState(napi_env env, napi_callback_info info) {
  BATCH_MEMBER(2);
  const int n = member_size(env, b, i);
  if (!n) return nullptr;
  void* data = nullptr;
  napi_value arr = new_typed(env, napi_float32_array, 2 * (size_t)n * n, 4, &data);
  if (!arr) return nullptr;
  S2T_CALL(env, s2t_batch_read_state(b, i, static_cast<float*>(data)));
  return arr;
}

napi_value BatchDiagnostics(napi_env env, napi_callback_info info) {
  BATCH_MEMBER(2);
  const int n = member_size(env, b, i);
  if (!n) return nullptr;
  DiagJob j;
  j.n = n;
  j.spectrum.assign(n / 2 + 1, 0.0);
  j.zonal.assign(n, 0.0);
  S2T_CALL(env, s2t_batch_diagnostics(b, i, &j.d, j.spectrum.data(), j.zonal.data()));
  return j.result(env);
}

napi_value BatchUseGraphs(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  s2t_batch* b = batch(env, a.argv[0]);
  if (!b) return nullptr;
  bool on = true;
  napi_get_value_bool(env, a.argv[1], &on);
  S2T_CALL(env, s2t_batch_use_graphs(b, on ? 1 : 0));
  return nullptr;
}

napi_value BatchUseGrouped(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  s2t_batch* b = batch(env, a.argv[0]);
  if (!b) return nullptr;
  bool on = false;
  napi_get_value_bool(env, a.argv[1], &on);
  S2T_CALL(env, s2t_batch_use_grouped(b, on ? 1 : 0));
  return nullptr;
}

// batchAutotune(h, steps); batchColumnWidth(h, i) -> width
napi_value BatchAutotune(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  s2t_batch* b = batch(env, a.argv[0]);
  if (!b) return nullptr;
  S2T_CALL(env, s2t_batch_autotune(b, (int)num(env, a.argv[1], 5)));
  return nullptr;
}

napi_value BatchColumnWidth(napi_env env, napi_callback_info info) {
  BATCH_MEMBER(2);
  int w = 0;
  S2T_CALL(env, s2t_batch_column_width(b, i, &w));
  return make_num(env, w);
}

napi_value BatchStep(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  s2t_batch* b = batch(env, a.argv[0]);
  if (!b) return nullptr;
  S2T_CALL(env, s2t_batch_step(b, (int)num(env, a.argv[1], 1)));
  return nullptr;
}

napi_value BatchSync(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 1)) return nullptr;
  s2t_batch* b = batch(env, a.argv[0]);
  if (!b) return nullptr;
  S2T_CALL(env, s2t_batch_sync(b));
  return nullptr;
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
  std::string error;  // the server's error when the request failed
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
  if (status == S2T_REQ_FAILED) {
    int st = 0;
    if (s2t_server_poll(b->sv, id, &st, nullptr, nullptr) != S2T_OK) f->error = s2t_last_error();
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
// Float32Array in fixture layout, for keepState) when done, the server's error text when failed,
// else undefined
void final_to_js(napi_env env, napi_value js_cb, void*, void* data) {
  Final* f = static_cast<Final*>(data);
  if (env && js_cb) {
    napi_value args[3], undef;
    napi_get_undefined(env, &undef);
    napi_create_double(env, (double)f->id, &args[0]);
    napi_create_int32(env, f->status, &args[1]);
    args[2] = undef;
    if (f->status == S2T_REQ_FAILED)
      napi_create_string_utf8(env, f->error.c_str(), NAPI_AUTO_LENGTH, &args[2]);
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
  napi_set_named_property(env, obj, "budgetBytes", make_num(env, st.budget_bytes));
  napi_set_named_property(env, obj, "usedBytes", make_num(env, st.used_bytes));
  napi_set_named_property(env, obj, "peakBytes", make_num(env, st.peak_bytes));
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
napi_value ServerInjectFailure(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  ServerBox* b = server_box(env, a.argv[0]);
  if (!b) return nullptr;
  S2T_CALL(env, s2t_server_inject_failure(b->sv, (int)num(env, a.argv[1], -1)));
  return nullptr;
}

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
             {"stepAsync", StepAsync},
             {"paintDye", PaintDye},     {"readDye", ReadDye},
             {"fft2", Fft2},             {"velocityPack", VelocityPack},
             {"assemble", Assemble},     {"rhs", Rhs},
             {"serverCreate", ServerCreate}, {"serverSubmit", ServerSubmit},
             {"serverPoll", ServerPoll}, {"serverCancel", ServerCancel},
             {"serverReadState", ServerReadState}, {"serverStats", ServerStats},
             {"serverRef", ServerRef}, {"serverDestroy", ServerDestroy},
             {"batchCreate", BatchCreate}, {"batchSize", BatchSize},
             {"batchConfigure", BatchConfigure}, {"batchSeedSpectrum", BatchSeedSpectrum},
             {"batchSeedPhysical", BatchSeedPhysical}, {"batchWriteState", BatchWriteState},
             {"batchReadState", BatchReadState}, {"batchDiagnostics", BatchDiagnostics},
             {"batchUseGraphs", BatchUseGraphs}, {"batchUseGrouped", BatchUseGrouped},
             {"batchStep", BatchStep},
             {"batchSync", BatchSync}, {"batchAutotune", BatchAutotune},
             {"batchColumnWidth", BatchColumnWidth}, {"autotune", Autotune},
             {"setClock", SetClock},    {"prime", Prime},
             {"primed", Primed},        {"resetDye", ResetDye},
             {"dyeStep", DyeStep},      {"encodeDiagnostics", EncodeDiagnostics},
             {"copyDiagnostics", CopyDiagnostics},
             {"readCopiedDiagnosticsAsync", ReadCopiedDiagnosticsAsync},
             {"bufferCreate", BufferCreate}, {"bufferDestroy", BufferDestroy},
             {"bufferWrite", BufferWrite}, {"bufferRead", BufferRead},
             {"bufferCopy", BufferCopy}, {"fft2Device", Fft2Device},
             {"serverInjectFailure", ServerInjectFailure}};
  for (const Fn& f : fns) {
    napi_value fn;
    NAPI_CALL(env, napi_create_function(env, f.name, NAPI_AUTO_LENGTH, f.cb, nullptr, &fn));
    NAPI_CALL(env, napi_set_named_property(env, exports, f.name, fn));
  }
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
