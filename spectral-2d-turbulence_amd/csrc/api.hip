// C ABI of libs2t.so (include/s2t.h): the solver object, buffer ownership, step scheduling and
// the test hooks. Mirrors the WebGPU Solver of src/gpu/solver.js.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/s2t.h"
#include "aux.hpp"
#include "solver.hpp"

namespace s2t {

static thread_local std::string g_err;

struct Fail : std::runtime_error {
  int code;
  Fail(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

static void hipchk(hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  if (e == hipErrorOutOfMemory) throw Fail(S2T_ENOMEM, std::string(what) + ": out of device memory");
  throw Fail(S2T_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

void check_launch(const char* what) { hipchk(hipGetLastError(), what); }

template <class T>
static T* dalloc(size_t elems, const char* what) {
  void* p = nullptr;
  hipchk(hipMalloc(&p, elems * sizeof(T) + 256), what);
  return (T*)p;
}

void prof_begin(Solver* s, int kind) {
  if (!s->prof) return;
  if (s->nev == s->cap_ev) {
    const int cap = s->cap_ev ? s->cap_ev * 2 : 256;
    Solver::Ev* ev = (Solver::Ev*)realloc(s->evs, sizeof(Solver::Ev) * cap);
    if (!ev) throw Fail(S2T_ENOMEM, "profiling event list");
    for (int i = s->cap_ev; i < cap; ++i) {
      hipchk(hipEventCreate(&ev[i].a), "hipEventCreate");
      hipchk(hipEventCreate(&ev[i].b), "hipEventCreate");
    }
    s->evs = ev;
    s->cap_ev = cap;
  }
  s->evs[s->nev].kind = kind;
  hipchk(hipEventRecord(s->evs[s->nev].a, s->stream), "hipEventRecord");
}

void prof_end(Solver* s) {
  if (!s->prof) return;
  hipchk(hipEventRecord(s->evs[s->nev].b, s->stream), "hipEventRecord");
  s->nev += 1;
}

static void prof_collect(Solver* s) {
  if (!s->nev) return;
  hipchk(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
  for (int i = 0; i < s->nev; ++i) {
    float ms = 0.f;
    hipchk(hipEventElapsedTime(&ms, s->evs[i].a, s->evs[i].b), "hipEventElapsedTime");
    s->prof_launches[s->evs[i].kind] += 1;
    s->prof_ms[s->evs[i].kind] += ms;
  }
  s->nev = 0;
}

static int radix(int n, int* stages, int* r2) {
  if (n < 1) return -1;
  int st = 0, m = n;
  while (m % 4 == 0) {
    ++st;
    m /= 4;
  }
  if (m != 1 && m != 2) return -1;
  if (stages) *stages = st;
  if (r2) *r2 = m == 2;
  return 0;
}

static Phys make_phys(int n, const s2t_params& p) {
  Phys ph{};
  const double kcut = n / 3.0;
  ph.beta = p.beta;
  ph.mu = p.drag;
  ph.p = p.hyper_p;
  ph.nu = p.viscosity > 0 ? p.viscosity / std::pow(kcut, 2.0 * p.hyper_p) : 0.0;
  const double ip = std::floor(p.hyper_p);
  ph.p_int = (ip == p.hyper_p && ip >= 1 && ip <= 16) ? (int)ip : 0;
  ph.force_amp = (float)p.force_amp;
  ph.force_k = (float)p.force_k;
  ph.force_width = (float)p.force_width;
  ph.force_norm = (float)(0.01 * n * n);  // FORCE_NORM, src/gpu/shaders.js:26
  ph.cfl = p.cfl;
  ph.dt_min = p.dt_min;
  ph.dt_max = p.dt_max;
  ph.dx = 2.0 * M_PI / n;
  ph.fixed_dt = p.fixed_dt;
  ph.dye_fade = (float)p.dye_fade;
  return ph;
}

// defaults of the reference's config object (src/main.js:9-36)
static s2t_params default_params() {
  s2t_params p{};
  p.beta = 0;
  p.drag = 0.01;
  p.viscosity = 0.67;
  p.hyper_p = 4;
  p.force_amp = 0;
  p.force_k = 22;
  p.force_width = 2.5;
  p.cfl = 0.5;
  p.dt_min = 1e-5;
  p.dt_max = 0.02;
  p.dye_fade = 0.999;
  p.fixed_dt = 0;
  return p;
}

static void need_full(Solver* s) {
  // three full fields plus one compact state (the fused rhs hook)
  if (!s->full)
    s->full = dalloc<float2>((size_t)s->n * s->n * 3 + s->ops->state_elems, "full scratch");
  s->full2 = s->full + (size_t)s->n * s->n;
}
static void need_force(Solver* s) {
  if (!s->mixed_f) s->mixed_f = dalloc<float2>(s->ops->plane_elems, "force tiles");
  if (!s->force_k) s->force_k = dalloc<float2>(s->ops->state_elems, "force spectrum");
}
static void need_vel(Solver* s) {
  if (!s->vel) {
    s->vel = dalloc<float2>((size_t)s->n * s->n, "velocity");
    fill_u32(s->stream, s->vel, (size_t)s->n * s->n * 2, 0u);
  }
}

static void destroy(Solver* s) {
  if (!s) return;
  if (s->stream) hipStreamSynchronize(s->stream);
  void* bufs[] = {s->state[0], s->state[1], s->state[2], s->mixed_u, s->mixed_z, s->mixed_f,
                  s->force_k,  s->vel,      s->full,     s->tw,      s->wtab,    s->dev,
                  s->diag_bins, s->diag_zonal, s->seed_sum, s->dye[0], s->dye[1], s->dye[2]};
  for (void* b : bufs)
    if (b) hipFree(b);
  for (int i = 0; i < s->cap_ev; ++i) {
    hipEventDestroy(s->evs[i].a);
    hipEventDestroy(s->evs[i].b);
  }
  free(s->evs);
  if (s->stream) hipStreamDestroy(s->stream);
  delete s;
}

static void reset_clock_s(Solver* s) {
  reset_clock(s->stream, s->dev, s->params.dt_max);
  s->primed = false;
}

static void h2d(Solver* s, void* dst, const void* src, size_t bytes) {
  hipchk(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s->stream), "hipMemcpyAsync H2D");
  hipchk(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
}
static void d2h(Solver* s, void* dst, const void* src, size_t bytes) {
  hipchk(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s->stream), "hipMemcpyAsync D2H");
  hipchk(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
}

static void ensure_primed(Solver* s) {
  if (!s->primed) {
    s->ops->prime(s);
    s->primed = true;
  }
}

static void step_once(Solver* s, bool dye) {
  const bool force = s->splats.count > 0;
  if (force) need_force(s);
  if (dye) need_vel(s);
  ensure_primed(s);
  // encodeStep (solver.js:488-521): three stages of [rows, cols]
  s->ops->rows(s, 0, true, dye, force);
  s->ops->cols(s, 0, force);
  s->ops->rows(s, 1, false, false, false);
  s->ops->cols(s, 1, force);
  s->ops->rows(s, 2, false, false, false);
  s->ops->cols(s, 2, force);
  s->phase = (s->phase + 1) % 3;
  if (dye) {
    prof_begin(s, S2T_PROF_DYE);
    dye_step(s->stream, s->dye, s->dye_parity, s->vel, s->dev, 0.0, s->phys.dye_fade, s->splats,
             s->n, s->m);
    prof_end(s);
    s->dye_parity = 1 - s->dye_parity;
  }
}

}  // namespace s2t

using namespace s2t;

#define S2T_API_BEGIN try {
#define S2T_API_END                         \
  return S2T_OK;                            \
  }                                         \
  catch (const Fail& f) {                   \
    g_err = f.what();                       \
    return f.code;                          \
  }                                         \
  catch (const std::exception& e) {         \
    g_err = e.what();                       \
    return S2T_EHIP;                        \
  }

static int bad(const std::string& m) {
  g_err = m;
  return S2T_EINVAL;
}

extern "C" {

const char* s2t_last_error(void) { return g_err.c_str(); }

int s2t_radix_schedule(int n, int* radix4_stages, int* has_radix2) {
  if (radix(n, radix4_stages, has_radix2) != 0)
    return bad("grid size " + std::to_string(n) + " must be 4^a or 2 * 4^a");
  return S2T_OK;
}

int s2t_create(int n, int dye_scale, s2t_solver** out) {
  if (!out) return bad("s2t_create: out is NULL");
  *out = nullptr;
  if (radix(n, nullptr, nullptr) != 0)
    return bad("grid size " + std::to_string(n) + " must be 4^a or 2 * 4^a");
  if (n < 16 || n > 4096) return bad("grid size " + std::to_string(n) + " outside 16..4096");
  if (dye_scale != 1 && dye_scale != 2) return bad("dye_scale must be 1 or 2");
  Solver* s = new Solver();
  S2T_API_BEGIN
  s->n = n;
  s->K = n / 3;
  s->dye_scale = dye_scale;
  s->m = n * dye_scale;
  s->ops = ops_for(n);
  hipchk(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "hipStreamCreate");
  for (int i = 0; i < 3; ++i) s->state[i] = dalloc<float2>(s->ops->state_elems, "state");
  s->mixed_u = dalloc<float2>(3 * s->ops->plane_elems, "mixed omega/u/v");
  s->mixed_z = dalloc<float2>(2 * s->ops->plane_elems, "mixed products");
  s->dev = dalloc<DevState>(1, "device scalars");
  const int nb = n / 2 + 1;
  s->diag_bins = dalloc<double>(2 * nb + 2, "diagnostic bins");
  s->diag_zonal = dalloc<double>(n, "zonal mean");
  s->seed_sum = dalloc<double>(s->K + 2, "seed sums");
  for (int i = 0; i < 3; ++i) s->dye[i] = dalloc<ushort4>((size_t)s->m * s->m, "dye");
  // twiddles in FP64, rounded once
  std::vector<float2> tw(s->ops->twiddle_elems > 0 ? s->ops->twiddle_elems : 1);
  s->ops->twiddles(tw.data());
  s->tw = dalloc<float2>(tw.size(), "twiddles");
  h2d(s, s->tw, tw.data(), tw.size() * sizeof(float2));
  std::vector<double2> wt(n);
  for (int m = 0; m < n; ++m) {
    const double a = 2.0 * M_PI * m / n;
    wt[m] = make_double2(std::cos(a), std::sin(a));
  }
  s->wtab = dalloc<double2>(n, "zonal table");
  h2d(s, s->wtab, wt.data(), n * sizeof(double2));
  s->params = default_params();
  s->phys = make_phys(n, s->params);
  for (int i = 0; i < 3; ++i) fill_u32(s->stream, s->state[i], s->ops->state_elems * 2, 0u);
  for (int i = 0; i < 3; ++i) fill_u32(s->stream, s->dye[i], (size_t)s->m * s->m * 2, 0u);
  fill_u32(s->stream, s->mixed_u, 3 * s->ops->plane_elems * 2, 0u);
  reset_clock_s(s);
  hipchk(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
  *out = reinterpret_cast<s2t_solver*>(s);
  return S2T_OK;
  }
  catch (const Fail& f) {
    g_err = f.what();
    destroy(s);
    return f.code;
  }
  catch (const std::exception& e) {
    g_err = e.what();
    destroy(s);
    return S2T_EHIP;
  }
}

int s2t_destroy(s2t_solver* h) {
  destroy(reinterpret_cast<Solver*>(h));
  return S2T_OK;
}

#define SOLVER(h)                                           \
  Solver* s = reinterpret_cast<Solver*>(h);                 \
  if (!s) return bad(std::string(__func__) + ": null solver")

int s2t_configure(s2t_solver* h, const s2t_params* p) {
  SOLVER(h);
  if (!p) return bad("s2t_configure: params is NULL");
  if (!(p->hyper_p > 0)) return bad("hyper_p must be positive");
  if (p->force_amp > 0 && !(p->force_width > 0)) return bad("force_width must be positive");
  S2T_API_BEGIN
  s->params = *p;
  s->phys = make_phys(s->n, *p);
  s->configured = true;
  S2T_API_END
}

int s2t_nu_h(s2t_solver* h, double* nu_h) {
  SOLVER(h);
  if (!nu_h) return bad("s2t_nu_h: out is NULL");
  *nu_h = s->phys.nu;
  return S2T_OK;
}

int s2t_set_splats(s2t_solver* h, const s2t_splat* sp, int count) {
  SOLVER(h);
  if (count < 0) return bad("splat count must be >= 0");
  if (count > 0 && !sp) return bad("splats is NULL");
  const int c = count < S2T_MAX_SPLATS ? count : S2T_MAX_SPLATS;
  Splats out{};
  out.count = c;
  for (int i = 0; i < c; ++i) {
    out.s[i] = Splat{sp[i].x,        sp[i].y,        sp[i].dx,     sp[i].dy,  sp[i].color[0],
                     sp[i].color[1], sp[i].color[2], sp[i].radius, sp[i].amp, sp[i].dye_amp};
  }
  s->splats = out;
  return S2T_OK;
}

int s2t_seed_spectrum(s2t_solver* h, double k0, double amplitude, uint32_t seed) {
  SOLVER(h);
  if (!(k0 > 0)) return bad("k0 must be positive");
  S2T_API_BEGIN
  s->phase = 0;
  seed_spectrum(s->stream, s->state[0], s->seed_sum, s->n, k0, amplitude, seed);
  for (int i = 1; i < 3; ++i) fill_u32(s->stream, s->state[i], s->ops->state_elems * 2, 0u);
  reset_clock_s(s);
  hipchk(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
  S2T_API_END
}

int s2t_seed_physical(s2t_solver* h, const float* field) {
  SOLVER(h);
  if (!field) return bad("field is NULL");
  S2T_API_BEGIN
  need_full(s);
  const size_t nn = (size_t)s->n * s->n;
  float* tmp = reinterpret_cast<float*>(s->full2);
  h2d(s, tmp, field, nn * sizeof(float));
  real_to_complex(s->stream, tmp, s->full, nn);
  s->ops->fft2(s, s->full, 0);
  s->phase = 0;
  s->ops->pack(s, s->full, s->state[0]);
  for (int i = 1; i < 3; ++i) fill_u32(s->stream, s->state[i], s->ops->state_elems * 2, 0u);
  reset_clock_s(s);
  hipchk(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
  S2T_API_END
}

int s2t_write_state(s2t_solver* h, const float* spectrum) {
  SOLVER(h);
  if (!spectrum) return bad("spectrum is NULL");
  S2T_API_BEGIN
  need_full(s);
  h2d(s, s->full, spectrum, (size_t)s->n * s->n * sizeof(float2));
  s->ops->pack(s, s->full, s->state[s->phase]);
  s->primed = false;
  hipchk(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
  S2T_API_END
}

int s2t_read_state(s2t_solver* h, float* spectrum) {
  SOLVER(h);
  if (!spectrum) return bad("spectrum is NULL");
  S2T_API_BEGIN
  need_full(s);
  s->ops->unpack(s, s->state[s->phase], s->full);
  d2h(s, spectrum, s->full, (size_t)s->n * s->n * sizeof(float2));
  S2T_API_END
}

int s2t_reset_clock(s2t_solver* h) {
  SOLVER(h);
  S2T_API_BEGIN
  reset_clock_s(s);
  hipchk(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
  S2T_API_END
}

int s2t_step(s2t_solver* h, int nsteps, uint32_t flags) {
  SOLVER(h);
  if (nsteps < 0) return bad("nsteps must be >= 0");
  S2T_API_BEGIN
  for (int i = 0; i < nsteps; ++i) step_once(s, (flags & S2T_STEP_DYE) != 0);
  S2T_API_END
}

int s2t_sync(s2t_solver* h) {
  SOLVER(h);
  S2T_API_BEGIN
  hipchk(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
  S2T_API_END
}

int s2t_diagnostics(s2t_solver* h, s2t_diag* out, double* spectrum, double* zonal) {
  SOLVER(h);
  S2T_API_BEGIN
  const int nb = s->n / 2 + 1;
  diagnostics(s->stream, s->state[s->phase], s->wtab, s->diag_bins, s->diag_zonal, s->n);
  std::vector<double> bins(2 * nb + 2);
  d2h(s, bins.data(), s->diag_bins, bins.size() * sizeof(double));
  DevState d{};
  d2h(s, &d, s->dev, sizeof(DevState));
  if (out) {
    out->energy = bins[2 * nb];
    out->enstrophy = bins[2 * nb + 1];
    out->dt = d.dt;
    out->max_speed = d.speed;
    out->time = d.t;
    out->step_index = d.step + 1u;
    out->nonfinite = d.nonfinite;
  }
  if (spectrum) memcpy(spectrum, bins.data(), nb * sizeof(double));
  if (zonal) d2h(s, zonal, s->diag_zonal, s->n * sizeof(double));
  S2T_API_END
}

int s2t_paint_dye(s2t_solver* h, int pattern) {
  SOLVER(h);
  if (pattern != 0 && pattern != 1) return bad("dye pattern must be 0 (none) or 1 (bands)");
  S2T_API_BEGIN
  if (pattern == 1) dye_paint(s->stream, s->dye[s->dye_parity], s->m);
  hipchk(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
  S2T_API_END
}

int s2t_dye_size(s2t_solver* h, int* m) {
  SOLVER(h);
  if (!m) return bad("m is NULL");
  *m = s->m;
  return S2T_OK;
}

int s2t_read_dye(s2t_solver* h, uint16_t* rgba) {
  SOLVER(h);
  if (!rgba) return bad("rgba is NULL");
  S2T_API_BEGIN
  d2h(s, rgba, s->dye[s->dye_parity], (size_t)s->m * s->m * sizeof(ushort4));
  S2T_API_END
}

int s2t_write_dye(s2t_solver* h, const uint16_t* rgba) {
  SOLVER(h);
  if (!rgba) return bad("rgba is NULL");
  S2T_API_BEGIN
  h2d(s, s->dye[s->dye_parity], rgba, (size_t)s->m * s->m * sizeof(ushort4));
  S2T_API_END
}

int s2t_write_velocity(s2t_solver* h, const float* uv) {
  SOLVER(h);
  if (!uv) return bad("uv is NULL");
  S2T_API_BEGIN
  need_vel(s);
  h2d(s, s->vel, uv, (size_t)s->n * s->n * sizeof(float2));
  S2T_API_END
}

int s2t_dye_step(s2t_solver* h, double dt) {
  SOLVER(h);
  S2T_API_BEGIN
  need_vel(s);
  dye_step(s->stream, s->dye, s->dye_parity, s->vel, nullptr, dt, s->phys.dye_fade, s->splats,
           s->n, s->m);
  s->dye_parity = 1 - s->dye_parity;
  hipchk(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
  S2T_API_END
}

int s2t_fft2(s2t_solver* h, const float* in, float* out, int inverse) {
  SOLVER(h);
  if (!in || !out) return bad("s2t_fft2: NULL array");
  S2T_API_BEGIN
  need_full(s);
  const size_t bytes = (size_t)s->n * s->n * sizeof(float2);
  h2d(s, s->full, in, bytes);
  s->ops->fft2(s, s->full, inverse ? 1 : 0);
  d2h(s, out, s->full, bytes);
  S2T_API_END
}

int s2t_velocity_pack(s2t_solver* h, const float* w, float* pack_a, float* pack_b) {
  SOLVER(h);
  if (!w || !pack_a || !pack_b) return bad("s2t_velocity_pack: NULL array");
  S2T_API_BEGIN
  need_full(s);
  const size_t nn = (size_t)s->n * s->n, bytes = nn * sizeof(float2);
  float2* a = s->full + nn;
  float2* b = s->full + 2 * nn;
  h2d(s, s->full, w, bytes);
  vpack_full(s->stream, s->full, a, b, s->n);
  d2h(s, pack_a, a, bytes);
  d2h(s, pack_b, b, bytes);
  S2T_API_END
}

int s2t_assemble(s2t_solver* h, const float* z, const float* fhat, uint32_t step_index, double dt,
                 float* rhs) {
  SOLVER(h);
  if (!z || !rhs) return bad("s2t_assemble: NULL array");
  S2T_API_BEGIN
  need_full(s);
  const size_t nn = (size_t)s->n * s->n, bytes = nn * sizeof(float2);
  float2* fh = nullptr;
  h2d(s, s->full, z, bytes);
  if (fhat) {
    fh = s->full + nn;
    h2d(s, fh, fhat, bytes);
  }
  float2* r = s->full + 2 * nn;
  assemble_full(s->stream, s->full, fh, r, s->phys, step_index, dt, s->n);
  d2h(s, rhs, r, bytes);
  S2T_API_END
}

int s2t_rhs(s2t_solver* h, uint32_t step_index, double dt, float* rhs) {
  SOLVER(h);
  if (!rhs) return bad("s2t_rhs: NULL array");
  S2T_API_BEGIN
  need_full(s);
  const bool force = s->splats.count > 0;
  if (force) need_force(s);
  ensure_primed(s);
  s->ops->rows(s, -1, false, false, force);
  float2* compact = s->full + 3 * (size_t)s->n * s->n;
  cols_rhs(s, force, compact, step_index, dt);
  s->ops->unpack(s, compact, s->full);
  d2h(s, rhs, s->full, (size_t)s->n * s->n * sizeof(float2));
  S2T_API_END
}

int s2t_profile(s2t_solver* h, int enable) {
  SOLVER(h);
  S2T_API_BEGIN
  if (enable) {
    hipchk(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
    s->nev = 0;
    for (int k = 0; k < S2T_PROF_KINDS; ++k) {
      s->prof_launches[k] = 0;
      s->prof_ms[k] = 0;
    }
  } else {
    prof_collect(s);
  }
  s->prof = enable != 0;
  S2T_API_END
}

int s2t_profile_read(s2t_solver* h, int kind, int* launches, double* total_ms) {
  SOLVER(h);
  if (kind < 0 || kind >= S2T_PROF_KINDS) return bad("profile kind out of range");
  S2T_API_BEGIN
  prof_collect(s);
  if (launches) *launches = s->prof_launches[kind];
  if (total_ms) *total_ms = s->prof_ms[kind];
  S2T_API_END
}

int s2t_kernel_bytes(s2t_solver* h, int kind, int stage, double* bytes) {
  SOLVER(h);
  if (!bytes) return bad("bytes is NULL");
  const double n = s->n, kc = s->K + 1, nj = 2 * s->K + 1;
  const double plane = n * kc * 8.0;  // one mixed field
  const double st = kc * nj * 8.0;    // one in-band half-spectrum field
  const bool force = s->splats.count > 0;
  double b = 0;
  if (kind == S2T_PROF_ROWS) {
    // read omega', u', v'; write P', Q' (+ force tile at stage 0)
    b = 3 * plane + 2 * plane + ((force && stage == 0) ? plane : 0);
  } else if (kind == S2T_PROF_COLS) {
    // read P', Q', w0 (+ wcur), write wout and omega', u', v'
    b = 2 * plane + st * (stage == 0 ? 2 : 3) + 3 * plane;
    if (force) b += stage == 0 ? plane + st : st;
  } else {
    return bad("no byte model for this kernel kind");
  }
  *bytes = b;
  return S2T_OK;
}

}  // extern "C"
