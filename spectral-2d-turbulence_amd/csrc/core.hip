// Slot-aware solver internals behind the C ABI: allocation, configuration, seeding, state I/O,
// diagnostics and step scheduling. Mirrors the buffer ownership and operations of the WebGPU
// Solver (src/gpu/solver.js:77-608).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "aux.hpp"
#include "core.hpp"

namespace s2t {

void hipchk(hipError_t e, const char* what) {
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

// ---------------------------------------------------------------------------------------------
// profiling: HIP events recorded on the solver's stream around each launch. They only time:
// without the system-scope fence a record does not write back and invalidate the caches, which
// would cost stream time and leave the next kernel's L2 cold
// ---------------------------------------------------------------------------------------------
// grow the profiling event list to at least `count` event pairs (creating events is host work of
// a few microseconds each: enabling profiling reserves a list, so that it stays out of the steps)
void prof_reserve(Solver* s, int count) {
  if (count <= s->cap_ev) return;
  const int cap = std::max(count, s->cap_ev * 2);
  Solver::Ev* ev = (Solver::Ev*)realloc(s->evs, sizeof(Solver::Ev) * cap);
  if (!ev) throw Fail(S2T_ENOMEM, "profiling event list");
  s->evs = ev;
  for (int i = s->cap_ev; i < cap; ++i) {
    hipchk(hipEventCreateWithFlags(&ev[i].a, hipEventDisableSystemFence), "hipEventCreate");
    hipchk(hipEventCreateWithFlags(&ev[i].b, hipEventDisableSystemFence), "hipEventCreate");
    s->cap_ev = i + 1;
  }
}

void prof_begin(Solver* s, int kind, hipStream_t st) {
  s->prof_open = ((s->prof_mask >> kind) & 1u) && !s->prof_skip;
  if (!s->prof_open) return;
  if (s->nev == s->cap_ev) prof_reserve(s, s->cap_ev ? s->cap_ev * 2 : 256);
  s->evs[s->nev].kind = kind;
  hipchk(hipEventRecord(s->evs[s->nev].a, st ? st : s->stream), "hipEventRecord");
}

void prof_end(Solver* s, hipStream_t st) {
  if (!s->prof_open) return;
  s->prof_open = false;
  hipchk(hipEventRecord(s->evs[s->nev].b, st ? st : s->stream), "hipEventRecord");
  s->nev += 1;
}

// ---------------------------------------------------------------------------------------------
// parameters
// ---------------------------------------------------------------------------------------------
int radix(int n, int* stages, int* r2) {
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

std::string size_error(int n) {
  if (radix(n, nullptr, nullptr) != 0)
    return "grid size " + std::to_string(n) + " must be 4^a or 2 * 4^a";  // fft.js:28
  if (n < 16 || n > 4096) return "grid size " + std::to_string(n) + " outside 16..4096";
  return "";
}

// defaults of the reference's config object (src/main.js:9-36)
s2t_params default_params() {
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

// The reference's controls (src/main.js:150-190) offer dissipation orders 1, 2, 4, beta <= 80,
// drag >= 0 and dtMax <= 0.05; the checks admit a superset that keeps the integrating factor's
// exponent non-positive and its phase |beta kx / k^2 dt / 2| below 1e4.
std::string check_params(const s2t_params& p) {
  if (!(p.hyper_p >= 1 && p.hyper_p <= 16 && std::floor(p.hyper_p) == p.hyper_p))
    return "hyper_p must be an integer in 1..16";
  if (!(p.drag >= 0)) return "drag must be non-negative";
  if (!(p.viscosity >= 0)) return "viscosity must be non-negative";
  if (p.force_amp > 0 && !(p.force_width > 0)) return "force_width must be positive";
  if (!(p.dt_max > 0)) return "dt_max must be positive";
  if (!(std::fabs(p.beta) * std::fmax(p.dt_max, p.fixed_dt) <= 1e4))
    return "|beta| * max(dt_max, fixed_dt) must not exceed 1e4";
  return "";
}

// Solver.configure (solver.js:306-324): nu_h = viscosity / kcut^(2p), FORCE_NORM = 0.01 N^2
static Phys make_phys(int n, const s2t_params& p) {
  Phys ph{};
  const double kcut = n / 3.0;
  ph.beta = p.beta;
  ph.mu = p.drag;
  ph.p = p.hyper_p;
  ph.nu = p.viscosity > 0 ? p.viscosity / std::pow(kcut, 2.0 * p.hyper_p) : 0.0;
  ph.p_int = (int)p.hyper_p;  // an integer in 1..16 (check_params)
  ph.force_amp = (float)p.force_amp;
  ph.force_k = (float)p.force_k;
  ph.force_width = (float)p.force_width;
  ph.force_norm = (float)(0.01 * n * n);  // src/gpu/shaders.js:26
  {
    const double lo = p.force_k - 3.05 * p.force_width, hi = p.force_k + 3.05 * p.force_width;
    ph.force_k2lo = lo > 0 ? (float)(lo * lo * (1.0 - 1e-5)) : -1.0f;
    ph.force_k2hi = (float)(hi * hi * (1.0 + 1e-5));
  }
  ph.cfl = p.cfl;
  ph.dt_min = p.dt_min;
  ph.dt_max = p.dt_max;
  ph.dx = 2.0 * M_PI / n;
  ph.fixed_dt = p.fixed_dt;
  ph.dye_fade = (float)p.dye_fade;
  return ph;
}

// ---------------------------------------------------------------------------------------------
// host <-> device helpers on the solver's stream
// ---------------------------------------------------------------------------------------------
static void h2d(Solver* s, void* dst, const void* src, size_t bytes) {
  hipchk(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s->stream), "hipMemcpyAsync H2D");
  hipchk(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
}
static void d2h(Solver* s, void* dst, const void* src, size_t bytes) {
  hipchk(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s->stream), "hipMemcpyAsync D2H");
  hipchk(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
}
void sync(Solver* s) { hipchk(hipStreamSynchronize(s->stream), "hipStreamSynchronize"); }

static void need_full(Solver* s) {
  // three full fields plus one compact state (the fused rhs hook)
  if (!s->full)
    s->full = dalloc<float2>((size_t)s->n * s->n * 3 + s->ops->state_elems, "full scratch");
  s->full2 = s->full + (size_t)s->n * s->n;
}
static void need_force(Solver* s) {
  if (!s->mixed_f) s->mixed_f = dalloc<float2>(s->ops->plane_elems * s->batch, "force tiles");
  if (!s->force_k) s->force_k = dalloc<float2>(s->ops->state_elems * s->batch, "force spectrum");
}
static void need_dye_stream(Solver* s) {
  if (s->dye_stream) return;
  hipchk(hipStreamCreateWithFlags(&s->dye_stream, hipStreamNonBlocking), "hipStreamCreate");
  // device-side ordering only: no timing, and an agent-scope release (no system-scope fence)
  const unsigned flags = hipEventDisableTiming | hipEventDisableSystemFence;
  hipchk(hipEventCreateWithFlags(&s->ev_cols0, flags), "hipEventCreate");
  hipchk(hipEventCreateWithFlags(&s->ev_dye, flags), "hipEventCreate");
}
static void need_vel(Solver* s) {
  if (!s->vel) {
    s->vel = dalloc<float2>((size_t)s->n * s->n, "velocity");
    fill_u32(s->stream, s->vel, (size_t)s->n * s->n * 2, 0u);
  }
}

static float2* slot_state(Solver* s, int slot, int phase = -1) {
  return s->state[phase >= 0 ? phase : s->phase] + (size_t)slot * s->ops->state_elems;
}

// ---------------------------------------------------------------------------------------------
// lifecycle (Solver.create / build / destroy, solver.js:71-140, 614-629)
// ---------------------------------------------------------------------------------------------
Solver* solver_new(int n, int dye_scale, int batch, bool with_dye, hipStream_t stream) {
  Solver* s = new Solver();
  try {
    s->n = n;
    s->K = n / 3;
    s->dye_scale = dye_scale;
    s->m = n * dye_scale;
    s->batch = batch;
    s->with_dye = with_dye;
    s->ops = ops_for(n);
    if (stream) {
      s->stream = stream;
      s->own_stream = false;
    } else {
      hipchk(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "hipStreamCreate");
    }
    const size_t st = s->ops->state_elems * batch, pl = s->ops->plane_elems * batch;
    for (int i = 0; i < 3; ++i) s->state[i] = dalloc<float2>(st, "state");
    s->mixed_u = dalloc<float2>(3 * pl, "mixed omega/u/v");
    s->mixed_z = dalloc<float2>(2 * pl, "mixed products");
    s->inst = dalloc<Inst>(batch, "instance scalars");
    const int nb = n / 2 + 1;
    s->diag_bins = dalloc<double>(2 * nb + 2, "diagnostic bins");
    s->diag_zonal = dalloc<double>(n, "zonal mean");
    s->seed_sum = dalloc<double>(s->K + 2, "seed sums");
    if (with_dye)
      for (int i = 0; i < 3; ++i) s->dye[i] = dalloc<ushort4>((size_t)s->m * s->m, "dye");
    // twiddles in FP64, rounded once
    std::vector<float2> tw(s->ops->twiddle_elems > 0 ? s->ops->twiddle_elems : 1);
    s->ops->twiddles(tw.data());
    s->tw = dalloc<float2>(tw.size(), "twiddles");
    h2d(s, s->tw, tw.data(), tw.size() * sizeof(float2));
    std::vector<double2> wt(n);
    for (int k = 0; k < n; ++k) {
      const double a = 2.0 * M_PI * k / n;
      wt[k] = make_double2(std::cos(a), std::sin(a));
    }
    s->wtab = dalloc<double2>(n, "zonal table");
    h2d(s, s->wtab, wt.data(), n * sizeof(double2));
    s->phys_slot.assign(batch, make_phys(n, default_params()));
    s->params_slot.assign(batch, default_params());
    s->primed_slot.assign(batch, 0);
    for (int i = 0; i < 3; ++i) fill_u32(s->stream, s->state[i], st * 2, 0u);
    if (with_dye)
      for (int i = 0; i < 3; ++i) fill_u32(s->stream, s->dye[i], (size_t)s->m * s->m * 2, 0u);
    fill_u32(s->stream, s->mixed_u, 3 * pl * 2, 0u);
    for (int i = 0; i < batch; ++i) {
      hipchk(hipMemcpyAsync(&s->inst[i].ph, &s->phys_slot[i], sizeof(Phys), hipMemcpyHostToDevice,
                            s->stream),
             "hipMemcpyAsync params");
    }
    reset_inst(s->stream, s->inst, 0, batch);
    sync(s);
    return s;
  } catch (...) {
    solver_free(s);
    throw;
  }
}

void solver_free(Solver* s) {
  if (!s) return;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  if (s->dye_stream) {
    (void)hipStreamSynchronize(s->dye_stream);
    (void)hipStreamDestroy(s->dye_stream);
  }
  if (s->ev_cols0) (void)hipEventDestroy(s->ev_cols0);
  if (s->ev_dye) (void)hipEventDestroy(s->ev_dye);
  void* bufs[] = {s->state[0], s->state[1], s->state[2], s->mixed_u,  s->mixed_z,    s->mixed_f,
                  s->force_k,  s->vel,      s->full,     s->tw,       s->wtab,       s->inst,
                  s->diag_bins, s->diag_zonal, s->seed_sum, s->dye[0], s->dye[1], s->dye[2]};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  if (s->enc_diag) (void)hipFree(s->enc_diag);
  if (s->enc_host) (void)hipHostFree(s->enc_host);
  if (s->enc_event) (void)hipEventDestroy(s->enc_event);
  for (int i = 0; i < s->cap_ev; ++i) {
    (void)hipEventDestroy(s->evs[i].a);
    (void)hipEventDestroy(s->evs[i].b);
  }
  free(s->evs);
  if (s->stream && s->own_stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

// ---------------------------------------------------------------------------------------------
// per-slot operations
// ---------------------------------------------------------------------------------------------
void configure_slot(Solver* s, int slot, const s2t_params& p) {
  s->params_slot[slot] = p;
  s->phys_slot[slot] = make_phys(s->n, p);
  h2d(s, &s->inst[slot].ph, &s->phys_slot[slot], sizeof(Phys));
}

void reset_slot(Solver* s, int slot) {
  reset_inst(s->stream, s->inst, slot, 1);
  s->primed_slot[slot] = 0;
}

// simulated time and step index of a slot (the forcing phases follow the step index), for resuming
// from a checkpoint; waits for queued work first
void set_clock_slot(Solver* s, int slot, double t, unsigned step) {
  h2d(s, &s->inst[slot].st.t, &t, sizeof(double));
  h2d(s, &s->inst[slot].st.step, &step, sizeof(unsigned));
}

void set_budget_slot(Solver* s, int slot, int steps) {
  h2d(s, &s->inst[slot].st.steps_left, &steps, sizeof(int));
}

DevState read_clock_slot(Solver* s, int slot) {
  DevState d{};
  d2h(s, &d, &s->inst[slot].st, sizeof(DevState));
  return d;
}

// Solver.seed via fillSpectrum (solver.js:347-359, 384-418), on the device
void seed_spectrum_slot(Solver* s, int slot, double k0, double amplitude, uint32_t seed) {
  seed_spectrum(s->stream, slot_state(s, slot), s->seed_sum, s->n, k0, amplitude, seed);
  reset_slot(s, slot);
  sync(s);
}

// Solver.seedPhysical (solver.js:361-382): forward FFT, 2/3 mask, k = 0 zeroed
void seed_physical_slot(Solver* s, int slot, const float* field) {
  need_full(s);
  const size_t nn = (size_t)s->n * s->n;
  float* tmp = reinterpret_cast<float*>(s->full2);
  h2d(s, tmp, field, nn * sizeof(float));
  real_to_complex(s->stream, tmp, s->full, nn);
  s->ops->fft2(s, s->full, 0);
  s->ops->pack(s, s->full, slot_state(s, slot));
  reset_slot(s, slot);
  sync(s);
}

// Seeding and admission for the serving loop, enqueued on the solver's stream with no host wait
// (behind whatever steps are queued there): the scenario built on the device, or fillSpectrum,
// into the slot's state; then its parameters, clock and step budget in one launch.
void seed_scenario_slot_async(Solver* s, int slot, int scenario, double param) {
  need_full(s);
  scenario_field(s->stream, s->full, s->n, scenario, param);
  s->ops->fft2(s, s->full, 0);
  s->ops->pack(s, s->full, slot_state(s, slot));
}

void seed_spectrum_slot_async(Solver* s, int slot, double k0, double amplitude, uint32_t seed) {
  seed_spectrum(s->stream, slot_state(s, slot), s->seed_sum, s->n, k0, amplitude, seed);
}

void admit_slot_async(Solver* s, int slot, const s2t_params& p, int steps) {
  s->params_slot[slot] = p;
  s->phys_slot[slot] = make_phys(s->n, p);
  admit_inst(s->stream, s->inst, slot, s->phys_slot[slot], steps);
  s->primed_slot[slot] = 0;
}

void write_state_slot(Solver* s, int slot, const float* spectrum) {
  need_full(s);
  h2d(s, s->full, spectrum, (size_t)s->n * s->n * sizeof(float2));
  s->ops->pack(s, s->full, slot_state(s, slot));
  s->primed_slot[slot] = 0;
  sync(s);
}

void read_state_slot(Solver* s, int slot, float* spectrum, int phase) {
  need_full(s);
  s->ops->unpack(s, slot_state(s, slot, phase), s->full);
  d2h(s, spectrum, s->full, (size_t)s->n * s->n * sizeof(float2));
}

// readDiagnostics (solver.js:547-608) in FP64 with a fixed reduction order
void diag_slot(Solver* s, int slot, s2t_diag* out, double* spectrum, double* zonal, int phase) {
  const int nb = s->n / 2 + 1;
  diagnostics(s->stream, slot_state(s, slot, phase), s->wtab, s->diag_bins, s->diag_zonal, s->n);
  std::vector<double> bins(2 * nb + 2);
  d2h(s, bins.data(), s->diag_bins, bins.size() * sizeof(double));
  const DevState d = read_clock_slot(s, slot);
  if (out) {
    out->energy = bins[2 * nb];
    out->enstrophy = bins[2 * nb + 1];
    out->dt = d.dt;
    out->max_speed = d.speed;
    out->time = d.t;
    out->step_index = d.step;
    out->nonfinite = d.nonfinite;
  }
  if (spectrum) memcpy(spectrum, bins.data(), nb * sizeof(double));
  if (zonal) d2h(s, zonal, s->diag_zonal, s->n * sizeof(double));
}

// Solver.encodeDiagnostics (solver.js:547-558): the FP64 diagnostics of slot 0's current state
// and a snapshot of its clock, into the solver's device record, enqueued only. Layout of the
// record: bins (2 nb + 2 doubles: E(k), Z(k), E, Z), the zonal mean (n doubles), the DevState.
static size_t enc_bytes(const Solver* s) {
  const int nb = s->n / 2 + 1;
  return (size_t)(2 * nb + 2 + s->n) * sizeof(double) + sizeof(DevState);
}

void encode_diag(Solver* s) {
  const int nb = s->n / 2 + 1;
  if (!s->enc_diag) s->enc_diag = dalloc<char>(enc_bytes(s), "diagnostics record");
  double* bins = reinterpret_cast<double*>(s->enc_diag);
  double* zonal = bins + 2 * nb + 2;
  diagnostics(s->stream, slot_state(s, 0), s->wtab, bins, zonal, s->n);
  hipchk(hipMemcpyAsync(zonal + s->n, &s->inst[0].st, sizeof(DevState), hipMemcpyDeviceToDevice,
                        s->stream),
         "hipMemcpyAsync clock snapshot");
}

// Solver.copyDiagnostics (solver.js:560-568): the last encoded record to pinned host memory,
// enqueued; an event marks its arrival
void copy_diag(Solver* s) {
  if (!s->enc_diag) throw Fail(S2T_ESTATE, "copy_diagnostics: no diagnostics were encoded");
  if (!s->enc_host) {
    void* p = nullptr;
    hipchk(hipHostMalloc(&p, enc_bytes(s), 0), "hipHostMalloc diagnostics staging");
    s->enc_host = static_cast<char*>(p);
  }
  if (!s->enc_event)
    hipchk(hipEventCreateWithFlags(&s->enc_event, hipEventDisableTiming), "hipEventCreate");
  hipchk(hipMemcpyAsync(s->enc_host, s->enc_diag, enc_bytes(s), hipMemcpyDeviceToHost, s->stream),
         "hipMemcpyAsync diagnostics record");
  hipchk(hipEventRecord(s->enc_event, s->stream), "hipEventRecord");
  s->enc_copied = true;
}

// readDiagnostics (solver.js:570-608) of the copied record: waits for the copy only
void read_copied_diag(Solver* s, s2t_diag* out, double* spectrum, double* zonal) {
  if (!s->enc_copied) throw Fail(S2T_ESTATE, "read_copied_diagnostics: nothing was copied");
  hipchk(hipEventSynchronize(s->enc_event), "hipEventSynchronize");
  s->enc_copied = false;
  const int nb = s->n / 2 + 1;
  const double* bins = reinterpret_cast<const double*>(s->enc_host);
  const double* zon = bins + 2 * nb + 2;
  DevState d;
  memcpy(&d, zon + s->n, sizeof(DevState));
  if (out) {
    out->energy = bins[2 * nb];
    out->enstrophy = bins[2 * nb + 1];
    out->dt = d.dt;
    out->max_speed = d.speed;
    out->time = d.t;
    out->step_index = d.step;
    out->nonfinite = d.nonfinite;
  }
  if (spectrum) memcpy(spectrum, bins, nb * sizeof(double));
  if (zonal) memcpy(zonal, zon, s->n * sizeof(double));
}

// ---------------------------------------------------------------------------------------------
// stepping
// ---------------------------------------------------------------------------------------------
// the velocity prime (encodePrime, solver.js:475-486) of every slot whose state was replaced
void ensure_primed(Solver* s) {
  int i = 0;
  while (i < s->batch) {
    if (s->primed_slot[i]) {
      ++i;
      continue;
    }
    int j = i;
    while (j < s->batch && !s->primed_slot[j]) ++j;
    s->ops->prime(s, i, j - i);
    for (int k = i; k < j; ++k) s->primed_slot[k] = 1;
    i = j;
  }
}

// encodeStep (+ encodeReduce, + encodeDye) of solver.js:488-545 for all active slots
void step_once(Solver* s, bool dye) {
  s->prof_skip = s->prof_every > 1 && s->prof_step % s->prof_every != 0;
  s->prof_step += 1;
  const bool force = s->splats.count > 0;
  if (force) need_force(s);
  if (dye) {
    need_vel(s);
    need_dye_stream(s);
  }
  ensure_primed(s);
  const bool sd = s->ops->rows(s, 0, true, dye, force);
  s->ops->cols(s, 0, force, sd);
  if (dye) {
    // The BFECC dye step (encodeDye, after the step in solver.js) needs the stage-0 velocity
    // and the step's dt, both final once the stage-0 column kernel has run: it runs on its own
    // stream beside stages 1 and 2, which write neither (the stage-2 epilogue advances the clock
    // but not dt), and fills the device where those launches leave it idle.
    hipchk(hipEventRecord(s->ev_cols0, s->stream), "hipEventRecord");
    hipchk(hipStreamWaitEvent(s->dye_stream, s->ev_cols0, 0), "hipStreamWaitEvent");
    prof_begin(s, S2T_PROF_DYE, s->dye_stream);
    dye_step(s->dye_stream, s->dye, s->dye_parity, s->vel, &s->inst[0].st, 0.0,
             s->phys_slot[0].dye_fade, s->splats, s->n, s->m);
    prof_end(s, s->dye_stream);
    hipchk(hipEventRecord(s->ev_dye, s->dye_stream), "hipEventRecord");
    s->dye_parity = 1 - s->dye_parity;
  }
  s->ops->rows(s, 1, false, false, false);
  s->ops->cols(s, 1, force, false);
  s->ops->rows(s, 2, false, false, false);
  s->ops->cols(s, 2, force, false);
  s->phase = (s->phase + 1) % 3;
  // whatever comes next on the solver's stream (the next step's stage-0 velocity, reads of the
  // dye) follows the dye step
  if (dye) hipchk(hipStreamWaitEvent(s->stream, s->ev_dye, 0), "hipStreamWaitEvent");
  s->prof_skip = false;
}

// Column-group width of the step's column kernel for this solver: each compiled width is timed
// over `steps` steps (after two untimed ones) of a scratch solver of the same size, slot count and
// slot parameters, seeded with turbulence; the fastest is kept. The solver's own buffers are not
// touched, and every width gives bit-identical results, so tuning never changes a result.
// ms[i] receives each width's milliseconds per step (in the order of widths[]).
int autotune_cols(Solver* s, int steps, int* widths, double* ms, int* count) {
  int w[MAX_COL_WIDTHS];
  const int nw = s->ops->col_widths(w);
  Solver* t = solver_new(s->n, 1, s->batch, false);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int best = w[0];
  try {
    for (int k = 0; k < s->batch; ++k) {
      configure_slot(t, k, s->params_slot[k]);
      seed_spectrum_slot(t, k, std::max(2.0, s->n / 32.0), 6.0, 1u + (uint32_t)k);
    }
    hipchk(hipEventCreate(&e0), "hipEventCreate");
    hipchk(hipEventCreate(&e1), "hipEventCreate");
    double best_ms = 0.0;
    for (int i = 0; i < nw; ++i) {
      t->cols_c = w[i];
      step_once(t, false);
      step_once(t, false);
      hipchk(hipEventRecord(e0, t->stream), "hipEventRecord");
      for (int k = 0; k < steps; ++k) step_once(t, false);
      hipchk(hipEventRecord(e1, t->stream), "hipEventRecord");
      hipchk(hipEventSynchronize(e1), "hipEventSynchronize");
      float f = 0.f;
      hipchk(hipEventElapsedTime(&f, e0, e1), "hipEventElapsedTime");
      const double per = (double)f / steps;
      if (widths) widths[i] = w[i];
      if (ms) ms[i] = per;
      if (i == 0 || per < best_ms) {
        best_ms = per;
        best = w[i];
      }
    }
  } catch (...) {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    solver_free(t);
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  solver_free(t);
  if (count) *count = nw;
  s->cols_c = best;
  return best;
}

// Column-group width for `instances` instances of one size stepped by the same per-class launches
// (a batch class, a serving chunk's busy slots): the tuner's choices for one-class batches. Four
// columns per workgroup once their workgroups fill the chip (2048^2 from four instances, 1024^2
// from eight, 512^2 from sixteen; 128^2 and 256^2, which have no field-parallel kernel, from two),
// two columns for two or three 2048^2 instances (9.5% and 7.3% faster than four), else 0: the
// size's single-instance kernel (field-parallel at 512^2 to 2048^2).
int class_width(const Solver* s, int instances) {
  int w[MAX_COL_WIDTHS];
  const int nw = s->ops->col_widths(w);
  auto has = [&](int x) {
    for (int i = 0; i < nw; ++i)
      if (w[i] == x) return true;
    return false;
  };
  const int n = s->n;
  const int full = n >= 2048 ? 4 : n == 1024 ? 8 : n == 512 ? 16 : 2;
  if (instances >= full && has(4)) return 4;
  if (n == 2048 && instances >= 2 && has(2)) return 2;
  return 0;
}

// Default width of a batch class (a class of one member keeps the size's default)
void default_class_width(Solver* s, int members) {
  if (members >= 2) s->cols_c = class_width(s, members);
}

void set_col_width(Solver* s, int width) {
  int w[MAX_COL_WIDTHS];
  const int nw = s->ops->col_widths(w);
  if (width == 0) {
    s->cols_c = 0;
    return;
  }
  for (int i = 0; i < nw; ++i)
    if (w[i] == width) {
      s->cols_c = width;
      return;
    }
  throw Fail(S2T_EINVAL, "column-group width " + std::to_string(width) +
                             " is not compiled for grid size " + std::to_string(s->n));
}

// Band-limited random vorticity (host, FP64, rounded once to FP32). Per canonical mode with
// kmin <= |k| <= kmax in the 2/3 band: three uniforms u_j = (hash32((3 idx + j) ^ hash32(seed))
// + 1/2) / 2^32, idx = row * n + col of the canonical mode; g = sqrt(-2 ln u0) cos(2 pi u1)
// (Box-Muller), phase 2 pi u2, value g (cos, sgn sin) with sgn = -1 on the mirror; then every
// value is scaled by amplitude / (sqrt(sum g^2) / n^2), the rms of the physical field.
void bandlimited_spectrum(int n, double kmin, double kmax, double amplitude, uint32_t seed,
                          float* out) {
  const double kcut = n / 3.0;
  const uint32_t salt = hash1(seed);
  const double two_pi = 6.283185307179586476925286766559;
  const size_t nn = (size_t)n * n;
  std::vector<double> re(nn, 0.0), im(nn, 0.0);
  double total = 0.0;
  for (int row = 0; row < n; ++row) {
    const int ky = row <= n / 2 ? row : row - n;
    for (int col = 0; col < n; ++col) {
      const int kx = col <= n / 2 ? col : col - n;
      if (!(std::fabs((double)kx) < kcut && std::fabs((double)ky) < kcut)) continue;
      const double k2 = (double)kx * kx + (double)ky * ky;
      const double kr = std::sqrt(k2);
      if (k2 == 0.0 || kr < kmin || kr > kmax) continue;
      const bool canon = ky > 0 || (ky == 0 && kx > 0);
      const int hr = canon ? row : (n - row) % n, hc = canon ? col : (n - col) % n;
      const uint32_t idx = (uint32_t)hr * (uint32_t)n + (uint32_t)hc;
      double u[3];
      for (int j = 0; j < 3; ++j)
        u[j] = ((double)hash1((3u * idx + (uint32_t)j) ^ salt) + 0.5) / 4294967296.0;
      const double g = std::sqrt(-2.0 * std::log(u[0])) * std::cos(two_pi * u[1]);
      const double ph = two_pi * u[2];
      const size_t i = (size_t)row * n + col;
      re[i] = g * std::cos(ph);
      im[i] = (canon ? 1.0 : -1.0) * g * std::sin(ph);
      total += g * g;
    }
  }
  const double rms = std::sqrt(total) / ((double)n * n);
  const double scale = rms > 0.0 ? amplitude / rms : 0.0;
  for (size_t i = 0; i < nn; ++i) {
    out[2 * i] = (float)(re[i] * scale);
    out[2 * i + 1] = (float)(im[i] * scale);
  }
}

}  // namespace s2t
