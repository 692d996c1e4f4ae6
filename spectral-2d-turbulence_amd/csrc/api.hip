// C ABI of a single solver (include/s2t.h): a Solver with one slot. Mirrors the WebGPU Solver of
// src/gpu/solver.js; the batch and serving entry points are in batch.hip.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/s2t.h"
#include "aux.hpp"
#include "core.hpp"

namespace s2t {

thread_local std::string g_err;

static void prof_collect(Solver* s) {
  if (!s->nev) return;
  sync(s);
  for (int i = 0; i < s->nev; ++i) {
    float ms = 0.f;
    hipchk(hipEventElapsedTime(&ms, s->evs[i].a, s->evs[i].b), "hipEventElapsedTime");
    s->prof_launches[s->evs[i].kind] += 1;
    s->prof_ms[s->evs[i].kind] += ms;
  }
  s->nev = 0;
}

int api_bad(const std::string& m) {
  g_err = m;
  return S2T_EINVAL;
}

}  // namespace s2t

using namespace s2t;

#define S2T_API_BEGIN try {
#define S2T_API_END                 \
  return S2T_OK;                    \
  }                                 \
  catch (const Fail& f) {           \
    g_err = f.what();               \
    return f.code;                  \
  }                                 \
  catch (const std::exception& e) { \
    g_err = e.what();               \
    return S2T_EHIP;                \
  }

#define SOLVER(h)                             \
  Solver* s = reinterpret_cast<Solver*>(h); \
  if (!s) return api_bad(std::string(__func__) + ": null solver")

extern "C" {

const char* s2t_last_error(void) { return g_err.c_str(); }

int s2t_radix_schedule(int n, int* radix4_stages, int* has_radix2) {
  if (radix(n, radix4_stages, has_radix2) != 0)
    return api_bad("grid size " + std::to_string(n) + " must be 4^a or 2 * 4^a");
  return S2T_OK;
}

int s2t_create(int n, int dye_scale, s2t_solver** out) {
  if (!out) return api_bad("s2t_create: out is NULL");
  *out = nullptr;
  const std::string e = size_error(n);
  if (!e.empty()) return api_bad(e);
  if (dye_scale != 1 && dye_scale != 2) return api_bad("dye_scale must be 1 or 2");
  S2T_API_BEGIN
  *out = reinterpret_cast<s2t_solver*>(solver_new(n, dye_scale, 1, true));
  S2T_API_END
}

int s2t_destroy(s2t_solver* h) {
  solver_free(reinterpret_cast<Solver*>(h));
  return S2T_OK;
}

int s2t_configure(s2t_solver* h, const s2t_params* p) {
  SOLVER(h);
  if (!p) return api_bad("s2t_configure: params is NULL");
  const std::string e = check_params(*p);
  if (!e.empty()) return api_bad(e);
  S2T_API_BEGIN
  configure_slot(s, 0, *p);
  S2T_API_END
}

int s2t_nu_h(s2t_solver* h, double* nu_h) {
  SOLVER(h);
  if (!nu_h) return api_bad("s2t_nu_h: out is NULL");
  *nu_h = s->phys_slot[0].nu;
  return S2T_OK;
}

int s2t_set_splats(s2t_solver* h, const s2t_splat* sp, int count) {
  SOLVER(h);
  if (count < 0) return api_bad("splat count must be >= 0");
  if (count > 0 && !sp) return api_bad("splats is NULL");
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
  if (!(k0 > 0)) return api_bad("k0 must be positive");
  S2T_API_BEGIN
  seed_spectrum_slot(s, 0, k0, amplitude, seed);
  S2T_API_END
}

int s2t_fill_spectrum_noalloc(int n, double k0, double amplitude, uint32_t seed,
                              float* spectrum_dev, void* stream) {
  const std::string e = size_error(n);
  if (n > noalloc_max_n())
    return api_bad("s2t_fill_spectrum_noalloc: grid size " + std::to_string(n) +
                   " exceeds its fixed scratch (at most " + std::to_string(noalloc_max_n()) + ")");
  if (!e.empty()) return api_bad(e);
  if (!(k0 > 0)) return api_bad("k0 must be positive");
  if (!spectrum_dev) return api_bad("s2t_fill_spectrum_noalloc: spectrum_dev is NULL");
  S2T_API_BEGIN
  fill_spectrum_noalloc(static_cast<hipStream_t>(stream), reinterpret_cast<float2*>(spectrum_dev),
                        n, k0, amplitude, seed);
  S2T_API_END
}

int s2t_seed_physical(s2t_solver* h, const float* field) {
  SOLVER(h);
  if (!field) return api_bad("field is NULL");
  S2T_API_BEGIN
  seed_physical_slot(s, 0, field);
  S2T_API_END
}

int s2t_bandlimited_spectrum(int n, double kmin, double kmax, double amplitude, uint32_t seed,
                             float* spectrum) {
  if (!spectrum) return api_bad("spectrum is NULL");
  const std::string e = size_error(n);
  if (!e.empty()) return api_bad(e);
  if (!(kmin >= 0.0) || !(kmax >= kmin)) return api_bad("need 0 <= kmin <= kmax");
  if (!(amplitude >= 0.0)) return api_bad("amplitude must be non-negative");
  try {
    bandlimited_spectrum(n, kmin, kmax, amplitude, seed, spectrum);
  } catch (const std::bad_alloc&) {
    g_err = "s2t_bandlimited_spectrum: out of host memory";
    return S2T_ENOMEM;
  }
  return S2T_OK;
}

int s2t_seed_bandlimited(s2t_solver* h, double kmin, double kmax, double amplitude,
                         uint32_t seed) {
  SOLVER(h);
  if (!(kmin >= 0.0) || !(kmax >= kmin)) return api_bad("need 0 <= kmin <= kmax");
  if (!(amplitude >= 0.0)) return api_bad("amplitude must be non-negative");
  S2T_API_BEGIN
  std::vector<float> spec((size_t)2 * s->n * s->n);
  bandlimited_spectrum(s->n, kmin, kmax, amplitude, seed, spec.data());
  write_state_slot(s, 0, spec.data());
  reset_slot(s, 0);
  sync(s);
  S2T_API_END
}

int s2t_write_state(s2t_solver* h, const float* spectrum) {
  SOLVER(h);
  if (!spectrum) return api_bad("spectrum is NULL");
  S2T_API_BEGIN
  write_state_slot(s, 0, spectrum);
  S2T_API_END
}

int s2t_read_state(s2t_solver* h, float* spectrum) {
  SOLVER(h);
  if (!spectrum) return api_bad("spectrum is NULL");
  S2T_API_BEGIN
  read_state_slot(s, 0, spectrum);
  S2T_API_END
}

int s2t_reset_clock(s2t_solver* h) {
  SOLVER(h);
  S2T_API_BEGIN
  reset_slot(s, 0);
  sync(s);
  S2T_API_END
}

int s2t_set_clock(s2t_solver* h, double time, uint32_t step_index) {
  SOLVER(h);
  if (!std::isfinite(time)) return api_bad("s2t_set_clock: time must be finite");
  S2T_API_BEGIN
  set_clock_slot(s, 0, time, step_index);
  S2T_API_END
}

int s2t_step(s2t_solver* h, int nsteps, uint32_t flags) {
  SOLVER(h);
  if (nsteps < 0) return api_bad("nsteps must be >= 0");
  S2T_API_BEGIN
  for (int i = 0; i < nsteps; ++i) step_once(s, (flags & S2T_STEP_DYE) != 0);
  S2T_API_END
}

int s2t_sync(s2t_solver* h) {
  SOLVER(h);
  S2T_API_BEGIN
  sync(s);
  S2T_API_END
}

int s2t_diagnostics(s2t_solver* h, s2t_diag* out, double* spectrum, double* zonal) {
  SOLVER(h);
  S2T_API_BEGIN
  diag_slot(s, 0, out, spectrum, zonal);
  S2T_API_END
}

int s2t_paint_dye(s2t_solver* h, int pattern) {
  SOLVER(h);
  if (pattern != 0 && pattern != 1) return api_bad("dye pattern must be 0 (none) or 1 (bands)");
  S2T_API_BEGIN
  if (pattern == 1) dye_paint(s->stream, s->dye[s->dye_parity], s->m);
  sync(s);
  S2T_API_END
}

int s2t_dye_size(s2t_solver* h, int* m) {
  SOLVER(h);
  if (!m) return api_bad("m is NULL");
  *m = s->m;
  return S2T_OK;
}

int s2t_read_dye(s2t_solver* h, uint16_t* rgba) {
  SOLVER(h);
  if (!rgba) return api_bad("rgba is NULL");
  S2T_API_BEGIN
  hipchk(hipMemcpyAsync(rgba, s->dye[s->dye_parity], (size_t)s->m * s->m * sizeof(ushort4),
                        hipMemcpyDeviceToHost, s->stream),
         "hipMemcpyAsync D2H");
  sync(s);
  S2T_API_END
}

int s2t_write_dye(s2t_solver* h, const uint16_t* rgba) {
  SOLVER(h);
  if (!rgba) return api_bad("rgba is NULL");
  S2T_API_BEGIN
  hipchk(hipMemcpyAsync(s->dye[s->dye_parity], rgba, (size_t)s->m * s->m * sizeof(ushort4),
                        hipMemcpyHostToDevice, s->stream),
         "hipMemcpyAsync H2D");
  sync(s);
  S2T_API_END
}

int s2t_write_velocity(s2t_solver* h, const float* uv) {
  SOLVER(h);
  if (!uv) return api_bad("uv is NULL");
  S2T_API_BEGIN
  if (!s->vel) {
    hipchk(hipMalloc(&s->vel, (size_t)s->n * s->n * sizeof(float2)), "velocity");
  }
  hipchk(hipMemcpyAsync(s->vel, uv, (size_t)s->n * s->n * sizeof(float2), hipMemcpyHostToDevice,
                        s->stream),
         "hipMemcpyAsync H2D");
  sync(s);
  S2T_API_END
}

int s2t_dye_step(s2t_solver* h, double dt) {
  SOLVER(h);
  if (!s->vel) return api_bad("s2t_dye_step: no velocity (s2t_write_velocity or a dye step)");
  if (!std::isfinite(dt)) return api_bad("s2t_dye_step: dt must be finite");
  S2T_API_BEGIN
  // dt <= 0: the step of the last solver step, read by the kernels on the device
  dye_step(s->stream, s->dye, s->dye_parity, s->vel, dt > 0 ? nullptr : &s->inst[0].st, dt,
           s->phys_slot[0].dye_fade, s->splats, s->n, s->m);
  s->dye_parity = 1 - s->dye_parity;
  S2T_API_END
}

int s2t_dye_step_passes(s2t_solver* h, double dt) {
  SOLVER(h);
  if (!s->vel) return api_bad("s2t_dye_step_passes: no velocity (s2t_write_velocity or a dye step)");
  if (!std::isfinite(dt)) return api_bad("s2t_dye_step_passes: dt must be finite");
  S2T_API_BEGIN
  dye_step_passes(s->stream, s->dye, s->dye_parity, s->vel, dt > 0 ? nullptr : &s->inst[0].st,
                  dt, s->phys_slot[0].dye_fade, s->splats, s->n, s->m);
  s->dye_parity = 1 - s->dye_parity;
  S2T_API_END
}

int s2t_reset_dye(s2t_solver* h) {
  SOLVER(h);
  S2T_API_BEGIN
  for (int i = 0; i < 3; ++i)
    if (s->dye[i]) fill_u32(s->stream, s->dye[i], (size_t)s->m * s->m * 2, 0u);
  s->dye_parity = 0;
  sync(s);
  S2T_API_END
}

int s2t_prime(s2t_solver* h) {
  SOLVER(h);
  S2T_API_BEGIN
  ensure_primed(s);
  S2T_API_END
}

int s2t_primed(s2t_solver* h, int* primed) {
  SOLVER(h);
  if (!primed) return api_bad("s2t_primed: out is NULL");
  *primed = s->primed_slot[0] ? 1 : 0;
  return S2T_OK;
}

int s2t_encode_diagnostics(s2t_solver* h) {
  SOLVER(h);
  S2T_API_BEGIN
  encode_diag(s);
  S2T_API_END
}

int s2t_copy_diagnostics(s2t_solver* h) {
  SOLVER(h);
  S2T_API_BEGIN
  copy_diag(s);
  S2T_API_END
}

int s2t_read_copied_diagnostics(s2t_solver* h, s2t_diag* out, double* spectrum, double* zonal) {
  SOLVER(h);
  S2T_API_BEGIN
  read_copied_diag(s, out, spectrum, zonal);
  S2T_API_END
}

int s2t_buffer_create(size_t bytes, void** dev) {
  if (!dev) return api_bad("s2t_buffer_create: out is NULL");
  *dev = nullptr;
  if (bytes == 0) return api_bad("s2t_buffer_create: size must be positive");
  S2T_API_BEGIN
  hipchk(hipMalloc(dev, bytes), "hipMalloc buffer");
  hipchk(hipMemset(*dev, 0, bytes), "hipMemset buffer");
  S2T_API_END
}

int s2t_buffer_destroy(void* dev) {
  if (dev) (void)hipFree(dev);
  return S2T_OK;
}

int s2t_buffer_write(void* dev, size_t offset, const void* src, size_t bytes) {
  if (!dev || (!src && bytes)) return api_bad("s2t_buffer_write: NULL pointer");
  S2T_API_BEGIN
  hipchk(hipMemcpy(static_cast<char*>(dev) + offset, src, bytes, hipMemcpyHostToDevice),
         "hipMemcpy H2D");
  S2T_API_END
}

int s2t_buffer_read(const void* dev, size_t offset, void* dst, size_t bytes) {
  if (!dev || (!dst && bytes)) return api_bad("s2t_buffer_read: NULL pointer");
  S2T_API_BEGIN
  hipchk(hipMemcpy(dst, static_cast<const char*>(dev) + offset, bytes, hipMemcpyDeviceToHost),
         "hipMemcpy D2H");
  S2T_API_END
}

int s2t_buffer_copy(void* dst, size_t dst_offset, const void* src, size_t src_offset,
                    size_t bytes) {
  if (!dst || !src) return api_bad("s2t_buffer_copy: NULL pointer");
  S2T_API_BEGIN
  hipchk(hipMemcpy(static_cast<char*>(dst) + dst_offset,
                   static_cast<const char*>(src) + src_offset, bytes, hipMemcpyDeviceToDevice),
         "hipMemcpy D2D");
  S2T_API_END
}

int s2t_fft2_device(s2t_solver* h, float* data_dev, int inverse) {
  SOLVER(h);
  if (!data_dev) return api_bad("s2t_fft2_device: NULL buffer");
  S2T_API_BEGIN
  s->ops->fft2(s, reinterpret_cast<float2*>(data_dev), inverse ? 1 : 0);
  S2T_API_END
}

int s2t_fft2(s2t_solver* h, const float* in, float* out, int inverse) {
  SOLVER(h);
  if (!in || !out) return api_bad("s2t_fft2: NULL array");
  S2T_API_BEGIN
  if (!s->full)
    hipchk(hipMalloc(&s->full, ((size_t)s->n * s->n * 3 + s->ops->state_elems) * sizeof(float2)),
           "full scratch");
  const size_t bytes = (size_t)s->n * s->n * sizeof(float2);
  hipchk(hipMemcpyAsync(s->full, in, bytes, hipMemcpyHostToDevice, s->stream), "H2D");
  s->ops->fft2(s, s->full, inverse ? 1 : 0);
  hipchk(hipMemcpyAsync(out, s->full, bytes, hipMemcpyDeviceToHost, s->stream), "D2H");
  sync(s);
  S2T_API_END
}

int s2t_velocity_pack(s2t_solver* h, const float* w, float* pack_a, float* pack_b) {
  SOLVER(h);
  if (!w || !pack_a || !pack_b) return api_bad("s2t_velocity_pack: NULL array");
  S2T_API_BEGIN
  if (!s->full)
    hipchk(hipMalloc(&s->full, ((size_t)s->n * s->n * 3 + s->ops->state_elems) * sizeof(float2)),
           "full scratch");
  const size_t nn = (size_t)s->n * s->n, bytes = nn * sizeof(float2);
  float2* a = s->full + nn;
  float2* b = s->full + 2 * nn;
  hipchk(hipMemcpyAsync(s->full, w, bytes, hipMemcpyHostToDevice, s->stream), "H2D");
  vpack_full(s->stream, s->full, a, b, s->n);
  hipchk(hipMemcpyAsync(pack_a, a, bytes, hipMemcpyDeviceToHost, s->stream), "D2H");
  hipchk(hipMemcpyAsync(pack_b, b, bytes, hipMemcpyDeviceToHost, s->stream), "D2H");
  sync(s);
  S2T_API_END
}

int s2t_assemble(s2t_solver* h, const float* z, const float* fhat, uint32_t step_index, double dt,
                 float* rhs) {
  SOLVER(h);
  if (!z || !rhs) return api_bad("s2t_assemble: NULL array");
  S2T_API_BEGIN
  if (!s->full)
    hipchk(hipMalloc(&s->full, ((size_t)s->n * s->n * 3 + s->ops->state_elems) * sizeof(float2)),
           "full scratch");
  const size_t nn = (size_t)s->n * s->n, bytes = nn * sizeof(float2);
  float2* fh = nullptr;
  hipchk(hipMemcpyAsync(s->full, z, bytes, hipMemcpyHostToDevice, s->stream), "H2D");
  if (fhat) {
    fh = s->full + nn;
    hipchk(hipMemcpyAsync(fh, fhat, bytes, hipMemcpyHostToDevice, s->stream), "H2D");
  }
  float2* r = s->full + 2 * nn;
  assemble_full(s->stream, s->full, fh, r, s->phys_slot[0], step_index, dt, s->n);
  hipchk(hipMemcpyAsync(rhs, r, bytes, hipMemcpyDeviceToHost, s->stream), "D2H");
  sync(s);
  S2T_API_END
}

int s2t_rhs(s2t_solver* h, uint32_t step_index, double dt, float* rhs) {
  SOLVER(h);
  if (!rhs) return api_bad("s2t_rhs: NULL array");
  S2T_API_BEGIN
  if (!s->full)
    hipchk(hipMalloc(&s->full, ((size_t)s->n * s->n * 3 + s->ops->state_elems) * sizeof(float2)),
           "full scratch");
  const bool force = s->splats.count > 0;
  if (force) {
    if (!s->mixed_f)
      hipchk(hipMalloc(&s->mixed_f, s->ops->plane_elems * sizeof(float2)), "force tiles");
    if (!s->force_k)
      hipchk(hipMalloc(&s->force_k, s->ops->state_elems * sizeof(float2)), "force spectrum");
  }
  ensure_primed(s);
  // the planes of a primed (or stepped) state are in the stage-0 form
  const bool sd = s->ops->rows(s, 0, false, false, force);
  float2* compact = s->full + 3 * (size_t)s->n * s->n;
  s->ops->rhs(s, force, sd, compact, step_index, dt);
  s->ops->unpack(s, compact, s->full);
  hipchk(hipMemcpyAsync(rhs, s->full, (size_t)s->n * s->n * sizeof(float2), hipMemcpyDeviceToHost,
                        s->stream),
         "D2H");
  sync(s);
  S2T_API_END
}

int s2t_profile(s2t_solver* h, int enable) {
  return s2t_profile_kinds(h, enable ? (1u << S2T_PROF_KINDS) - 1 : 0u);
}

int s2t_profile_kinds(s2t_solver* h, unsigned mask) {
  return s2t_profile_sampled(h, mask, 1);
}

int s2t_profile_sampled(s2t_solver* h, unsigned mask, int every) {
  SOLVER(h);
  if (every < 1) return api_bad("profile sampling period must be >= 1");
  S2T_API_BEGIN
  mask &= (1u << S2T_PROF_KINDS) - 1;
  s->prof_every = every;
  s->prof_step = 0;
  if (mask) {
    sync(s);
    s->nev = 0;
    prof_reserve(s, 256);
    for (int k = 0; k < S2T_PROF_KINDS; ++k) {
      s->prof_launches[k] = 0;
      s->prof_ms[k] = 0;
    }
  } else {
    prof_collect(s);
  }
  s->prof_mask = mask;
  S2T_API_END
}

int s2t_profile_read(s2t_solver* h, int kind, int* launches, double* total_ms) {
  SOLVER(h);
  if (kind < 0 || kind >= S2T_PROF_KINDS) return api_bad("profile kind out of range");
  S2T_API_BEGIN
  prof_collect(s);
  if (launches) *launches = s->prof_launches[kind];
  if (total_ms) *total_ms = s->prof_ms[kind];
  S2T_API_END
}

int s2t_column_widths(int n, int* widths, int max, int* count) {
  const Ops* ops = ops_for(n);
  if (!ops) return api_bad(size_error(n).empty() ? "grid size " + std::to_string(n) +
                                                       " is not supported"
                                                 : size_error(n));
  int w[MAX_COL_WIDTHS];
  const int nw = ops->col_widths(w);
  if (count) *count = nw;
  for (int i = 0; i < nw && i < max && widths; ++i) widths[i] = w[i];
  return S2T_OK;
}

int s2t_set_column_width(s2t_solver* h, int width) {
  SOLVER(h);
  S2T_API_BEGIN
  set_col_width(s, width);
  S2T_API_END
}

int s2t_set_row_kernel(s2t_solver* h, int field_parallel) {
  SOLVER(h);
  if (field_parallel != 0 && field_parallel != 1) return api_bad("s2t_set_row_kernel: 0 or 1");
  S2T_API_BEGIN
  s->rows_fp = field_parallel != 0;
  S2T_API_END
}

int s2t_autotune(s2t_solver* h, int steps, int* width, double* ms_per_step, int max) {
  SOLVER(h);
  if (steps < 1) return api_bad("s2t_autotune: steps must be >= 1");
  S2T_API_BEGIN
  int w[MAX_COL_WIDTHS], nw = 0;
  double ms[MAX_COL_WIDTHS];
  const int best = autotune_cols(s, steps, w, ms, &nw);
  if (width) *width = best;
  for (int i = 0; i < nw && i < max && ms_per_step; ++i) ms_per_step[i] = ms[i];
  S2T_API_END
}

int s2t_copy_probe(double bytes, int reps, double* ms_per_copy) {
  if (!(bytes >= 16) || reps <= 0) return api_bad("s2t_copy_probe: bytes >= 16 and reps > 0");
  void *a = nullptr, *b = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = S2T_OK;
  try {
    const size_t n = (size_t)bytes / 16 * 16;
    hipchk(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
    hipchk(hipMalloc(&a, n), "hipMalloc");
    hipchk(hipMalloc(&b, n), "hipMalloc");
    fill_u32(st, a, n / 4, 0x3f800000u);
    hipchk(hipEventCreate(&e0), "hipEventCreate");
    hipchk(hipEventCreate(&e1), "hipEventCreate");
    copy_probe(st, a, b, n);  // warm-up
    hipchk(hipEventRecord(e0, st), "hipEventRecord");
    for (int i = 0; i < reps; ++i) copy_probe(st, a, b, n);
    hipchk(hipEventRecord(e1, st), "hipEventRecord");
    hipchk(hipStreamSynchronize(st), "hipStreamSynchronize");
    float ms = 0.f;
    hipchk(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    if (ms_per_copy) *ms_per_copy = ms / reps;
  } catch (const Fail& f) {
    g_err = f.what();
    rc = f.code;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (st) (void)hipStreamDestroy(st);
  return rc;
}

int s2t_kernel_bytes(s2t_solver* h, int kind, int stage, double* bytes) {
  SOLVER(h);
  if (!bytes) return api_bad("bytes is NULL");
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
  } else if (kind == S2T_PROF_DYE) {
    // one BFECC dye step (the fused kernel): phi0 read and the result written once (rgba16f,
    // 8 B per texel each), the N x N float2 velocity read once; phi1 and phi_hat stay on chip
    const double m = s->m;
    b = 2.0 * m * m * 8.0 + n * n * 8.0;
  } else {
    return api_bad("no byte model for this kernel kind");
  }
  *bytes = b;
  return S2T_OK;
}

}  // extern "C"
