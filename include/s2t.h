/*
 * s2t.h -- C ABI of the MI355X spectral 2-D turbulence step (libs2t.so).
 *
 * This is the boundary that replaces the reference's WebGPU `Solver` / `Fft2D` objects
 * (src/gpu/solver.js, src/gpu/fft.js). Every entry point is plain C: opaque handles, plain
 * pointers and sizes, int status (0 = ok, < 0 = error) and a thread-local message from
 * s2t_last_error(). Host arrays are copied at the call, like the reference's
 * queue.writeBuffer (src/gpu/solver.js:344, 355). Work is enqueued on the solver's own HIP stream;
 * only the read and diagnostics calls wait for it.
 *
 * Array layout used at this boundary (the fixture layout, reference/make_fixtures.py:11-16 and
 * src/shaders/solver.wgsl:27-28): a complex N x N field is row-major, index = row*N + col, with
 * kx = wn(col), ky = wn(row) (fftfreq order, Nyquist at -N/2), stored as interleaved float
 * (re, im) pairs: 2*N*N floats.
 */
#ifndef S2T_H
#define S2T_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct s2t_solver s2t_solver;

/* status codes */
#define S2T_OK 0
#define S2T_EINVAL -1   /* bad argument (size, pointer, count) */
#define S2T_EHIP -2     /* HIP runtime error */
#define S2T_ENOMEM -3   /* device allocation failed */
#define S2T_ESTATE -4   /* call not valid in the solver's current state */

/*
 * Physical and integration parameters. Mirrors the reference config keys (src/main.js:9-36) as
 * mapped by Solver.configure (src/gpu/solver.js:306-324) into Params (src/shaders/common.wgsl:8-21).
 * `viscosity` is the damping rate at the cut-off wavenumber kcut = N/3; the hyperviscosity
 * coefficient is viscosity / kcut^(2 hyper_p) (solver.js:311). fixed_dt > 0 bypasses the CFL
 * reduction and uses that step (the reference harness writes dt directly, tests/index.html:188).
 */
typedef struct s2t_params {
  double beta;        /* beta-plane parameter */
  double drag;        /* linear drag mu */
  double viscosity;   /* damping rate at kcut (0 = inviscid) */
  double hyper_p;     /* dissipation order p (1, 2, 4) */
  double force_amp;   /* stochastic forcing amplitude (0 = off) */
  double force_k;     /* forcing annulus centre */
  double force_width; /* forcing annulus width */
  double cfl;         /* CFL number */
  double dt_min;      /* lower clamp of the CFL step */
  double dt_max;      /* upper clamp of the CFL step */
  double dye_fade;    /* dye persistence per step */
  double fixed_dt;    /* > 0: use this dt instead of the CFL step */
} s2t_params;

/* Pointer splat: the record of src/ui/pointer.js:58-67 / Splat in common.wgsl:23-32. */
typedef struct s2t_splat {
  float x, y;        /* position in unit coordinates [0, 1) */
  float dx, dy;      /* pointer velocity (unit coordinates per frame) */
  float color[3];    /* dye colour */
  float radius;      /* Gaussian radius (unit coordinates) */
  float amp;         /* body-force amplitude */
  float dye_amp;     /* dye injection amplitude */
} s2t_splat;

#define S2T_MAX_SPLATS 8 /* src/gpu/solver.js:9 */

/* Diagnostics record: the fields of readDiagnostics (solver.js:591-602), in FP64. */
typedef struct s2t_diag {
  double energy;     /* 1/2 sum |w|^2/k^2 / N^4 */
  double enstrophy;  /* 1/2 sum |w|^2 / N^4 */
  double dt;         /* step used by the last step */
  double max_speed;  /* max |u|+|v| of the last step's initial field */
  double time;       /* simulated time */
  uint32_t step_index; /* number of steps taken since the last reset */
  uint32_t nonfinite;  /* nonzero once a non-finite velocity was seen */
} s2t_diag;

/* step flags */
#define S2T_STEP_DYE 1u /* advect the dye after each step (Solver.encodeDye, solver.js:523-536) */

/* ------------------------------------------------------------------------------------------ */
/* lifecycle                                                                                  */
/* ------------------------------------------------------------------------------------------ */

/* Solver.create(device, {n, dyeScale}) (solver.js:71-140). n must be 4^a or 2*4^a with
 * 16 <= n <= 4096 (fft.js:20-31 gives the error text); dye_scale in {1, 2} (src/main.js:199). */
int s2t_create(int n, int dye_scale, s2t_solver** out);
/* Solver.destroy() (solver.js:614-629). Waits for queued work. NULL is accepted. */
int s2t_destroy(s2t_solver* h);
/* Thread-local text of the last error on this thread ("" when none). */
const char* s2t_last_error(void);
/* radixSchedule(n) (fft.js:20-31): number of radix-4 stages and whether a radix-2 stage
 * follows; returns S2T_EINVAL with the reference's message for other n. */
int s2t_radix_schedule(int n, int* radix4_stages, int* has_radix2);

/* ------------------------------------------------------------------------------------------ */
/* parameters and inputs                                                                      */
/* ------------------------------------------------------------------------------------------ */

/* Solver.configure(cfg) (solver.js:306-324). */
int s2t_configure(s2t_solver* h, const s2t_params* p);
/* Hyperviscosity coefficient in effect (Solver.paramsF[2], tests/index.html:136-140). */
int s2t_nu_h(s2t_solver* h, double* nu_h);
/* Solver.setSplats(list) (solver.js:326-345); count is clamped to S2T_MAX_SPLATS. */
int s2t_set_splats(s2t_solver* h, const s2t_splat* splats, int count);

/* ------------------------------------------------------------------------------------------ */
/* seeding (Solver.reset and its two routes, solver.js:347-444)                               */
/* ------------------------------------------------------------------------------------------ */

/* Solver.seed({k0, amplitude, seed}) via fillSpectrum (solver.js:347-359, 384-418); amplitude 0
 * gives a fluid at rest. Resets step index, time and dt like Solver.reset (solver.js:420-444). */
int s2t_seed_spectrum(s2t_solver* h, double k0, double amplitude, uint32_t seed);
/* fillSpectrum(data, k0, amplitude, seed) into caller memory with no allocation. No counterpart
 * in the reference (its fillSpectrum, src/gpu/solver.js:384-418, fills a host array); for callers
 * that cannot allocate device memory: code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc. Writes the spectrum of s2t_seed_spectrum, bit for bit, as 2*n*n floats
 * in the fixture layout (both halves, zero out of band) to spectrum_dev, device memory of the
 * current device; enqueued on `stream` (a hipStream_t; NULL = the null stream), no wait. Calls no
 * hipMalloc, hipFree or allocating library routine: the per-column sums live in a fixed device
 * array of the library (one per GPU, about 11 KB, held until the process ends), so n is at most
 * 4096. Calls that overlap on one GPU share that array and race: order them (one stream, or
 * events). */
int s2t_fill_spectrum_noalloc(int n, double k0, double amplitude, uint32_t seed,
                              float* spectrum_dev, void* stream);
/* Solver.seedPhysical(values) (solver.js:361-382): n*n physical vorticity -> forward FFT ->
 * 2/3 mask, k = 0 zeroed (src/shaders/seed.wgsl:4-17). Also resets like Solver.reset. */
int s2t_seed_physical(s2t_solver* h, const float* field);
/* Band-limited random vorticity (no counterpart in the reference; the initial condition of the
 * 1024^2 decaying-turbulence case): every in-band mode (2/3 rule, k != 0) with
 * kmin <= |k| <= kmax gets a standard-normal amplitude and a uniform phase, drawn per canonical
 * mode (ky > 0, or ky = 0 and kx > 0) from hash32 streams of the seed (src/gpu/solver.js:43-47);
 * the mirror mode is the conjugate, and the field is scaled to rms vorticity `amplitude`.
 * s2t_bandlimited_spectrum fills n*n interleaved floats on the host (no GPU needed);
 * s2t_seed_bandlimited stores the same spectrum as the state and resets like s2t_seed_spectrum. */
int s2t_bandlimited_spectrum(int n, double kmin, double kmax, double amplitude, uint32_t seed,
                             float* spectrum);
int s2t_seed_bandlimited(s2t_solver* h, double kmin, double kmax, double amplitude, uint32_t seed);
/* Write the vorticity spectrum (2*n*n floats, fixture layout). The field must be the spectrum of
 * a real field; it is stored masked to the 2/3 band with k = 0 zeroed. Step index, time and dt
 * are left as they are (the reference harness writes state[0] directly, tests/index.html:187). */
int s2t_write_state(s2t_solver* h, const float* spectrum);
/* Read the vorticity spectrum (2*n*n floats, fixture layout). Waits for queued work. */
int s2t_read_state(s2t_solver* h, float* spectrum);
/* Zero the step index and time and set dt = dt_max (part of Solver.reset). */
int s2t_reset_clock(s2t_solver* h);
/* Set the simulated time and the step index (the stochastic forcing phases follow the step
 * index). No counterpart in the reference, whose stepIndex is a host field (solver.js:520): with
 * s2t_read_state / s2t_write_state and s2t_diagnostics' time and step_index it makes a
 * checkpoint that resumes a forced run exactly. */
int s2t_set_clock(s2t_solver* h, double time, uint32_t step_index);

/* ------------------------------------------------------------------------------------------ */
/* stepping                                                                                   */
/* ------------------------------------------------------------------------------------------ */

/* nsteps x Solver.encodeStep (solver.js:488-521) with the CFL dt of each step taken from its
 * initial field (reference/spectral2d.py:66-71, 77-78), followed by encodeDye when
 * (flags & S2T_STEP_DYE). The velocity prime of encodePrime (solver.js:475-486) is implicit.
 * Enqueues only; does not wait. */
int s2t_step(s2t_solver* h, int nsteps, uint32_t flags);
/* Wait for all queued work on the solver's stream. */
int s2t_sync(s2t_solver* h);
/* Solver.encodePrime (solver.js:475-486): the velocity planes of the current state (every step
 * primes a replaced state by itself; this does it now). Enqueues only. */
int s2t_prime(s2t_solver* h);
/* Solver.primed: 1 once the current state's velocity planes exist (after a prime or a step),
 * 0 after seeding or writing the state. */
int s2t_primed(s2t_solver* h, int* primed);

/* ------------------------------------------------------------------------------------------ */
/* diagnostics and dye                                                                        */
/* ------------------------------------------------------------------------------------------ */

/* readDiagnostics (solver.js:547-608) computed in FP64 with a fixed reduction order:
 * spectrum[n/2+1] = E(k) in bins round(|k|) (reference/spectral2d.py:98-104), zonal[n] =
 * zonal-mean u per row (spectral2d.py:106-108). Either array may be NULL. Waits. */
int s2t_diagnostics(s2t_solver* h, s2t_diag* out, double* spectrum, double* zonal);
/* Solver.encodeDiagnostics (solver.js:547-558): the FP64 diagnostics of the current state and a
 * snapshot of its clock into the solver's device record. Enqueues only. */
int s2t_encode_diagnostics(s2t_solver* h);
/* Solver.copyDiagnostics (solver.js:560-568): the last encoded record to pinned host memory.
 * Enqueues only; S2T_ESTATE when nothing was encoded. */
int s2t_copy_diagnostics(s2t_solver* h);
/* readDiagnostics (solver.js:570-608) of the copied record: waits for that copy only (not for
 * work queued after it). Fills the fields of s2t_diagnostics; S2T_ESTATE when nothing was copied
 * since the last read. */
int s2t_read_copied_diagnostics(s2t_solver* h, s2t_diag* out, double* spectrum, double* zonal);
/* Solver.paintDye(pattern) (solver.js:446-473): 0 = 'none', 1 = 'bands'. */
int s2t_paint_dye(s2t_solver* h, int pattern);
/* Dye texture side M = n * dye_scale. */
int s2t_dye_size(s2t_solver* h, int* m);
/* Read the current dye field (Solver.currentDye) as M*M*4 binary16 values (rgba). Waits. */
int s2t_read_dye(s2t_solver* h, uint16_t* rgba);
/* Write the current dye field from M*M*4 binary16 values. */
int s2t_write_dye(s2t_solver* h, const uint16_t* rgba);
/* Write the physical velocity (n*n*2 floats, (u, v) per cell) that the dye advection reads;
 * s2t_step with S2T_STEP_DYE writes it from each step's initial field. */
int s2t_write_velocity(s2t_solver* h, const float* uv);
/* One dye update (advect_fwd, advect_back, advect_corr; src/shaders/dye.wgsl:58-113) with the
 * current velocity, splats and fade: Solver.encodeDye (solver.js:523-536) when dt <= 0, which
 * takes the step of the last solver step from the device (dtbuf, dye.wgsl:63); a given dt > 0 is
 * used as is. Enqueues only. */
int s2t_dye_step(s2t_solver* h, double dt);
/* The same dye update as three launches, one per pass of dye.wgsl:58-113 (advect_fwd, advect_back,
 * advect_corr), with the intermediate fields stored as textures: the fused update of
 * s2t_dye_step, bit for bit (test hook). Enqueues only. */
int s2t_dye_step_passes(s2t_solver* h, double dt);
/* Blank all three dye textures and reset their parity (part of Solver.reset, solver.js:420-436,
 * before paintDye). Waits. */
int s2t_reset_dye(s2t_solver* h);

/* ------------------------------------------------------------------------------------------ */
/* test hooks: the product kernels on host data (the reference harness reaches the same stages  */
/* through Solver internals, tests/index.html:145-201)                                         */
/* ------------------------------------------------------------------------------------------ */

/* Fft2D.forward / inverse on a full complex n x n field (fft.js:81-100): row pass then column
 * pass, forward unscaled, inverse scaled by 1/n per pass. in and out may alias. */
int s2t_fft2(s2t_solver* h, const float* in, float* out, int inverse);
/* Fft2D.run (fft.js:81-92) on a device buffer: the n x n complex field at data_dev (2*n*n floats,
 * fixture layout, device memory) transformed in place on the solver's stream, row pass then
 * column pass, forward unscaled, inverse scaled by 1/n per pass. Enqueues only. */
int s2t_fft2_device(s2t_solver* h, float* data_dev, int inverse);
/* Device buffers for the Fft2D.register route (GPUDevice.createBuffer, queue.writeBuffer,
 * copyBufferToBuffer and a mapped read, fft.js:77-79): zero-filled allocation, free, and
 * host-blocking copies. They are not ordered against a solver's stream: call s2t_sync on a
 * solver whose queued work touches the buffer first. */
int s2t_buffer_create(size_t bytes, void** dev);
int s2t_buffer_destroy(void* dev);
int s2t_buffer_write(void* dev, size_t offset, const void* src, size_t bytes);
int s2t_buffer_read(const void* dev, size_t offset, void* dst, size_t bytes);
int s2t_buffer_copy(void* dst, size_t dst_offset, const void* src, size_t src_offset,
                    size_t bytes);
/* velocity_pack (solver.wgsl:21-40) on a full spectrum: packA = w + i u_hat, packB = v_hat. */
int s2t_velocity_pack(s2t_solver* h, const float* w, float* pack_a, float* pack_b);
/* assemble (solver.wgsl:56-95) on a full product spectrum z and pointer-forcing spectrum fhat
 * (NULL = zero), with the stochastic forcing of the configured amplitude at step index
 * `step_index` and step `dt`: rhs = (fhat - N_hat) * mask + forcing. */
int s2t_assemble(s2t_solver* h, const float* z, const float* fhat, uint32_t step_index, double dt,
                 float* rhs);
/* Right-hand side -N_hat(w) (+ forcing) of the current state through the fused product path
 * (row kernel + column kernel), full fixture layout. Uses step index `step_index` and `dt` for
 * the stochastic forcing. */
int s2t_rhs(s2t_solver* h, uint32_t step_index, double dt, float* rhs);

/* ------------------------------------------------------------------------------------------ */
/* batched ensembles: many independent instances of mixed grid sizes stepped together           */
/* (no counterpart in the reference, which runs one solver)                                     */
/* ------------------------------------------------------------------------------------------ */

typedef struct s2t_batch s2t_batch;

/* Members 0..count-1 with grid sizes sizes[i]. Members of one size share buffers and one launch
 * per fused kernel; each member has its own parameters, seed, clock and forcing phases, and its
 * results are bit-identical to the same member run as a plain solver. */
int s2t_batch_create(const int* sizes, int count, s2t_batch** out);
int s2t_batch_destroy(s2t_batch* b);
int s2t_batch_size(s2t_batch* b, int member, int* n);
int s2t_batch_configure(s2t_batch* b, int member, const s2t_params* p);
int s2t_batch_seed_spectrum(s2t_batch* b, int member, double k0, double amplitude, uint32_t seed);
int s2t_batch_seed_physical(s2t_batch* b, int member, const float* field);
int s2t_batch_write_state(s2t_batch* b, int member, const float* spectrum);
int s2t_batch_read_state(s2t_batch* b, int member, float* spectrum);
int s2t_batch_diagnostics(s2t_batch* b, int member, s2t_diag* out, double* spectrum, double* zonal);
/* Capture each RK3 step of all members as a hipGraph (default on) or launch eagerly. */
int s2t_batch_use_graphs(s2t_batch* b, int enable);
/* Grouped launches (no counterpart in the reference, whose instances are single solvers): the
 * classes of sizes 256, 512 and 1024 are stepped by six launches per step that each cover every
 * instance of those classes, instead of six launches per class on concurrent graph branches;
 * other sizes keep their own launches. Results are bit-identical either way. Default on; the
 * column-group widths of s2t_batch_autotune / s2t_batch_column_width apply to the classes on
 * their own launches. */
int s2t_batch_use_grouped(s2t_batch* b, int enable);
/* nsteps steps of every member; enqueues only. */
int s2t_batch_step(s2t_batch* b, int nsteps);
int s2t_batch_sync(s2t_batch* b);
/* s2t_autotune for every size class of the batch, each with its own member count (replaces the
 * captured steps). */
int s2t_batch_autotune(s2t_batch* b, int steps);
/* Column-group width in effect for a member's size class: the tuned width, else for a class of
 * two or more members the width measured fastest for that many (4 once the class fills the chip:
 * 2048 from four members, 1024 from eight, 512 from sixteen, 128 and 256 from two; 2 for two or
 * three of 2048), else the size's default (s2t_column_widths, first). */
int s2t_batch_column_width(s2t_batch* b, int member, int* width);

/* ------------------------------------------------------------------------------------------ */
/* serving: continuous batching of simulation requests (no counterpart in the reference)        */
/* ------------------------------------------------------------------------------------------ */

typedef struct s2t_server s2t_server;

/* scenarios (src/scenarios.js:57-165, Solver.seed of solver.js:347-359) */
#define S2T_SCENARIO_TURBULENCE 0  /* fillSpectrum(k0, amplitude, seed) */
#define S2T_SCENARIO_TAYLOR_GREEN 1
#define S2T_SCENARIO_DIPOLE 2
#define S2T_SCENARIO_MERGER 3      /* scenario_param = core separation */

/* request status */
#define S2T_REQ_QUEUED 0
#define S2T_REQ_RUNNING 1
#define S2T_REQ_DONE 2
#define S2T_REQ_REJECTED 3
#define S2T_REQ_CANCELLED 4
#define S2T_REQ_FAILED 5 /* the server's worker stopped on an error; it accepts nothing more */

typedef struct s2t_request {
  int n;                 /* grid size (4^a or 2*4^a, 16..4096) */
  int steps;             /* RK3 steps to run */
  int scenario;          /* S2T_SCENARIO_* */
  int keep_state;        /* nonzero: keep the final spectrum for s2t_server_read_state */
  double scenario_param; /* merger separation */
  double k0, amplitude;  /* turbulence seeding */
  uint32_t seed;
  s2t_params params;
} s2t_request;

typedef struct s2t_server_stats {
  uint64_t submitted, completed, rejected, cancelled;
  double p50_latency_s, p99_latency_s, mean_latency_s;
  double mpts_steps_per_s; /* completed work / time from first submit to last completion */
  double budget_bytes, used_bytes; /* the budget (0 = none) and the device bytes held now */
  double peak_bytes;               /* the most held at once */
} s2t_server_stats;

/* A server with a device-memory budget (budget_bytes; 0 = no limit). The instances of one grid size
 * live in chunks of slots allocated as requests need them; sizes[i] (any supported size) gets
 * chunks of slots[i] instances and one chunk at once (nclasses may be 0; other sizes get chunks
 * of about 2^22 points, at most 32 instances). A request takes a free slot, or a new chunk if the
 * budget has room (idle chunks are freed to make room); otherwise it waits, and later requests
 * that fit are admitted past it. A worker thread admits and steps. */
int s2t_server_create(const int* sizes, const int* slots, int nclasses, double budget_bytes,
                      s2t_server** out);
/* Cancels the requests still queued or running (their callbacks run), stops the worker and
 * frees everything. */
int s2t_server_destroy(s2t_server* s);
/* Thread-safe. An unsupported size, or one of which not even a single instance fits in the budget,
 * is rejected (status S2T_REQ_REJECTED). */
int s2t_server_submit(s2t_server* s, const s2t_request* r, uint64_t* id);
/* Status, FP64 diagnostics of the final state and submit-to-completion latency (-1 until done). */
int s2t_server_poll(s2t_server* s, uint64_t id, int* status, s2t_diag* diag, double* latency_s);
/* Final spectrum (2*n*n floats, fixture layout) of a finished request submitted with keep_state. */
int s2t_server_read_state(s2t_server* s, uint64_t id, float* spectrum);
/* Thread-safe. A queued request is dropped at once; a running one is stopped by the worker
 * before its next steps (its slot is freed, no result is kept). Status S2T_REQ_CANCELLED.
 * S2T_ESTATE if the request has already finished (done, rejected or cancelled). */
int s2t_server_cancel(s2t_server* s, uint64_t id);
/* Completion callback, called once per request when it reaches a final status (done, rejected,
 * cancelled, failed): on the worker thread, or on the submitting / cancelling thread when the
 * status is final at that call. If the worker stops on an error, every request still queued or
 * running fails (S2T_REQ_FAILED) and is called back; s2t_server_poll then reports the error. It runs without the server's lock held, so it may call s2t_server_poll and
 * s2t_server_read_state; it must not call s2t_server_destroy. NULL removes it. */
typedef void (*s2t_server_callback)(uint64_t id, int status, void* user);
int s2t_server_set_callback(s2t_server* s, s2t_server_callback cb, void* user);
/* Failure injection, with no counterpart in the reference, for callers that test their handling
 * of a failed server: the worker fails (as on a device error) at the admission after
 * `after_admissions` more admissions counted from the server's start; < 0 turns it off. */
int s2t_server_inject_failure(s2t_server* s, int after_admissions);
/* Wait until every submitted request is done, rejected or cancelled (timeout_s <= 0: no limit). */
int s2t_server_drain(s2t_server* s, double timeout_s);
int s2t_server_get_stats(s2t_server* s, s2t_server_stats* out);

/* ------------------------------------------------------------------------------------------ */
/* profiling: per-kernel HIP-event timing on the solver's stream                                */
/* ------------------------------------------------------------------------------------------ */

/* ------------------------------------------------------------------------------------------ */
/* Tuning, with no counterpart in the reference (its schedules are fixed per grid size). The step's
 * column kernel is compiled for a few column-group widths (columns per workgroup) per grid size;
 * every width gives bit-identical results, and the best one depends on the size and on how many
 * instances share the launches. */
/* Compiled widths for grid size n, the default first (host only). */
int s2t_column_widths(int n, int* widths, int max, int* count);
/* Use a compiled width for this solver's steps (0: the size's default). */
int s2t_set_column_width(s2t_solver* h, int width);
/* Row kernel of this solver's steps: 1 (the default) the field-parallel row kernel where it is
 * compiled (512 and 1024: three thread groups per row pair), 0 one thread group per row pair.
 * Both give bit-identical results; other sizes have one row kernel and ignore the choice. */
int s2t_set_row_kernel(s2t_solver* h, int field_parallel);
/* Times `steps` steps per compiled width on a scratch solver of the same size and parameters (the
 * solver's own state is not touched) and keeps the fastest: *width, and the milliseconds per step
 * of each width in s2t_column_widths order in ms_per_step[0 .. max-1]. */
int s2t_autotune(s2t_solver* h, int steps, int* width, double* ms_per_step, int max);

#define S2T_PROF_ROWS 0  /* fused row kernel */
#define S2T_PROF_COLS 1  /* fused column kernel */
#define S2T_PROF_DYE 2   /* dye kernels */
#define S2T_PROF_OTHER 3 /* prime and everything else */
#define S2T_PROF_KINDS 4

/* Enable (1) or disable (0) event timing around every kernel launched by s2t_step; enabling
 * clears the totals. */
int s2t_profile(s2t_solver* h, int enable);
/* As s2t_profile, timing only the kernel kinds in `mask` (bit k = kind k; 0 disables). Each timed
 * launch adds two event records to the stream (a few microseconds), so a throughput measurement
 * times only the kinds it reports. */
int s2t_profile_kinds(s2t_solver* h, unsigned mask);
/* As s2t_profile_kinds, timing the launches of one s2t_step step in `every` (the first, then
 * every `every`-th step after it), so that a throughput measurement carries fewer event records
 * while every kernel stage is still sampled. */
int s2t_profile_sampled(s2t_solver* h, unsigned mask, int every);
/* Launch count and summed device time (ms) of one kernel kind since profiling was enabled.
 * Waits for queued work. */
int s2t_profile_read(s2t_solver* h, int kind, int* launches, double* total_ms);
/* Algorithmic HBM bytes of one launch of a kernel kind for this solver's size and the current
 * options (stage 0 of a step when the kinds differ by stage); for S2T_PROF_DYE, of one dye step
 * (the fused update: 16 B per dye texel + 8 B per grid point). */
int s2t_kernel_bytes(s2t_solver* h, int kind, int stage, double* bytes);
/* Measurement helper with no counterpart in the reference: `reps` launches of a plain 16-byte
 * streaming copy kernel (k_copy) of `bytes` bytes on a private stream; average ms per copy. Its
 * counter readings calibrate the HBM traffic counters of the step kernels. */
int s2t_copy_probe(double bytes, int reps, double* ms_per_copy);

#ifdef __cplusplus
}
#endif

#endif /* S2T_H */
