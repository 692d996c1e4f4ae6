// Host-side solver object behind the C ABI, and the table of size-specialised launchers.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/s2t.h"
#include "physics.hpp"

namespace s2t {

struct Solver;

// Launchers specialised on the grid size N (one table per supported N).
struct Ops {
  size_t state_elems;   // complex values per state field
  size_t plane_elems;   // complex values per mixed-space field (N * (K+1))
  int twiddle_elems;    // size of the line-FFT twiddle table
  void (*twiddles)(float2* host);  // fill the FP64-derived twiddle table
  // fused step kernels
  // rows returns true when the products were written as S', D' (sum / difference form), which
  // the following cols launch must be told
  bool (*rows)(Solver* s, int stage, bool speed, bool vel, bool force);
  void (*cols)(Solver* s, int stage, bool force, bool sd);
  void (*prime)(Solver* s, int i0, int ni);
  void (*rhs)(Solver* s, bool force, bool sd, float2* rhs_out, uint32_t step, double dt);
  // full-field helpers
  void (*fft2)(Solver* s, float2* data, int inverse);
  void (*pack)(Solver* s, const float2* full, float2* state);
  void (*unpack)(Solver* s, const float2* state, float2* full);
  // column-group widths compiled for the step's column kernel (the default first); returns the
  // count, at most MAX_COL_WIDTHS
  int (*col_widths)(int* out);
};
constexpr int MAX_COL_WIDTHS = 6;

const Ops* ops_for(int n);

// Grouped launches (step_kernels.hip): one step of up to four size classes of sizes 256, 512 or
// 1024 (no pointer forcing, no dye, one buffer phase, primed) in six launches, each covering every
// instance of every class.
bool group_size(int n);
void group_step(Solver* const* cls, int count, hipStream_t stream);

enum ColsMode { COLS_STEP = 0, COLS_PRIME = 1, COLS_RHS = 2 };

struct Solver {
  int n = 0, K = 0, dye_scale = 1, m = 0;
  const Ops* ops = nullptr;
  hipStream_t stream = nullptr;
  bool own_stream = true;  // false: the stream is the caller's (a serving lane), not destroyed here

  // device buffers
  float2* state[3] = {nullptr, nullptr, nullptr};  // compact half spectra, rotating
  float2* mixed_u = nullptr;   // omega', u', v' tiles (F = 3)
  float2* mixed_z = nullptr;   // P', Q' tiles (F = 2)
  float2* mixed_f = nullptr;   // pointer-force F' tiles (F = 1), lazily
  float2* force_k = nullptr;   // pointer-force spectrum, compact, lazily
  float2* vel = nullptr;       // physical (u, v), N*N, lazily (dye)
  float2* full = nullptr;      // full N*N complex scratch for seeding and hooks, lazily
  float2* full2 = nullptr;     // second full scratch (hooks), lazily
  float2* tw = nullptr;        // line-FFT twiddles
  double2* wtab = nullptr;     // exp(2 pi i m / N), FP64, for the zonal mean
  Inst* inst = nullptr;        // per-instance parameters and scalars (batch entries)
  int batch = 1;               // instances sharing this solver's buffers
  int cols_c = 0;              // column-kernel width in a step (0: the default; < 0: field-parallel)
  bool rows_fp = true;         // field-parallel row kernel where it is compiled (512, 1024)
  double* diag_bins = nullptr; // [2][NBINS] energy and enstrophy per bin, then 2 totals
  double* diag_zonal = nullptr;
  double* seed_sum = nullptr;  // per-column partial sums for fillSpectrum normalisation
  ushort4* dye[3] = {nullptr, nullptr, nullptr};
  // encodeDiagnostics / copyDiagnostics: the record of the last encode on the device (bins,
  // zonal mean, clock), its pinned host copy and the event that marks the copy, lazily
  char* enc_diag = nullptr;
  char* enc_host = nullptr;
  hipEvent_t enc_event = nullptr;
  bool enc_copied = false;

  // host mirror of the configuration of each slot
  std::vector<Phys> phys_slot;
  std::vector<s2t_params> params_slot;
  std::vector<char> primed_slot;  // mixed_u of the slot holds the prime of its state[phase]
  Splats splats{};                // pointer splats (slot 0 of a plain solver)
  int phase = 0;                  // state[phase] holds the current spectra of all slots
  int dye_parity = 0;
  bool with_dye = false;
  // The stepped dye runs on a stream of its own, beside stages 1 and 2 of its step: it reads the
  // stage-0 velocity and the step's dt, which they do not write. ev_cols0 marks the stage-0 column
  // kernel on `stream`, ev_dye the dye step on dye_stream; `stream` waits for ev_dye at the end of
  // the step, so whatever is enqueued on it afterwards follows the dye step. Lazily.
  hipStream_t dye_stream = nullptr;
  hipEvent_t ev_cols0 = nullptr, ev_dye = nullptr;

  // profiling
  unsigned prof_mask = 0;  // kernel kinds timed with events (bit k = S2T_PROF kind k)
  bool prof_open = false;  // prof_begin recorded an event that prof_end closes
  int prof_every = 1;      // time the launches of one step in prof_every
  int prof_step = 0;       // steps since profiling was enabled
  bool prof_skip = false;  // the current step is not timed
  struct Ev { hipEvent_t a, b; int kind; };
  Ev* evs = nullptr;
  int nev = 0, cap_ev = 0;
  int prof_launches[S2T_PROF_KINDS] = {0, 0, 0, 0};
  double prof_ms[S2T_PROF_KINDS] = {0, 0, 0, 0};
};

// profiling brackets around a launch (on the solver's stream, or on `st`)
void prof_reserve(Solver* s, int count);
void prof_begin(Solver* s, int kind, hipStream_t st = nullptr);
void prof_end(Solver* s, hipStream_t st = nullptr);

// launch-error check used by the launchers; records the error text and throws
void check_launch(const char* what);

}  // namespace s2t
