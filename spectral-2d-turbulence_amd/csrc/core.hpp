// Internal solver operations shared by the plain C ABI (api.hip) and the batch / serving C ABI
// (batch.hip). A Solver holds `batch` instances ("slots") of one grid size; every operation on
// one instance names its slot, and stepping advances all active slots with one launch per kernel.
#pragma once
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "../../include/s2t.h"
#include "solver.hpp"

namespace s2t {

struct Fail : std::runtime_error {
  int code;
  Fail(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void hipchk(hipError_t e, const char* what);
// 0 when n is 4^a or 2*4^a (src/gpu/fft.js:20-31)
int radix(int n, int* stages, int* has_radix2);
std::string size_error(int n);  // "" when n is a supported grid size
s2t_params default_params();
std::string check_params(const s2t_params& p);  // "" when valid

// stream: the caller's stream to run on (kept by the caller), or null for a stream of its own
Solver* solver_new(int n, int dye_scale, int batch, bool with_dye, hipStream_t stream = nullptr);
void solver_free(Solver* s);

void configure_slot(Solver* s, int slot, const s2t_params& p);
void reset_slot(Solver* s, int slot);
// band-limited random spectrum (s2t_bandlimited_spectrum), n*n interleaved floats, host only
void bandlimited_spectrum(int n, double kmin, double kmax, double amplitude, uint32_t seed,
                          float* out);
void seed_spectrum_slot(Solver* s, int slot, double k0, double amplitude, uint32_t seed);
void seed_physical_slot(Solver* s, int slot, const float* field);
void write_state_slot(Solver* s, int slot, const float* spectrum);
// phase >= 0 reads the slot from state[phase] instead of the solver's current phase (a slot that
// retired early keeps its final state in the buffer of the step it retired on)
void read_state_slot(Solver* s, int slot, float* spectrum, int phase = -1);
void diag_slot(Solver* s, int slot, s2t_diag* out, double* spectrum, double* zonal,
               int phase = -1);
// limit the slot to `steps` more steps, after which it retires itself (steps < 0: no limit)
void set_budget_slot(Solver* s, int slot, int steps);
// the serving loop's admission, enqueued with no host wait: seed the slot (a scenario built on the
// device, or fillSpectrum), then set its parameters, clock and step budget and make it active
void seed_scenario_slot_async(Solver* s, int slot, int scenario, double param);
void seed_spectrum_slot_async(Solver* s, int slot, double k0, double amplitude, uint32_t seed);
void admit_slot_async(Solver* s, int slot, const s2t_params& p, int steps);
void set_clock_slot(Solver* s, int slot, double t, unsigned step);
DevState read_clock_slot(Solver* s, int slot);

// encodeDiagnostics / copyDiagnostics / readDiagnostics of slot 0 (enqueue, enqueue, wait)
void encode_diag(Solver* s);
void copy_diag(Solver* s);
void read_copied_diag(Solver* s, s2t_diag* out, double* spectrum, double* zonal);

void ensure_primed(Solver* s);
void step_once(Solver* s, bool dye);
// column-group width tuning (s2t_autotune): returns the chosen width, also kept in s->cols_c
int autotune_cols(Solver* s, int steps, int* widths, double* ms, int* count);
void set_col_width(Solver* s, int width);  // 0: the size's default; else a compiled width
// width for `instances` instances of one size in per-class launches (0: the size's default)
int class_width(const Solver* s, int instances);
void default_class_width(Solver* s, int members);  // class_width for classes of two or more
void sync(Solver* s);

}  // namespace s2t
