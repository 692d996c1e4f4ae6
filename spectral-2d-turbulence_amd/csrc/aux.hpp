// Host launchers of the kernels in aux_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "physics.hpp"
#include "solver.hpp"

namespace s2t {

void fill_u32(hipStream_t st, void* p, size_t words, uint32_t v);
void reset_inst(hipStream_t st, Inst* in, int i0, int ni);
// parameters, clock reset and step budget of instance i in one launch (no host copy or wait)
void admit_inst(hipStream_t st, Inst* in, int i, const Phys& ph, int steps);
// physical vorticity of a serving scenario (S2T_SCENARIO_*) as a full complex field
void scenario_field(hipStream_t st, float2* out, int n, int scenario, double param);
void vpack_full(hipStream_t st, const float2* w, float2* pa, float2* pb, int n);
void assemble_full(hipStream_t st, const float2* z, const float2* fhat, float2* rhs,
                   const Phys& ph, uint32_t step, double dt, int n);
void real_to_complex(hipStream_t st, const float* r, float2* c, size_t n);
void seed_spectrum(hipStream_t st, float2* state, double* sums, int n, double k0,
                   double amplitude, uint32_t seed);
// fillSpectrum into a full n x n device array with no allocation (its scratch is a fixed device
// array of the library); n up to noalloc_max_n()
void fill_spectrum_noalloc(hipStream_t st, float2* full, int n, double k0, double amplitude,
                           uint32_t seed);
int noalloc_max_n();
void diagnostics(hipStream_t st, const float2* state, const double2* wtab, double* bins,
                 double* zonal, int n);
void dye_step(hipStream_t st, ushort4* const dye[3], int parity, const float2* vel,
              const DevState* dev, double dt, float fade, const Splats& sp, int n, int m);
// the same update as three launches (the unfused passes; test hook)
void dye_step_passes(hipStream_t st, ushort4* const dye[3], int parity, const float2* vel,
                     const DevState* dev, double dt, float fade, const Splats& sp, int n, int m);
void dye_paint(hipStream_t st, ushort4* dst, int m);
void copy_probe(hipStream_t st, const void* src, void* dst, size_t bytes);

}  // namespace s2t
