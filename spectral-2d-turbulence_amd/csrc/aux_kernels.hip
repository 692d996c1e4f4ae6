// Kernels off the fused step path: seeding, FP64 diagnostics, BFECC dye and the full-field
// test hooks. Grid size is a runtime argument here (no line FFTs).
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>

#include "aux.hpp"
#include "core.hpp"
#include "physics.hpp"

namespace s2t {

// compact-state geometry at runtime
struct RGeo {
  int n, K, KC, NJ, pitch;
  __host__ __device__ int jrow(int i) const {
    return i <= K ? i : (i >= n - K ? i - (n - NJ) : -1);
  }
};
static RGeo rgeo(int n) {
  RGeo g;
  g.n = n;
  g.K = n / 3;
  g.KC = g.K + 1;
  g.NJ = 2 * g.K + 1;
  g.pitch = (g.NJ + 63) / 64 * 64;
  return g;
}

static unsigned blocks(size_t n, int b) { return (unsigned)((n + b - 1) / b); }

// ---------------------------------------------------------------------------------------------
// fill kernels (used instead of memset so a captured graph never holds a memset node)
// ---------------------------------------------------------------------------------------------
__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
void fill_u32(hipStream_t st, void* p, size_t words, uint32_t v) {
  if (words == 0) return;
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks(words, 256)), dim3(256), 0, st, (uint32_t*)p, words,
                     v);
  check_launch("k_fill_u32");
}

// zero the clock of instances i0 .. i0+ni-1 and make them active (part of Solver.reset,
// solver.js:420-444: stepIndex = 0, simTime = 0, dt = dtMax)
__global__ void k_reset_inst(Inst* in, int i0, int ni) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ni) return;
  DevState& d = in[i0 + i].st;
  d.dt = in[i0 + i].ph.dt_max;
  d.t = 0.0;
  d.speed = 0.f;
  for (int k = 0; k < DevState::SPEED_SLOTS; ++k) d.speed_bits[k] = 0u;
  d.step = 0u;
  d.nonfinite = 0u;
  d.active = 1;
  d.steps_left = -1;
  d.cols_done = 0u;
}
void reset_inst(hipStream_t st, Inst* in, int i0, int ni) {
  hipLaunchKernelGGL(k_reset_inst, dim3(blocks(ni, 64)), dim3(64), 0, st, in, i0, ni);
  check_launch("k_reset_inst");
}

// A slot's admission in one launch: its parameters, a clock reset as k_reset_inst's and its step
// budget, all from the kernel's arguments, so the host neither copies nor waits (the serving
// loop admits requests this way, ordered behind the steps already queued on the stream)
__global__ void k_admit_inst(Inst* in, int i, Phys ph, int steps) {
  if (threadIdx.x != 0) return;
  in[i].ph = ph;
  DevState& d = in[i].st;
  d.dt = ph.dt_max;
  d.t = 0.0;
  d.speed = 0.f;
  for (int k = 0; k < DevState::SPEED_SLOTS; ++k) d.speed_bits[k] = 0u;
  d.step = 0u;
  d.nonfinite = 0u;
  d.active = 1;
  d.steps_left = steps;
  d.cols_done = 0u;
}
void admit_inst(hipStream_t st, Inst* in, int i, const Phys& ph, int steps) {
  hipLaunchKernelGGL(k_admit_inst, dim3(1), dim3(64), 0, st, in, i, ph, steps);
  check_launch("k_admit_inst");
}

// Physical vorticity of a serving scenario on x_j = 2 pi j / N (src/scenarios.js:57-165), in FP64,
// as the real part of a full complex field: Taylor-Green 6 cos 2x cos 2y; the Gaussian dipole
// (cores of radius 0.3 and circulation +-12 at x = 0.6 pi, y = pi -+ 0.45); the two-vortex merger
// (amplitude 30, radius 0.22, separation `param`, 0.7 when not positive). Blobs are cut at
// r^2 < 25 a^2 with periodic separations (common.wgsl:46-48).
__device__ double pdelta(double a, double b) {
  double d = a - b;
  if (d > M_PI) d -= 2 * M_PI;
  else if (d < -M_PI) d += 2 * M_PI;
  return d;
}
__device__ double blob(double x, double y, double cx, double cy, double a, double amp) {
  const double dx = pdelta(x, cx), dy = pdelta(y, cy);
  const double r2 = dx * dx + dy * dy, a2 = a * a;
  return r2 < 25 * a2 ? amp * exp(-r2 / a2) : 0.0;
}
__global__ void k_scenario(float2* out, int n, int scenario, double param) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)n * n) return;
  const double h = 2 * M_PI / n;
  const double x = (double)(i % n) * h, y = (double)(i / n) * h;
  double v = 0.0;
  if (scenario == S2T_SCENARIO_TAYLOR_GREEN) {
    v = 6 * cos(2 * x) * cos(2 * y);
  } else if (scenario == S2T_SCENARIO_DIPOLE) {
    const double a = 0.3, amp = 12 / (M_PI * a * a), cx = 0.6 * M_PI;
    v = blob(x, y, cx, M_PI - 0.45, a, amp) + blob(x, y, cx, M_PI + 0.45, a, -amp);
  } else if (scenario == S2T_SCENARIO_MERGER) {
    const double d = param > 0 ? param : 0.7;
    v = blob(x, y, M_PI - d / 2, M_PI, 0.22, 30) + blob(x, y, M_PI + d / 2, M_PI, 0.22, 30);
  }
  out[i] = make_float2((float)v, 0.f);
}
void scenario_field(hipStream_t st, float2* out, int n, int scenario, double param) {
  hipLaunchKernelGGL(k_scenario, dim3(blocks((size_t)n * n, 256)), dim3(256), 0, st, out, n,
                     scenario, param);
  check_launch("k_scenario");
}

// plain streaming copy (16-byte vectors, grid-stride, four vectors in flight per thread): the
// reference kernel whose counter readings calibrate the HBM traffic counters of the step kernels,
// and a measure of the HBM bandwidth a plain copy reaches (about 5.5 TB/s on an MI355X at 1 GiB)
__global__ void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
  constexpr int U = 4;
  const size_t stride = (size_t)gridDim.x * blockDim.x * U;
  for (size_t base = (size_t)blockIdx.x * blockDim.x * U + threadIdx.x; base < n; base += stride) {
    float4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const size_t i = base + (size_t)u * blockDim.x;
      if (i < n) v[u] = src[i];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const size_t i = base + (size_t)u * blockDim.x;
      if (i < n) dst[i] = v[u];
    }
  }
}
void copy_probe(hipStream_t st, const void* src, void* dst, size_t bytes) {
  const size_t n = bytes / 16;
  hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, st, (const float4*)src, (float4*)dst, n);
  check_launch("k_copy");
}

// ---------------------------------------------------------------------------------------------
// test hooks on full N x N fields
// ---------------------------------------------------------------------------------------------
// velocity_pack (solver.wgsl:21-40)
__global__ void k_vpack_full(const float2* __restrict__ w, float2* __restrict__ pa,
                             float2* __restrict__ pb, int n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)n * n) return;
  const float kx = (float)wn((int)(i % n), n), ky = (float)wn((int)(i / n), n);
  const float k2 = kx * kx + ky * ky;
  const float inv = k2 > 0.f ? 1.f / k2 : 0.f;
  const float2 v = w[i];
  float2 uh, vh;
  vel_pack(v, kx, ky, inv, uh, vh);
  pa[i] = cadd(v, cmuli(uh));
  pb[i] = vh;
}
void vpack_full(hipStream_t st, const float2* w, float2* pa, float2* pb, int n) {
  hipLaunchKernelGGL(k_vpack_full, dim3(blocks((size_t)n * n, 256)), dim3(256), 0, st, w, pa, pb,
                     n);
  check_launch("k_vpack_full");
}

// assemble (solver.wgsl:56-95)
__global__ void k_assemble_full(const float2* __restrict__ z, const float2* __restrict__ fhat,
                                float2* __restrict__ rhs, Phys ph, uint32_t step, float dt,
                                int n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)n * n) return;
  const int col = (int)(i % n), row = (int)(i / n);
  const float kx = (float)wn(col, n), ky = (float)wn(row, n);
  const float k2 = kx * kx + ky * ky;
  const float2 zk = z[i];
  const float2 zm = z[(size_t)((n - row) % n) * n + (n - col) % n];
  float2 q1, q2;
  herm_unpack(zk, zm, q1, q2);
  const float2 nh = nhat(q1, q2, kx, ky);
  const float kcut = (float)n / 3.f;
  const bool band = fabsf(kx) < kcut && fabsf(ky) < kcut;
  float2 f = make_float2(0.f, 0.f);
  if (band) {
    const float2 fh = fhat ? fhat[i] : make_float2(0.f, 0.f);
    f = csub(fh, nh);
    if (ph.force_amp > 0.f && k2 > 0.f) f = cadd(f, stoch_force(row, col, kx, ky, k2, ph, step, dt, n));
  }
  rhs[i] = f;
}
void assemble_full(hipStream_t st, const float2* z, const float2* fhat, float2* rhs,
                   const Phys& ph, uint32_t step, double dt, int n) {
  hipLaunchKernelGGL(k_assemble_full, dim3(blocks((size_t)n * n, 256)), dim3(256), 0, st, z, fhat,
                     rhs, ph, step, (float)dt, n);
  check_launch("k_assemble_full");
}

// real physical field -> complex (imaginary part zero)
__global__ void k_real_to_complex(const float* __restrict__ r, float2* __restrict__ c, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) c[i] = make_float2(r[i], 0.f);
}
void real_to_complex(hipStream_t st, const float* r, float2* c, size_t n) {
  hipLaunchKernelGGL(k_real_to_complex, dim3(blocks(n, 256)), dim3(256), 0, st, r, c, n);
  check_launch("k_real_to_complex");
}

// ---------------------------------------------------------------------------------------------
// seeding: fillSpectrum (src/gpu/solver.js:384-418) on the device, FP64 like the JS host loop
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ double seed_amp(int kx, int ky, double k0) {
  const double kr = hypot((double)kx, (double)ky);
  const double env = (kr * kr) / (1.0 + pow(kr / k0, 8.0));
  return sqrt(env);
}

// per-column sum of amp^2 over the stored modes (k = 0 excluded)
__global__ void k_seed_colsum(RGeo g, double k0, double* __restrict__ colsum) {
  const int kx = blockIdx.x * blockDim.x + threadIdx.x;
  if (kx >= g.KC) return;
  double s = 0.0;
  for (int j = 0; j < g.NJ; ++j) {
    const int ky = j <= g.K ? j : j - g.NJ;
    if (kx == 0 && ky == 0) continue;
    const double a = seed_amp(kx, ky, k0);
    s += a * a;
  }
  colsum[kx] = kx == 0 ? s : 2.0 * s;  // kx > 0 columns also stand for their mirror
}

// scale = amplitude / (sqrt(sum)/N^2), or 0 when the sum is 0; one block, fixed order
__global__ void k_seed_scale(RGeo g, double amplitude, double* __restrict__ colsum) {
  __shared__ double sh[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < g.KC; i += 256) s += colsum[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double rms = sqrt(sh[0]) / ((double)g.n * (double)g.n);
    colsum[g.KC] = rms > 0.0 ? amplitude / rms : 0.0;
  }
}

// fillSpectrum value of the in-band mode (kx, ky) != 0 at full-grid position (row, col): magnitude
// sqrt(k^2 / (1 + (|k|/k0)^8)), phase 2 pi hash32(canonicalIdx ^ salt) / 2^32 (the mirror of a
// non-canonical mode gives the index, with the sine negated), stored as f32 and then rescaled by
// the FP64 factor sc (solver.js:396-417). A mode and its mirror come out as exact conjugates.
__device__ __forceinline__ float2 seed_mode(int n, int kx, int ky, int row, int col, double k0,
                                            uint32_t salt, double sc) {
  const double amp = seed_amp(kx, ky, k0);
  const bool canon = ky > 0 || (ky == 0 && kx > 0);
  const int hr = canon ? row : (n - row) % n;
  const int hc = canon ? col : (n - col) % n;
  const double sgn = canon ? 1.0 : -1.0;
  const double phase = 2.0 * M_PI * (double)hash1((uint32_t)(hr * n + hc) ^ salt) / 4294967296.0;
  float re = (float)(amp * cos(phase));
  float im = (float)(sgn * amp * sin(phase));
  if (sc > 0.0) {
    re = (float)((double)re * sc);
    im = (float)((double)im * sc);
  }
  return make_float2(re, im);
}

// the solver's compact half-spectrum layout
__global__ void k_seed_fill(RGeo g, double k0, uint32_t salt, const double* __restrict__ scale,
                            float2* __restrict__ st) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)g.KC * g.pitch) return;
  const int kx = (int)(i / g.pitch), j = (int)(i % g.pitch);
  float2 v = make_float2(0.f, 0.f);
  const int ky = j <= g.K ? j : j - g.NJ;
  if (j < g.NJ && !(kx == 0 && ky == 0))
    v = seed_mode(g.n, kx, ky, (ky + g.n) % g.n, kx, k0, salt, *scale);
  st[i] = v;
}

// the full N x N fixture layout (both halves, zero out of band)
__global__ void k_seed_fill_full(RGeo g, double k0, uint32_t salt, const double* __restrict__ scale,
                                 float2* __restrict__ full) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)g.n * g.n) return;
  const int row = (int)(i / g.n), col = (int)(i % g.n);
  const int kx = wn(col, g.n), ky = wn(row, g.n);
  float2 v = make_float2(0.f, 0.f);
  if (abs(kx) <= g.K && abs(ky) <= g.K && !(kx == 0 && ky == 0))
    v = seed_mode(g.n, kx, ky, row, col, k0, salt, *scale);
  full[i] = v;
}

void seed_spectrum(hipStream_t stm, float2* st, double* scratch_sums, int n, double k0,
                   double amplitude, uint32_t seed) {
  const RGeo g = rgeo(n);
  if (amplitude > 0.0) {
    hipLaunchKernelGGL(k_seed_colsum, dim3(blocks(g.KC, 64)), dim3(64), 0, stm, g, k0,
                       scratch_sums);
    check_launch("k_seed_colsum");
    hipLaunchKernelGGL(k_seed_scale, dim3(1), dim3(256), 0, stm, g, amplitude, scratch_sums);
    check_launch("k_seed_scale");
    hipLaunchKernelGGL(k_seed_fill, dim3(blocks((size_t)g.KC * g.pitch, 256)), dim3(256), 0, stm,
                       g, k0, hash1(seed), scratch_sums + g.KC, st);
    check_launch("k_seed_fill");
  } else {
    fill_u32(stm, st, (size_t)g.KC * g.pitch * 2, 0u);
  }
}

// ---------------------------------------------------------------------------------------------
// fillSpectrum into caller memory with no allocation (s2t_fill_spectrum_noalloc): the column
// sums and the scale live in this fixed device array, one copy per GPU, held until the process
// ends. Host code never passes the array's name (a host-side stand-in) to a kernel: it uses the
// device address that hipGetSymbolAddress gives on each device.
// ---------------------------------------------------------------------------------------------
constexpr int NOALLOC_MAX_N = 4096;
constexpr int NOALLOC_DOUBLES = NOALLOC_MAX_N / 3 + 3;  // KC column sums + the scale, at most
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_dev[NOALLOC_MAX_DEVICES] = {};

// device address of `scratch` on the current device, looked up on the first call there
static double* scratch_on_device() {
  int dev = 0;
  hipchk(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
    throw Fail(S2T_EINVAL, "device index " + std::to_string(dev) + " beyond the no-alloc table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    hipchk(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    g_scratch_dev[dev] = static_cast<double*>(p);
  }
  return g_scratch_dev[dev];
}

int noalloc_max_n() { return NOALLOC_MAX_N; }

void fill_spectrum_noalloc(hipStream_t stm, float2* full, int n, double k0, double amplitude,
                           uint32_t seed) {
  if (n > NOALLOC_MAX_N)
    throw Fail(S2T_EINVAL, "s2t_fill_spectrum_noalloc: grid size " + std::to_string(n) +
                               " exceeds its fixed scratch (at most " +
                               std::to_string(NOALLOC_MAX_N) + ")");
  const RGeo g = rgeo(n);
  if (!(amplitude > 0.0)) {
    fill_u32(stm, full, (size_t)n * n * 2, 0u);
    return;
  }
  double* sums = scratch_on_device();
  // every column sum and the scale are written before they are read: nothing to clear
  hipLaunchKernelGGL(k_seed_colsum, dim3(blocks(g.KC, 64)), dim3(64), 0, stm, g, k0, sums);
  check_launch("k_seed_colsum");
  hipLaunchKernelGGL(k_seed_scale, dim3(1), dim3(256), 0, stm, g, amplitude, sums);
  check_launch("k_seed_scale");
  hipLaunchKernelGGL(k_seed_fill_full, dim3(blocks((size_t)n * n, 256)), dim3(256), 0, stm, g, k0,
                     hash1(seed), sums + g.KC, full);
  check_launch("k_seed_fill_full");
}

// ---------------------------------------------------------------------------------------------
// diagnostics in FP64 with a fixed reduction order (readDiagnostics, solver.js:547-608;
// reference/spectral2d.py:92-108)
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ int isqrt_floor(long v) {
  if (v <= 0) return 0;
  long r = (long)sqrt((double)v);
  while (r * r > v) --r;
  while ((r + 1) * (r + 1) <= v) ++r;
  return (int)r;
}

// block b sums |w|^2 (weight 2 for kx > 0, the mirror half) over modes with round(|k|) = b,
// i.e. b^2 - b < k^2 <= b^2 + b (|k| is never a half integer)
__global__ void k_diag_bins(RGeo g, const float2* __restrict__ st, double* __restrict__ be,
                            double* __restrict__ bz) {
  __shared__ double she[256], shz[256];
  const int b = blockIdx.x;
  const long lo = b == 0 ? 0 : (long)b * b - b + 1, hi = (long)b * b + b;
  double se = 0.0, sz = 0.0;
  const int kxmax = min(g.K, b);
  for (int kx = threadIdx.x; kx <= kxmax; kx += 256) {
    const long kx2 = (long)kx * kx;
    const long a = lo - kx2, c = hi - kx2;
    if (c < 0) continue;
    int ky0 = a <= 0 ? 0 : isqrt_floor(a - 1) + 1;
    int ky1 = min(isqrt_floor(c), g.K);
    const double wgt = kx == 0 ? 1.0 : 2.0;
    for (int ky = ky0; ky <= ky1; ++ky) {
      const double k2 = (double)(kx2 + (long)ky * ky);
      for (int sgn = 0; sgn < (ky == 0 ? 1 : 2); ++sgn) {
        const int kys = sgn ? -ky : ky;
        const int j = kys >= 0 ? kys : kys + g.NJ;
        const float2 w = st[(size_t)kx * g.pitch + j];
        const double m2 = (double)w.x * w.x + (double)w.y * w.y;
        sz += wgt * m2;
        if (k2 > 0.0) se += wgt * m2 / k2;
      }
    }
  }
  she[threadIdx.x] = se;
  shz[threadIdx.x] = sz;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      she[threadIdx.x] += she[threadIdx.x + w];
      shz[threadIdx.x] += shz[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double n4 = (double)g.n * g.n * (double)g.n * g.n;
    be[b] = 0.5 * she[0] / n4;
    bz[b] = 0.5 * shz[0] / n4;
  }
}

__global__ void k_diag_totals(int nbins, const double* __restrict__ be,
                              const double* __restrict__ bz, double* __restrict__ out) {
  __shared__ double she[256], shz[256];
  double se = 0.0, sz = 0.0;
  for (int i = threadIdx.x; i < nbins; i += 256) {
    se += be[i];
    sz += bz[i];
  }
  she[threadIdx.x] = se;
  shz[threadIdx.x] = sz;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      she[threadIdx.x] += she[threadIdx.x + w];
      shz[threadIdx.x] += shz[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = she[0];
    out[1] = shz[0];
  }
}

// zonal mean of u in row y: (1/N^2) Re sum_ky u_hat(0, ky) exp(2 pi i ky y / N)
__global__ void k_diag_zonal(RGeo g, const float2* __restrict__ st,
                             const double2* __restrict__ wtab, double* __restrict__ zonal) {
  __shared__ double sh[256];
  const int y = blockIdx.x;
  double s = 0.0;
  for (int j = threadIdx.x; j < g.NJ; j += 256) {
    const int ky = j <= g.K ? j : j - g.NJ;
    if (ky == 0) continue;
    const float2 w = st[j];
    // u_hat = i ky w / ky^2 = i w / ky
    const double ur = -(double)w.y / ky, ui = (double)w.x / ky;
    const long m = (((long)ky * y) % g.n + g.n) % g.n;
    const double2 e = wtab[m];
    s += ur * e.x - ui * e.y;
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) zonal[y] = sh[0] / ((double)g.n * (double)g.n);
}

void diagnostics(hipStream_t stm, const float2* st, const double2* wtab, double* bins,
                 double* zonal, int n) {
  const RGeo g = rgeo(n);
  const int nb = n / 2 + 1;
  hipLaunchKernelGGL(k_diag_bins, dim3(nb), dim3(256), 0, stm, g, st, bins, bins + nb);
  check_launch("k_diag_bins");
  hipLaunchKernelGGL(k_diag_totals, dim3(1), dim3(256), 0, stm, nb, bins, bins + nb,
                     bins + 2 * nb);
  check_launch("k_diag_totals");
  hipLaunchKernelGGL(k_diag_zonal, dim3(n), dim3(256), 0, stm, g, st, wtab, zonal);
  check_launch("k_diag_zonal");
}

// ---------------------------------------------------------------------------------------------
// dye: BFECC with MacCormack limiter (src/shaders/dye.wgsl:1-113) on binary16 rgba texels with
// FP32 bilinear filtering and repeat addressing
// ---------------------------------------------------------------------------------------------
// Repeat addressing by masks: the grid and dye sizes are powers of two (dye_step checks), and
// i & (m - 1) is the non-negative residue of i mod m for negative i too (two's complement).
__device__ __forceinline__ float4 texel(const ushort4* __restrict__ tex, int m, int ix, int iy) {
  const int x = ix & (m - 1), y = iy & (m - 1);
  const ushort4 v = tex[(size_t)y * m + x];
  return make_float4(__half2float(__ushort_as_half(v.x)), __half2float(__ushort_as_half(v.y)),
                     __half2float(__ushort_as_half(v.z)), __half2float(__ushort_as_half(v.w)));
}
__device__ __forceinline__ float4 lerp4(float4 a, float4 b, float f) {
  return make_float4(a.x + (b.x - a.x) * f, a.y + (b.y - a.y) * f, a.z + (b.z - a.z) * f,
                     a.w + (b.w - a.w) * f);
}
// the same with alpha 1 (the dye's alpha is set to 1 after every step, so the fused kernel
// carries only colour through its intermediate values)
__device__ __forceinline__ float4 lerp3(float4 a, float4 b, float f) {
  return make_float4(a.x + (b.x - a.x) * f, a.y + (b.y - a.y) * f, a.z + (b.z - a.z) * f, 1.f);
}
// Bilinear sample at unit position (px, py) from a corner source `at4(ix, iy, c)` that fills
// c[0..3] with the texels (ix, iy), (ix+1, iy), (ix, iy+1), (ix+1, iy+1) (float4; repeat
// addressing is the source's), texel centres at (i + 0.5)/m. RGB: alpha is 1, not interpolated.
template <bool RGB = false, class At4>
__device__ __forceinline__ float4 sample4_at(At4&& at4, int m, float px, float py) {
  const float gx = px * (float)m - 0.5f, gy = py * (float)m - 0.5f;
  const float bx = floorf(gx), by = floorf(gy);
  const float fx = gx - bx, fy = gy - by;
  float4 c[4];
  at4((int)bx, (int)by, c);
  if constexpr (RGB) return lerp3(lerp3(c[0], c[1], fx), lerp3(c[2], c[3], fx), fy);
  return lerp4(lerp4(c[0], c[1], fx), lerp4(c[2], c[3], fx), fy);
}
// the same from a texel source `at(ix, iy)`
template <bool RGB = false, class At>
__device__ __forceinline__ float4 sample_at(At&& at, int m, float px, float py) {
  return sample4_at<RGB>([&](int ix, int iy, float4(&c)[4]) {
    c[0] = at(ix, iy);
    c[1] = at(ix + 1, iy);
    c[2] = at(ix, iy + 1);
    c[3] = at(ix + 1, iy + 1);
  }, m, px, py);
}
__device__ __forceinline__ float4 sample_tex(const ushort4* __restrict__ tex, int m, float px,
                                             float py) {
  return sample_at([&](int ix, int iy) { return texel(tex, m, ix, iy); }, m, px, py);
}
__device__ __forceinline__ float2 vel_at(const float2* __restrict__ vel, int n, int ix, int iy) {
  const int x = ix & (n - 1), y = iy & (n - 1);
  return vel[(size_t)y * n + x];
}
// bilinear velocity at unit position (px, py) from a corner source `va4(ix, iy, c)` (cells
// (ix, iy), (ix+1, iy), (ix, iy+1), (ix+1, iy+1) into c[0..3]; cell centres at (i + 0.5)/n,
// repeat addressing is the source's)
template <class Va4>
__device__ __forceinline__ float2 sample_vel4_at(Va4&& va4, int n, float px, float py) {
  const float gx = px * (float)n - 0.5f, gy = py * (float)n - 0.5f;
  const float bx = floorf(gx), by = floorf(gy);
  const float fx = gx - bx, fy = gy - by;
  float2 v[4];
  va4((int)bx, (int)by, v);
  const float2 a = make_float2(v[0].x + (v[1].x - v[0].x) * fx, v[0].y + (v[1].y - v[0].y) * fx);
  const float2 b = make_float2(v[2].x + (v[3].x - v[2].x) * fx, v[2].y + (v[3].y - v[2].y) * fx);
  return make_float2(a.x + (b.x - a.x) * fy, a.y + (b.y - a.y) * fy);
}
__device__ __forceinline__ float fract1(float x) { return x - floorf(x); }
// RK2 midpoint back-trace (dye.wgsl:40-45) through a corner source, from the velocity v1 at the
// point itself (its first sample, the same for a forward and a backward trace of the point)
template <class Va4>
__device__ __forceinline__ float2 trace4_from(Va4&& va4, int n, float px, float py, float2 v1,
                                              float h) {
  const float mx = fract1(px - 0.5f * v1.x * h + 1.f), my = fract1(py - 0.5f * v1.y * h + 1.f);
  const float2 v2 = sample_vel4_at(va4, n, mx, my);
  return make_float2(fract1(px - v2.x * h + 1.f), fract1(py - v2.y * h + 1.f));
}
template <class Va4>
__device__ __forceinline__ float2 trace4_at(Va4&& va4, int n, float px, float py, float h) {
  return trace4_from(va4, n, px, py, sample_vel4_at(va4, n, px, py), h);
}
// the same through a cell source `va(ix, iy)`
template <class Va>
__device__ __forceinline__ float2 trace_at(Va&& va, int n, float px, float py, float h) {
  return trace4_at([&](int ix, int iy, float2(&c)[4]) {
    c[0] = va(ix, iy);
    c[1] = va(ix + 1, iy);
    c[2] = va(ix, iy + 1);
    c[3] = va(ix + 1, iy + 1);
  }, n, px, py, h);
}
__device__ __forceinline__ float2 trace(const float2* __restrict__ vel, int n, float px, float py,
                                        float h) {
  return trace_at([=](int ix, int iy) { return vel_at(vel, n, ix, iy); }, n, px, py, h);
}
__device__ __forceinline__ ushort4 to_half4(float4 v) {
  return make_ushort4(__half_as_ushort(__float2half(v.x)), __half_as_ushort(__float2half(v.y)),
                      __half_as_ushort(__float2half(v.z)), __half_as_ushort(__float2half(v.w)));
}

__device__ __forceinline__ double dye_dt(const DevState* dev, double dt) {
  return dev ? dev->dt : dt;
}

// Texel of this thread. Consecutive workgroups go round-robin over the 8 XCDs, so each XCD is
// given a contiguous band of block rows (when the block count divides by 8): the bilinear gathers
// of neighbouring blocks read the same cache lines, which then stay in one L2.
__device__ __forceinline__ int2 dye_texel() {
  const int gx = (int)gridDim.x, total = gx * (int)gridDim.y;
  int b = (int)blockIdx.y * gx + (int)blockIdx.x;
  if (total % 8 == 0) b = (b % 8) * (total / 8) + b / 8;
  return make_int2((b % gx) * (int)blockDim.x + (int)threadIdx.x,
                   (b / gx) * (int)blockDim.y + (int)threadIdx.y);
}

// advect_fwd (sign +1) and advect_back (sign -1), dye.wgsl:58-74
__global__ void k_dye_advect(const ushort4* __restrict__ src, ushort4* __restrict__ dst,
                             const float2* __restrict__ vel, const DevState* dev, double dt,
                             float sign, int n, int m) {
  const int2 xy = dye_texel();
  const int x = xy.x, y = xy.y;
  if (x >= m || y >= m) return;
  const float h = sign * (float)dye_dt(dev, dt) / TWO_PI_F;
  // m is a power of two: the product by 1/m is the quotient, exactly
  const float im = 1.f / (float)m;
  const float px = ((float)x + 0.5f) * im, py = ((float)y + 0.5f) * im;
  const float2 s = trace(vel, n, px, py, h);
  dst[(size_t)y * m + x] = to_half4(sample_tex(src, m, s.x, s.y));
}

// The MacCormack limiter (clamp to the four phi0 texels around the foot), fade, splat dye and
// alpha 1 (dye.wgsl:86-112)
__device__ __forceinline__ float4 dye_correct(float4 corr, float4 c00, float4 c10, float4 c01,
                                              float4 c11, float fade, const Splats& sp, float px,
                                              float py) {
  const float4 lo = make_float4(fminf(fminf(c00.x, c10.x), fminf(c01.x, c11.x)),
                                fminf(fminf(c00.y, c10.y), fminf(c01.y, c11.y)),
                                fminf(fminf(c00.z, c10.z), fminf(c01.z, c11.z)),
                                fminf(fminf(c00.w, c10.w), fminf(c01.w, c11.w)));
  const float4 hi = make_float4(fmaxf(fmaxf(c00.x, c10.x), fmaxf(c01.x, c11.x)),
                                fmaxf(fmaxf(c00.y, c10.y), fmaxf(c01.y, c11.y)),
                                fmaxf(fmaxf(c00.z, c10.z), fmaxf(c01.z, c11.z)),
                                fmaxf(fmaxf(c00.w, c10.w), fmaxf(c01.w, c11.w)));
  float4 res = make_float4(fminf(fmaxf(corr.x, lo.x), hi.x) * fade,
                           fminf(fmaxf(corr.y, lo.y), hi.y) * fade,
                           fminf(fmaxf(corr.z, lo.z), hi.z) * fade,
                           fminf(fmaxf(corr.w, lo.w), hi.w) * fade);
  for (int i = 0; i < sp.count; ++i) {
    const Splat q = sp.s[i];
    const float rx = sep(px, q.x), ry = sep(py, q.y);
    const float sig2 = fmaxf(q.radius * q.radius, 1e-8f);
    const float r2 = rx * rx + ry * ry;
    if (r2 < 9.f * sig2) {
      const float w = q.dye_amp * expf(-r2 / sig2);
      res.x += q.r * w;
      res.y += q.g * w;
      res.z += q.b * w;
    }
  }
  res.w = 1.f;
  return res;
}

// advect_corr, dye.wgsl:76-113
__global__ void k_dye_corr(const ushort4* __restrict__ src, const ushort4* __restrict__ aux,
                           ushort4* __restrict__ dst, const float2* __restrict__ vel,
                           const DevState* dev, double dt, float fade, Splats sp, int n, int m) {
  const int2 xy = dye_texel();
  const int x = xy.x, y = xy.y;
  if (x >= m || y >= m) return;
  const float h = (float)dye_dt(dev, dt) / TWO_PI_F;
  // m is a power of two: the product by 1/m is the quotient, exactly
  const float im = 1.f / (float)m;
  const float px = ((float)x + 0.5f) * im, py = ((float)y + 0.5f) * im;
  const float2 s = trace(vel, n, px, py, h);
  const float4 a = sample_tex(src, m, s.x, s.y);
  const float4 b = sample_tex(aux, m, s.x, s.y);
  float4 corr = make_float4(1.5f * a.x - 0.5f * b.x, 1.5f * a.y - 0.5f * b.y,
                            1.5f * a.z - 0.5f * b.z, 1.5f * a.w - 0.5f * b.w);
  const float gx = s.x * (float)m - 0.5f, gy = s.y * (float)m - 0.5f;
  const int ix = (int)floorf(gx), iy = (int)floorf(gy);
  const float4 c00 = texel(src, m, ix, iy), c10 = texel(src, m, ix + 1, iy);
  const float4 c01 = texel(src, m, ix, iy + 1), c11 = texel(src, m, ix + 1, iy + 1);
  dst[(size_t)y * m + x] = to_half4(dye_correct(corr, c00, c10, c01, c11, fade, sp, px, py));
}


// ---------------------------------------------------------------------------------------------
// The BFECC update fused into one kernel over 32 x 32 tiles of the dye. A tile's output needs
// phi_hat at the trace feet of its texels, phi_hat needs phi1 at its own back-trace feet and phi1
// needs phi0 at its forward-trace feet. A foot less than one texel from its texel centre lies in
// the bilinear stencil of that texel and its neighbours, so phi0, phi1 and phi_hat are staged in
// LDS on the tile grown by 3, 2 and 1 texels (phi0 loaded, phi1 and phi_hat computed there,
// rounded to binary16 like the three-pass textures), with the velocity cells under the traced
// points. CFL 0.5 on twice the spectral grid moves a point at most one texel per step, and less
// wherever |u| and |v| do not peak together. Each tile checks the bound on its staged velocity;
// where it fails (that maximum, a dt_min clamp, a fixed dt) every sample outside a staged region
// is computed on the spot from the same functions, so the result is the three-pass result bit
// for bit whatever the step.
// ---------------------------------------------------------------------------------------------
constexpr int DYE_T = 32;
constexpr int DYE_E2 = 1, DYE_E1 = 2, DYE_E0 = 3;
constexpr int DYE_R0 = DYE_T + 2 * DYE_E0, DYE_R1 = DYE_T + 2 * DYE_E1, DYE_R2 = DYE_T + 2 * DYE_E2;
constexpr int DYE_THREADS = 256;
// velocity cells staged for dye scale r: the cells under the points traced (the phi1 region, the
// widest) and their midpoints, with a cell of margin each side
template <int R>
struct DyeVel {
  static constexpr int EV = DYE_E1 + 1;                  // dye texels around the tile
  static constexpr int LO = (EV + R - 1) / R + 1;        // cells before the tile's first cell
  static constexpr int CELLS = (DYE_T + R - 1) / R + 2 * LO + 1;
};

// colour of a binary16 texel, alpha 1
__device__ __forceinline__ float4 half3(ushort4 v) {
  return make_float4(__half2float(__ushort_as_half(v.x)), __half2float(__ushort_as_half(v.y)),
                     __half2float(__ushort_as_half(v.z)), 1.f);
}
// colour of a point staged in LDS as three FP32 planes per row: red at q[0], green at q[W], blue
// at q[2 W] (W points per row), alpha 1
template <int W>
__device__ __forceinline__ float4 rgb_at(const float* q) {
  return make_float4(q[0], q[W], q[2 * W], 1.f);
}
// binary16 of a colour, alpha 1
__device__ __forceinline__ ushort4 to_half3(float4 v) {
  return make_ushort4(__half_as_ushort(__float2half(v.x)), __half_as_ushort(__float2half(v.y)),
                      __half_as_ushort(__float2half(v.z)), 0x3c00);
}

// The stages of one tile, after phi0 and the velocity cells are staged. CHECKED: every lookup
// checks its staged region and computes a sample outside it on the spot; otherwise (the tile's
// feet are known to stay within one texel) lookups go straight to LDS, the four corners of a
// stencil from one address. Only colour is carried: every output has alpha 1 whatever the alpha
// of the intermediate values, so they are given alpha 1 (the colour is the three-pass colour bit
// for bit).
template <int R, bool CHECKED>
__device__ __forceinline__ void dye_tile(const ushort4* __restrict__ src, ushort4* __restrict__ dst,
                                         const float2* __restrict__ vel, const float* s0,
                                         ushort4* s1, ushort4* s2, const float2* sv,
                                         float2* sfoot, float hf,
                                         float hb, float fade, const Splats& sp, int n, int m,
                                         int x0, int y0, int vx0, int vy0) {
  using V = DyeVel<R>;
  const int tid = (int)threadIdx.x;
  const int mask = m - 1, nmask = n - 1;
  // A staged region holds its points at relative index (i - origin) & mask in [0, size) (a grid
  // smaller than the region is staged periodically), so with the stencil's first corner inside,
  // its neighbours are the next index and the next row.
  auto va = [=](int ix, int iy) -> float2 {
    const int rx = (ix - vx0) & nmask, ry = (iy - vy0) & nmask;
    if (!CHECKED || (rx < V::CELLS && ry < V::CELLS)) return sv[ry * V::CELLS + rx];
    return vel_at(vel, n, ix, iy);
  };
  auto va4 = [=](int ix, int iy, float2(&c)[4]) {
    if constexpr (!CHECKED) {
      const float2* q = sv + ((iy - vy0) & nmask) * V::CELLS + ((ix - vx0) & nmask);
      c[0] = q[0];
      c[1] = q[1];
      c[2] = q[V::CELLS];
      c[3] = q[V::CELLS + 1];
    } else {
      c[0] = va(ix, iy);
      c[1] = va(ix + 1, iy);
      c[2] = va(ix, iy + 1);
      c[3] = va(ix + 1, iy + 1);
    }
  };
  auto phi0 = [=](int ix, int iy) -> float4 {
    const int rx = (ix - x0 + DYE_E0) & mask, ry = (iy - y0 + DYE_E0) & mask;
    if (!CHECKED || (rx < DYE_R0 && ry < DYE_R0)) return rgb_at<DYE_R0>(s0 + ry * 3 * DYE_R0 + rx);
    return half3(src[(size_t)(iy & mask) * m + (ix & mask)]);
  };
  auto phi0_4 = [=](int ix, int iy, float4(&c)[4]) {
    if constexpr (!CHECKED) {
      const float* q = s0 + ((iy - y0 + DYE_E0) & mask) * 3 * DYE_R0 + ((ix - x0 + DYE_E0) & mask);
      c[0] = rgb_at<DYE_R0>(q);
      c[1] = rgb_at<DYE_R0>(q + 1);
      c[2] = rgb_at<DYE_R0>(q + 3 * DYE_R0);
      c[3] = rgb_at<DYE_R0>(q + 3 * DYE_R0 + 1);
    } else {
      c[0] = phi0(ix, iy);
      c[1] = phi0(ix + 1, iy);
      c[2] = phi0(ix, iy + 1);
      c[3] = phi0(ix + 1, iy + 1);
    }
  };
  // texel centre; m is a power of two, so the product by 1/m is the quotient, exactly
  const float im = 1.f / (float)m;
  auto centre = [=](int i) { return ((float)(i & mask) + 0.5f) * im; };
  // phi1 of texel (ix, iy): the forward advection of phi0, rounded to binary16
  auto phi1_of = [=](int ix, int iy) -> ushort4 {
    const float2 f = trace4_at(va4, n, centre(ix), centre(iy), hf);
    return to_half3(sample4_at<true>(phi0_4, m, f.x, f.y));
  };
  // the same at point (rx, ry) of the phi1 region from the velocity v1 at the point, keeping the
  // forward foot of the tile's own texels (the trace of the correction stage)
  auto phi1_keep = [=](int rx, int ry, float2 v1) -> ushort4 {
    const int ix = x0 - DYE_E1 + rx, iy = y0 - DYE_E1 + ry;
    const float2 f = trace4_from(va4, n, centre(ix), centre(iy), v1, hf);
    const int tx = rx - DYE_E1, ty = ry - DYE_E1;
    if (tx >= 0 && tx < DYE_T && ty >= 0 && ty < DYE_T) sfoot[ty * DYE_T + tx] = f;
    return to_half3(sample4_at<true>(phi0_4, m, f.x, f.y));
  };
  auto phi1 = [=](int ix, int iy) -> float4 {
    const int rx = (ix - x0 + DYE_E1) & mask, ry = (iy - y0 + DYE_E1) & mask;
    if (!CHECKED || (rx < DYE_R1 && ry < DYE_R1)) return half3(s1[ry * DYE_R1 + rx]);
    return half3(phi1_of(ix, iy));
  };
  auto phi1_4 = [=](int ix, int iy, float4(&c)[4]) {
    if constexpr (!CHECKED) {
      const ushort4* q = s1 + ((iy - y0 + DYE_E1) & mask) * DYE_R1 + ((ix - x0 + DYE_E1) & mask);
      c[0] = half3(q[0]);
      c[1] = half3(q[1]);
      c[2] = half3(q[DYE_R1]);
      c[3] = half3(q[DYE_R1 + 1]);
    } else {
      c[0] = phi1(ix, iy);
      c[1] = phi1(ix + 1, iy);
      c[2] = phi1(ix, iy + 1);
      c[3] = phi1(ix + 1, iy + 1);
    }
  };
  // phi1 without the staged values (once phi_hat has taken their space)
  auto phi1_fresh = [=](int ix, int iy) -> float4 { return half3(phi1_of(ix, iy)); };
  auto phi1_fresh4 = [=](int ix, int iy, float4(&c)[4]) {
    c[0] = phi1_fresh(ix, iy);
    c[1] = phi1_fresh(ix + 1, iy);
    c[2] = phi1_fresh(ix, iy + 1);
    c[3] = phi1_fresh(ix + 1, iy + 1);
  };
  // phi_hat of texel (ix, iy): the backward advection of phi1 (from the corner source p1),
  // rounded to binary16
  auto phih_from = [=](auto p1, int ix, int iy) -> ushort4 {
    const float2 f = trace4_at(va4, n, centre(ix), centre(iy), hb);
    return to_half3(sample4_at<true>(p1, m, f.x, f.y));
  };
  auto phih = [=](int ix, int iy) -> float4 {
    const int rx = (ix - x0 + DYE_E2) & mask, ry = (iy - y0 + DYE_E2) & mask;
    if (!CHECKED || (rx < DYE_R2 && ry < DYE_R2)) return half3(s2[ry * DYE_R2 + rx]);
    return half3(phih_from(phi1_fresh4, ix, iy));
  };
  auto phih4 = [=](int ix, int iy, float4(&c)[4]) {
    if constexpr (!CHECKED) {
      const ushort4* q = s2 + ((iy - y0 + DYE_E2) & mask) * DYE_R2 + ((ix - x0 + DYE_E2) & mask);
      c[0] = half3(q[0]);
      c[1] = half3(q[1]);
      c[2] = half3(q[DYE_R2]);
      c[3] = half3(q[DYE_R2 + 1]);
    } else {
      c[0] = phih(ix, iy);
      c[1] = phih(ix + 1, iy);
      c[2] = phih(ix, iy + 1);
      c[3] = phih(ix + 1, iy + 1);
    }
  };
  // The phi1 stage as NALL items: items below NIN are the points of the phi_hat region (row-major,
  // one texel in from the phi1 region's edge), the rest the ring of the phi1 region around it.
  // Item i of a thread is point i of the phi_hat stage as well, so the velocity at the point (the
  // first sample of both its forward and its backward trace) stays in registers between the two.
  constexpr int NIN = DYE_R2 * DYE_R2, NALL = DYE_R1 * DYE_R1;
  constexpr int NH = (NIN + DYE_THREADS - 1) / DYE_THREADS;   // phi_hat points per thread
  constexpr int NI = (NALL + DYE_THREADS - 1) / DYE_THREADS;  // phi1 items per thread
  constexpr int ED = DYE_E1 - DYE_E2;
  static_assert(ED == 1, "the ring is one texel wide");
  auto item = [](int i) -> int2 {
    if (i < NIN) return make_int2(i % DYE_R2 + ED, i / DYE_R2 + ED);
    const int r = i - NIN;
    return r < DYE_R1       ? make_int2(r, 0)
         : r < 2 * DYE_R1   ? make_int2(r - DYE_R1, DYE_R1 - 1)
         : r < 2 * DYE_R1 + DYE_R2 ? make_int2(0, r - 2 * DYE_R1 + ED)
                                   : make_int2(DYE_R1 - 1, r - 2 * DYE_R1 - DYE_R2 + ED);
  };
  float2 v1k[NH];
#pragma unroll
  for (int k = 0; k < NI; ++k) {
    const int i = tid + k * DYE_THREADS;
    // every thread has the first items (the compiler then interleaves their independent chains)
    if ((k + 1) * DYE_THREADS <= NALL || i < NALL) {
      const int2 p = item(i);
      const float2 v1 = sample_vel4_at(va4, n, centre(x0 - DYE_E1 + p.x), centre(y0 - DYE_E1 + p.y));
      if (k < NH) v1k[k] = v1;
      s1[p.y * DYE_R1 + p.x] = phi1_keep(p.x, p.y, v1);
    }
  }
  __syncthreads();
  // phi_hat into registers (all of a thread's points at once: independent chains), then over
  // phi1's LDS space once every thread is done reading phi1
  ushort4 hv[NH];
#pragma unroll
  for (int k = 0; k < NH; ++k) {
    const int i = tid + k * DYE_THREADS;
    if (i < NIN) {
      const int ix = x0 - DYE_E2 + i % DYE_R2, iy = y0 - DYE_E2 + i / DYE_R2;
      const float2 f = trace4_from(va4, n, centre(ix), centre(iy), v1k[k], hb);
      hv[k] = to_half3(sample4_at<true>(phi1_4, m, f.x, f.y));
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NH; ++k) {
    const int i = tid + k * DYE_THREADS;
    if (i < DYE_R2 * DYE_R2) s2[i] = hv[k];
  }
  __syncthreads();
  auto out = [&](int i) {
    const int x = x0 + i % DYE_T, y = y0 + i / DYE_T;
    if (x >= m || y >= m) return;
    const float px = ((float)x + 0.5f) * im, py = ((float)y + 0.5f) * im;
    const float2 s = sfoot[i];  // trace_at(va, n, px, py, hf), from the phi1 stage
    // the phi0 stencil at the foot: the advected value and the limiter's bounds
    float4 c[4];
    const float4 a = sample4_at<true>([&](int ix, int iy, float4(&cc)[4]) {
      phi0_4(ix, iy, cc);
      c[0] = cc[0];
      c[1] = cc[1];
      c[2] = cc[2];
      c[3] = cc[3];
    }, m, s.x, s.y);
    const float4 bb = sample4_at<true>(phih4, m, s.x, s.y);
    const float4 corr = make_float4(1.5f * a.x - 0.5f * bb.x, 1.5f * a.y - 0.5f * bb.y,
                                    1.5f * a.z - 0.5f * bb.z, 1.f);
    dst[(size_t)y * m + x] = to_half4(dye_correct(corr, c[0], c[1], c[2], c[3], fade, sp, px, py));
  };
  for (int i = tid; i < DYE_T * DYE_T; i += 2 * DYE_THREADS) {
    out(i);
    if (i + DYE_THREADS < DYE_T * DYE_T) out(i + DYE_THREADS);
  }
}

// R = m / n, the dye scale (1 or 2)
template <int R>
__global__ void __launch_bounds__(DYE_THREADS)
k_dye_fused(const ushort4* __restrict__ src, ushort4* __restrict__ dst,
            const float2* __restrict__ vel, const DevState* dev, double dt, float fade, Splats sp,
            int n, int m) {
  using V = DyeVel<R>;
  static_assert(DYE_THREADS % 64 == 0, "whole waves");
  // phi0 in FP32 (converted once here rather than at each of the stencils that read it): per row,
  // the red, green and blue planes of its DYE_R0 points
  __shared__ float s0[3 * DYE_R0 * DYE_R0];
  // phi_hat (R2 x R2) takes phi1's space (R1 x R1) once phi1 is no longer read
  __shared__ ushort4 s1[DYE_R1 * DYE_R1];
  ushort4* s2 = s1;
  static_assert(DYE_R2 <= DYE_R1, "phi_hat fits in phi1's space");
  __shared__ float2 sv[V::CELLS * V::CELLS];
  __shared__ float2 sfoot[DYE_T * DYE_T];
  __shared__ float wmax[DYE_THREADS / 64];
  // tiles: bands of tile rows per XCD (consecutive workgroups go round-robin over the 8 XCDs;
  // neighbouring tiles share their halos through one L2)
  const int tiles_x = (m + DYE_T - 1) / DYE_T, total = tiles_x * tiles_x;
  int b = (int)blockIdx.x;
  if ((total & 7) == 0) b = (b & 7) * (total >> 3) + (b >> 3);
  const int x0 = (b % tiles_x) * DYE_T, y0 = (b / tiles_x) * DYE_T;
  const int tid = (int)threadIdx.x;
  const int mask = m - 1, nmask = n - 1;
  const int vx0 = x0 / R - V::LO, vy0 = y0 / R - V::LO;
  const float dtf = (float)dye_dt(dev, dt);
  const float hf = 1.f * dtf / TWO_PI_F, hb = -1.f * dtf / TWO_PI_F;

  // stage the velocity cells (and their largest component) and phi0
  float vmax = 0.f;
  for (int i = tid; i < V::CELLS * V::CELLS; i += DYE_THREADS) {
    const int rx = i % V::CELLS, ry = i / V::CELLS;
    const float2 v = vel[(size_t)((vy0 + ry) & nmask) * n + ((vx0 + rx) & nmask)];
    sv[i] = v;
    vmax = fmaxf(vmax, fmaxf(fabsf(v.x), fabsf(v.y)));
  }
  for (int i = tid; i < DYE_R0 * DYE_R0; i += DYE_THREADS) {
    const int rx = i % DYE_R0, ry = i / DYE_R0;
    const float4 c = half3(src[(size_t)((y0 - DYE_E0 + ry) & mask) * m + ((x0 - DYE_E0 + rx) & mask)]);
    float* q = s0 + ry * 3 * DYE_R0 + rx;
    q[0] = c.x;
    q[DYE_R0] = c.y;
    q[2 * DYE_R0] = c.z;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) vmax = fmaxf(vmax, __shfl_xor(vmax, off));
  if ((tid & 63) == 0) wmax[tid >> 6] = vmax;
  __syncthreads();
  vmax = wmax[0];
#pragma unroll
  for (int w = 1; w < DYE_THREADS / 64; ++w) vmax = fmaxf(vmax, wmax[w]);
  // Bilinear velocities never exceed the largest staged component, so every foot of the tile's
  // traces (and every midpoint) lies within vmax |h| m texels of its point; below one texel, with
  // room for rounding, all lookups stay in the staged regions. A non-finite velocity fails the
  // test and takes the checked path.
  const bool inside = vmax * fabsf(hf) * (float)m < 0.999f;
  if (inside)
    dye_tile<R, false>(src, dst, vel, s0, s1, s2, sv, sfoot, hf, hb, fade, sp, n, m, x0, y0, vx0,
                       vy0);
  else
    dye_tile<R, true>(src, dst, vel, s0, s1, s2, sv, sfoot, hf, hb, fade, sp, n, m, x0, y0, vx0,
                      vy0);
}

void dye_step(hipStream_t st, ushort4* const dye[3], int parity, const float2* vel,
              const DevState* dev, double dt, float fade, const Splats& sp, int n, int m) {
  if (n <= 0 || m <= 0 || (n & (n - 1)) != 0 || (m & (m - 1)) != 0 || (m != n && m != 2 * n))
    throw Fail(S2T_EINVAL, "dye step: grid and dye sizes must be powers of two, dye = 1 or 2 x grid");
  const int tiles_x = (m + DYE_T - 1) / DYE_T;
  // bind-group parity of src/gpu/solver.js:299-303: read dye[p], write dye[1 - p]
  if (m == 2 * n)
    hipLaunchKernelGGL(k_dye_fused<2>, dim3(tiles_x * tiles_x), dim3(DYE_THREADS), 0, st,
                       dye[parity], dye[1 - parity], vel, dev, dt, fade, sp, n, m);
  else
    hipLaunchKernelGGL(k_dye_fused<1>, dim3(tiles_x * tiles_x), dim3(DYE_THREADS), 0, st,
                       dye[parity], dye[1 - parity], vel, dev, dt, fade, sp, n, m);
  check_launch("k_dye_fused");
}

// The unfused passes, one launch each (the structure of dye.wgsl:58-113): the fused kernel's
// result, bit for bit (test hook).
void dye_step_passes(hipStream_t st, ushort4* const dye[3], int parity, const float2* vel,
                     const DevState* dev, double dt, float fade, const Splats& sp, int n, int m) {
  if (n <= 0 || m <= 0 || (n & (n - 1)) != 0 || (m & (m - 1)) != 0)
    throw Fail(S2T_EINVAL, "dye step: grid and dye sizes must be powers of two");
  const dim3 blk(16, 16), grd((m + 15) / 16, (m + 15) / 16);
  const int p = parity;
  // bind-group parity of src/gpu/solver.js:299-303
  hipLaunchKernelGGL(k_dye_advect, grd, blk, 0, st, dye[p], dye[1 - p], vel, dev, dt, 1.f, n, m);
  check_launch("k_dye_advect(fwd)");
  hipLaunchKernelGGL(k_dye_advect, grd, blk, 0, st, dye[1 - p], dye[2], vel, dev, dt, -1.f, n, m);
  check_launch("k_dye_advect(back)");
  hipLaunchKernelGGL(k_dye_corr, grd, blk, 0, st, dye[p], dye[2], dye[1 - p], vel, dev, dt, fade,
                     sp, n, m);
  check_launch("k_dye_corr");
}

// toHalf of src/gpu/solver.js:20-41 (mantissa truncation)
__device__ __forceinline__ uint16_t to_half_trunc(float value) {
  const uint32_t x = __float_as_uint(value);
  const uint32_t sign = (x >> 16) & 0x8000u;
  int e = (int)((x >> 23) & 0xffu);
  uint32_t mant = x & 0x7fffffu;
  if (e == 255) return (uint16_t)(sign | 0x7c00u | (mant ? 0x200u : 0u));
  e = e - 127 + 15;
  if (e >= 31) return (uint16_t)(sign | 0x7c00u);
  if (e <= 0) {
    if (e < -10) return (uint16_t)sign;
    mant = (mant | 0x800000u) >> (1 - e);
    return (uint16_t)(sign | (mant >> 13));
  }
  return (uint16_t)(sign | ((uint32_t)e << 10) | (mant >> 13));
}

// paintDye('bands') of src/gpu/solver.js:446-473, FP64 like the JS loop
__global__ void k_dye_paint(ushort4* __restrict__ dst, int m) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
  if (x >= m || y >= m) return;
  const double fy = (double)y / m, fx = (double)x / m;
  const double band = 0.5 + 0.5 * tanh(6.0 * sin(fy * M_PI * 12.0));
  const double mix = 0.5 + 0.5 * sin(fx * M_PI * 2.0);
  const double warm[3] = {1.0, 0.45, 0.22}, cool[3] = {0.28, 0.62, 1.0};
  uint16_t c[3];
  for (int k = 0; k < 3; ++k)
    c[k] = to_half_trunc((float)(band * (warm[k] * mix + cool[k] * (1.0 - mix)) * 0.9));
  dst[(size_t)y * m + x] = make_ushort4(c[0], c[1], c[2], to_half_trunc(1.0f));
}
void dye_paint(hipStream_t st, ushort4* dst, int m) {
  const dim3 blk(16, 16), grd((m + 15) / 16, (m + 15) / 16);
  hipLaunchKernelGGL(k_dye_paint, grd, blk, 0, st, dst, m);
  check_launch("k_dye_paint");
}

}  // namespace s2t
