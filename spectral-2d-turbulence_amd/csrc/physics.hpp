// Shared device definitions: data layout in HBM, parameters, the integer hash and the pointwise
// physics used both by the fused step kernels and by the test-hook kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "line_fft.hpp"

namespace s2t {

// ---------------------------------------------------------------------------------------------
// Layout.
//
// The vorticity state is the spectrum of a real field, masked to the 2/3 band
// (|kx|, |ky| <= K = floor(N/3), reference/spectral2d.py:22-23, common.wgsl:69-74). Only the
// kx >= 0 half is stored (the kx < 0 half is its Hermitian mirror), transposed so that one
// column kx is contiguous, and only the 2K+1 in-band rows:
//   state[kx * PITCH + j], kx = 0..K, j = 0..2K; ky = j for j <= K, ky = j - (2K+1) otherwise.
// Mixed-space fields (kx spectral, y physical) between the row and column kernels are stored for
// kx = 0..K in 4-row tiles so that both a row and a group of columns read long runs:
//   tile(y, kx, f) = ((y/4 * KC + kx) * F + f) * 4 + y%4  with F fields per tile.
// ---------------------------------------------------------------------------------------------
template <int N>
struct Geo {
  static constexpr int K = N / 3;             // N is never a multiple of 3
  static constexpr int KC = K + 1;            // stored columns
  static constexpr int NJ = 2 * K + 1;        // stored rows
  static constexpr int PITCH = (NJ + 63) / 64 * 64;
  static constexpr size_t STATE = (size_t)KC * PITCH;   // complex values per state field
  static constexpr int KP = (KC + 15) / 16 * 16;        // mixed-field row pitch (128-byte lines)
  static constexpr size_t PLANE = (size_t)N * KP;       // complex values per mixed field
  // compact row of row index i in [0, N); -1 when out of band
  static __device__ __forceinline__ int jrow(int i) {
    return i <= K ? i : (i >= N - K ? i - (N - NJ) : -1);
  }
  // Mixed field f of an F-field buffer as planes of RT-row tiles [f][y/RT][kx][RT]: the values of
  // rows RT m .. RT m + RT - 1 at one kx are adjacent. The row kernel works on row pairs: a pair is
  // one 16-byte element per kx, at a stride of RT * 8 bytes; the pairs of one tile are processed by
  // workgroups that run back to back on one XCD, so they share the tile's cache lines in L2. A
  // column group of the column kernel touches one cache line per RT rows instead of one per row.
  // 32-bit element offsets (every buffer here holds fewer than 2^31 elements).
  static constexpr int RT = 4;
  static constexpr unsigned KXB = RT * 8;  // bytes from kx to kx + 1 within a tile row
  template <int F>
  static __device__ __forceinline__ unsigned tile(int y, int kx, int f) {
    return (unsigned)f * (unsigned)PLANE + ((unsigned)(y / RT) * KP + (unsigned)kx) * RT +
           (unsigned)(y % RT);
  }
  // byte offset of (row y, kx) within a plane
  static __device__ __forceinline__ unsigned off(int y, int kx) {
    return (((unsigned)(y / RT) * KP + (unsigned)kx) * RT + (unsigned)(y % RT)) * 8u;
  }
};

__host__ __device__ __forceinline__ int wn(int i, int n) { return i < n / 2 ? i : i - n; }

// ---------------------------------------------------------------------------------------------
// parameters (by value in kernel arguments)
// ---------------------------------------------------------------------------------------------
struct Phys {
  double beta, mu, nu;   // nu = hyperviscosity coefficient nu_h
  double p;              // dissipation order
  int p_int;             // p as an integer (1 .. 16)
  float force_amp, force_k, force_width, force_norm;
  // |k|^2 outside [force_k2lo, force_k2hi] has envelope exp(-x^2) < 1e-4 by a margin
  // (|x| >= 3.05 > sqrt(ln 1e4) = 3.035): skipped without evaluating it
  float force_k2lo, force_k2hi;
  double cfl, dt_min, dt_max, dx, fixed_dt;
  float dye_fade;
};

struct Splat {
  float x, y, dx, dy, r, g, b, radius, amp, dye_amp;
};
struct Splats {
  int count;
  Splat s[8];
};

// Per-instance device scalars. dt and dt-derived values never leave the device during stepping
// (README.md:132-134 of the reference); the host reads them with the diagnostics. Only the
// per-step epilogue (step_epilogue, run by the last column workgroup of stage 2 to finish, after
// every other workgroup of that launch has counted itself done) advances step, t and steps_left,
// so no workgroup reads a field that another workgroup of the same launch is writing.
struct DevState {
  double dt;            // step of the step in flight / last step
  double t;             // simulated time
  float speed;          // max |u|+|v| of the last step's initial field
  // running max of |u|+|v| for the step in flight (float bits, non-negative), spread over slots
  // by row block so the per-workgroup atomics of a launch do not all hit one address
  static constexpr int SPEED_SLOTS = 64;
  unsigned speed_bits[SPEED_SLOTS];
  unsigned step;        // steps completed since the last reset = stepIndex of the step in flight
  unsigned nonfinite;   // sticky flag: a non-finite speed was seen
  int active;           // 0: the step kernels skip this instance
  int steps_left;       // steps until the instance retires itself (< 0: no limit)
  unsigned cols_done;   // column workgroups of the last stage that have finished (then reset)
};

// Per-step epilogue of one instance: advance stepIndex and time, clear the speed maximum for the
// next step and retire the instance when its step budget is spent (stepIndex++ of solver.js:520).
// Run by the last column workgroup of stage 2 to finish (which also clears speed_bits).
__device__ __forceinline__ void step_epilogue(DevState& d) {
  d.step += 1u;
  d.t += d.dt;
  if (d.steps_left > 0) {
    d.steps_left -= 1;
    if (d.steps_left == 0) d.active = 0;
  }
}

// One simulation instance's parameters and scalars, an array of these per solver (a plain solver
// has one; a batch has one per member). Kernels index it with blockIdx.y.
struct Inst {
  Phys ph;
  DevState st;
};

// ---------------------------------------------------------------------------------------------
// integer hash (src/shaders/common.wgsl:76-84, src/gpu/solver.js:43-47)
// ---------------------------------------------------------------------------------------------
__host__ __device__ __forceinline__ uint32_t hash1(uint32_t x) {
  uint32_t v = x * 747796405u + 2891336453u;
  v = ((v >> ((v >> 28u) + 4u)) ^ v) * 277803737u;
  return (v >> 22u) ^ v;
}
__device__ __forceinline__ float hashf(uint32_t a, uint32_t b) {
  return (float)hash1(a ^ hash1(b)) * 2.3283064365386963e-10f;
}

constexpr float TWO_PI_F = 6.28318530717958647f;

// ---------------------------------------------------------------------------------------------
// pointwise physics
// ---------------------------------------------------------------------------------------------

// u_hat = i ky w / k^2, v_hat = -i kx w / k^2 (solver.wgsl velocity_pack:21-40)
__device__ __forceinline__ void vel_pack(float2 w, float kx, float ky, float inv, float2& uh,
                                         float2& vh) {
  const float2 iw = cmuli(w);
  uh = cscale(iw, ky * inv);
  vh = cscale(iw, -kx * inv);
}

// Hermitian unpack of a transform of two real fields: q1 = (z + conj zm)/2, q2 = -i(z - conj zm)/2
__device__ __forceinline__ void herm_unpack(float2 z, float2 zm, float2& q1, float2& q2) {
  q1 = make_float2(0.5f * (z.x + zm.x), 0.5f * (z.y - zm.y));
  q2 = make_float2(0.5f * (z.y + zm.y), 0.5f * (zm.x - z.x));
}

// N_hat = i kx q1 + i ky q2 (solver.wgsl assemble:156)
__device__ __forceinline__ float2 nhat(float2 q1, float2 q2, float kx, float ky) {
  return make_float2(kx * (-q1.y) + ky * (-q2.y), kx * q1.x + ky * q2.x);
}

// Stochastic forcing of one in-band mode (solver.wgsl assemble:161-176). row, col are the full
// grid indices of the mode; the canonical half is ky > 0, or ky = 0 and kx > 0.
__device__ __forceinline__ float2 stoch_force(int row, int col, float kx, float ky, float k2,
                                              const Phys& ph, uint32_t step, float dt, int n) {
  if (k2 < ph.force_k2lo || k2 > ph.force_k2hi) return make_float2(0.f, 0.f);
  const float x = (sqrtf(k2) - ph.force_k) / ph.force_width;
  const float env = expf(-(x * x));
  if (!(env > 1e-4f)) return make_float2(0.f, 0.f);
  int hr = row, hc = col;
  float sgn = 1.f;
  if (!(ky > 0.f || (ky == 0.f && kx > 0.f))) {
    hr = (n - row) % n;
    hc = (n - col) % n;
    sgn = -1.f;
  }
  const float phase = TWO_PI_F * hashf((uint32_t)(hr * n + hc), step);
  const float amp = ph.force_amp * env * ph.force_norm / sqrtf(fmaxf(dt, 1e-6f));
  float s, c;
  sincosf(phase, &s, &c);
  return make_float2(amp * c, sgn * amp * s);
}

// CFL step from the grid max of |u|+|v| (src/shaders/reduce.wgsl:82-88)
__device__ __forceinline__ double cfl_dt(float speed, const Phys& ph) {
  double dt = ph.dt_max;
  if (speed > 1e-8f) dt = fmin(ph.dt_max, ph.cfl * ph.dx / (double)speed);
  return fmax(dt, ph.dt_min);
}

// FP64 exp and sincos for the integrating factor, a few double ulp accurate, so the factor is
// still correctly rounded once to FP32. Method: Cody-Waite reduction by ln2 (resp. pi/2) with
// split constants (high parts carry 32-33 fractional bits so n*hi is exact), then the Taylor
// series of exp to r^12 on |r| <= ln2/2 and of sin / cos to r^15 / r^16 on |r| <= pi/4 (Horner).
// exp_f64 is accurate over the whole double range (the 2^n scale is applied in two halves so
// results up to the overflow threshold stay finite); sincos_f64 is accurate for |x| < 1e5 and
// stays in [-1, 1] beyond (the solver's parameter checks keep the integrating factor's phase
// below 1e4). No library fallback: its code would set the register budget of the step kernels.
__host__ __device__ __forceinline__ double exp_f64(double y) {
  if (fabs(y) < 0.0625) {
    // small arguments (the damping of all but the highest modes): Taylor to y^8, truncation
    // below 0.0625^9 / 9! = 4e-17, no reduction
    double p = 2.48015873015873e-05;
    p = fma(p, y, 0.0001984126984126984);
    p = fma(p, y, 0.001388888888888889);
    p = fma(p, y, 0.008333333333333333);
    p = fma(p, y, 0.041666666666666664);
    p = fma(p, y, 0.16666666666666666);
    p = fma(p, y, 0.5);
    p = fma(p, y, 1.0);
    return fma(p, y, 1.0);
  }
  if (!(y > -745.0)) return y == y ? 0.0 : y;
  if (!(y < 710.0)) return y * HUGE_VAL;  // +inf, or NaN
  const double n = rint(y * 1.4426950408889634);
  double r = fma(-n, 0.6931471806019545, y);
  r = fma(-n, -4.2009150726810846e-11, r);
  double p = 2.08767569878681e-09;
  p = fma(p, r, 2.505210838544172e-08);
  p = fma(p, r, 2.755731922398589e-07);
  p = fma(p, r, 2.7557319223985893e-06);
  p = fma(p, r, 2.48015873015873e-05);
  p = fma(p, r, 0.0001984126984126984);
  p = fma(p, r, 0.001388888888888889);
  p = fma(p, r, 0.008333333333333333);
  p = fma(p, r, 0.041666666666666664);
  p = fma(p, r, 0.16666666666666666);
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  const int ni = (int)n;
  return ldexp(ldexp(p, ni >> 1), ni - (ni >> 1));
}

__host__ __device__ __forceinline__ void sincos_f64(double x, double* s, double* c) {
  if (fabs(x) < 0.125) {
    // small phases (beta kx / k^2 dt / 2 for all but the lowest modes): sin to x^11 and cos to
    // x^12, truncation below 3e-18 relative, no reduction
    const double x2 = x * x;
    double ps = -2.505210838544172e-08;
    ps = fma(ps, x2, 2.7557319223985893e-06);
    ps = fma(ps, x2, -0.0001984126984126984);
    ps = fma(ps, x2, 0.008333333333333333);
    ps = fma(ps, x2, -0.16666666666666666);
    *s = fma(ps * x2, x, x);
    double pc = 2.08767569878681e-09;
    pc = fma(pc, x2, -2.755731922398589e-07);
    pc = fma(pc, x2, 2.48015873015873e-05);
    pc = fma(pc, x2, -0.001388888888888889);
    pc = fma(pc, x2, 0.041666666666666664);
    pc = fma(pc, x2, -0.5);
    *c = fma(pc, x2, 1.0);
    return;
  }
  const double q = rint(x * 0.6366197723675814);
  double r = fma(-q, 1.570796326850541, x);
  r = fma(-q, -5.5644316763895155e-11, r);
  r = fma(-q, 2.0222662487959506e-21, r);
  const double r2 = r * r;
  double ps = -7.647163731819816e-13;
  ps = fma(ps, r2, 1.6059043836821613e-10);
  ps = fma(ps, r2, -2.505210838544172e-08);
  ps = fma(ps, r2, 2.7557319223985893e-06);
  ps = fma(ps, r2, -0.0001984126984126984);
  ps = fma(ps, r2, 0.008333333333333333);
  ps = fma(ps, r2, -0.16666666666666666);
  const double sr = fma(ps * r2, r, r);
  double pc = 4.779477332387385e-14;
  pc = fma(pc, r2, -1.1470745597729725e-11);
  pc = fma(pc, r2, 2.08767569878681e-09);
  pc = fma(pc, r2, -2.755731922398589e-07);
  pc = fma(pc, r2, 2.48015873015873e-05);
  pc = fma(pc, r2, -0.001388888888888889);
  pc = fma(pc, r2, 0.041666666666666664);
  pc = fma(pc, r2, -0.5);
  const double cr = fma(pc, r2, 1.0);
  switch (((long long)q) & 3) {
    case 0: *s = sr; *c = cr; break;
    case 1: *s = cr; *c = -sr; break;
    case 2: *s = -sr; *c = -cr; break;
    default: *s = -cr; *c = sr; break;
  }
}

// Integrating factor E(s) = exp(L s), L = -mu - nu (k^2)^p + i beta kx / k^2
// (reference/spectral2d.py:26, solver.wgsl expL:97-101), evaluated in FP64 once per mode and
// stage as E(dt/2); E(dt) = E(dt/2)^2 and E(-dt/2) = 1/E(dt/2).
struct IFactor {
  double mag, c, s;  // E(dt/2) = mag * (c + i s)
};
// 1/x from the hardware reciprocal estimate and two Newton steps (within 1 double ulp; used in
// place of a correctly rounded division, whose scaling and fix-up steps lengthen the per-mode chain)
__device__ __forceinline__ double rcp_f64(double x) {
  double y = __builtin_amdgcn_rcp(x);
  double e = fma(-x, y, 1.0);
  y = fma(y, e, y);
  e = fma(-x, y, 1.0);
  return fma(y, e, y);
}

__device__ __forceinline__ IFactor ifactor_half(double kx, double k2, const Phys& ph, double dt) {
  // (k^2)^p for the integer order p >= 1 (the configuration checks reject other orders)
  double k2p = k2;
  for (int i = 1; i < ph.p_int; ++i) k2p *= k2;
  const double re = -ph.mu - ph.nu * k2p;
  IFactor e;
  e.mag = exp_f64(re * 0.5 * dt);
  if (ph.beta != 0.0) {
    sincos_f64(ph.beta * kx * rcp_f64(k2) * 0.5 * dt, &e.s, &e.c);
  } else {
    e.s = 0.0;
    e.c = 1.0;
  }
  return e;
}

// RK coefficients [cA, a, cB, b] of the three IF-SSP-RK3 stages (src/gpu/solver.js:11-15):
//   wout = cA E(a dt) w0 + cB E(b dt) (wcur + dt rhs)
// evaluated in FP64 from the FP64 E(dt/2) and rounded once: rounding E to FP32 instead would bias
// a pure decay by ~2e-8 per step, the same every step (2e-5 over 1000 steps).
__device__ __forceinline__ float2 rk_update(int stage, const IFactor& e, float2 w0, float2 wcur,
                                            float2 rhs, double dt) {
  const double ax = (double)wcur.x + dt * (double)rhs.x;
  const double ay = (double)wcur.y + dt * (double)rhs.y;
  const double hr = e.mag * e.c, hi = e.mag * e.s;  // E(dt/2)
  if (stage == 0) {
    const double er = hr * hr - hi * hi, ei = 2.0 * hr * hi;  // E(dt)
    return make_float2((float)(er * ax - ei * ay), (float)(er * ay + ei * ax));
  }
  if (stage == 1) {
    const double im = rcp_f64(e.mag);
    const double ir = e.c * im, ii = -e.s * im;  // E(-dt/2)
    const double x = 0.75 * (hr * w0.x - hi * w0.y) + 0.25 * (ir * ax - ii * ay);
    const double y = 0.75 * (hr * w0.y + hi * w0.x) + 0.25 * (ir * ay + ii * ax);
    return make_float2((float)x, (float)y);
  }
  const double er = hr * hr - hi * hi, ei = 2.0 * hr * hi;
  const double x = (1.0 / 3.0) * (er * w0.x - ei * w0.y) + (2.0 / 3.0) * (hr * ax - hi * ay);
  const double y = (1.0 / 3.0) * (er * w0.y + ei * w0.x) + (2.0 / 3.0) * (hr * ay + hi * ax);
  return make_float2((float)x, (float)y);
}

// periodic separation in unit coordinates (common.wgsl:42-48)
__device__ __forceinline__ float sep(float a, float b) {
  const float d = a - b + 0.5f;
  return d - floorf(d) - 0.5f;
}

// curl of the Gaussian body forces at unit position (x, y) (src/shaders/force.wgsl:11-26)
__device__ __forceinline__ float pointer_curl(float x, float y, const Splats& sp) {
  float total = 0.f;
  for (int i = 0; i < sp.count; ++i) {
    const Splat s = sp.s[i];
    const float rx = sep(x, s.x), ry = sep(y, s.y);
    const float r2 = rx * rx + ry * ry;
    const float sig2 = fmaxf(s.radius * s.radius, 1e-8f);
    if (r2 < 9.f * sig2) {
      const float g = expf(-r2 / sig2);
      total = total + 2.f * s.amp * g * (s.dx * ry - s.dy * rx) / sig2;
    }
  }
  return total;
}

}  // namespace s2t
