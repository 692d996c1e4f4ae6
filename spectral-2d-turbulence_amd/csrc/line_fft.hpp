// Workgroup line FFT for gfx950: register radix-R butterflies, Stockham exchanges through LDS.
//
// Replaces the in-place radix-4 DIT of src/shaders/fft.wgsl:42-79 (one workgroup per line,
// digit-reversal scatter on load, one LDS round trip per radix-4 stage, FP32 twiddles from
// cos/sin each butterfly). Here a line of N complex values is held as N/T values per thread in
// registers; each pass does one radix-R (R <= 16) DFT per thread in registers and the passes
// exchange through LDS, so a 4096-point line needs 3 passes and 2 LDS exchanges instead of 6.
// The Stockham (autosort) order needs no digit reversal: thread t holds x[t + q*T] on entry and
// X[t + q*T] on exit, so loads and stores of a contiguous line are coalesced.
//
// Conventions match the reference: forward exp(-2 pi i jk/N) unscaled; the inverse transform is
// exp(+2 pi i jk/N) and is left unscaled here (callers fold the 1/N per pass of
// src/gpu/fft.js:59 into an exact power-of-two scale elsewhere).
//
// Twiddles between passes come from a table computed in FP64 on the host and rounded once to
// FP32 (table layout in twiddle_table(); inverse uses the conjugate).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "bufio.hpp"

namespace s2t {

// ---------------------------------------------------------------------------------------------
// complex helpers (float2 = interleaved re, im, the layout of the reference's vec2<f32>)
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
// multiply by i
__device__ __forceinline__ float2 cmuli(float2 a) { return make_float2(-a.y, a.x); }

// multiply by (DIR * i): DIR = -1 forward (-i), +1 inverse (+i)
template <int DIR>
__device__ __forceinline__ float2 mul_dir_i(float2 a) {
  return DIR < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

// multiply by (c, DIR*s)
template <int DIR>
__device__ __forceinline__ float2 mul_cs(float2 a, float c, float s) {
  const float sd = DIR < 0 ? -s : s;
  return make_float2(a.x * c - a.y * sd, a.x * sd + a.y * c);
}

// cos/sin of k*pi/8 rounded once from their FP64 values
constexpr float C8_1 = 0.92387953251128675613f;  // cos(pi/8)
constexpr float S8_1 = 0.38268343236508977173f;  // sin(pi/8)
constexpr float RSQRT2 = 0.70710678118654752440f;

// ---------------------------------------------------------------------------------------------
// in-register DFTs, natural-order in and out
// ---------------------------------------------------------------------------------------------
template <int DIR>
__device__ __forceinline__ void dft2(float2& a, float2& b) {
  const float2 t = csub(a, b);
  a = cadd(a, b);
  b = t;
}

template <int DIR>
__device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
  const float2 s02 = cadd(a0, a2), d02 = csub(a0, a2);
  const float2 s13 = cadd(a1, a3), d13 = mul_dir_i<DIR>(csub(a1, a3));
  a0 = cadd(s02, s13);
  a1 = cadd(d02, d13);
  a2 = csub(s02, s13);
  a3 = csub(d02, d13);
}

// W8^1 = (1, DIR)/sqrt2 and W8^3 = (-1, DIR)/sqrt2
template <int DIR>
__device__ __forceinline__ float2 mul_w8_1(float2 a) {
  return DIR < 0 ? make_float2((a.x + a.y) * RSQRT2, (a.y - a.x) * RSQRT2)
                 : make_float2((a.x - a.y) * RSQRT2, (a.y + a.x) * RSQRT2);
}
template <int DIR>
__device__ __forceinline__ float2 mul_w8_3(float2 a) {
  return DIR < 0 ? make_float2((a.y - a.x) * RSQRT2, -(a.x + a.y) * RSQRT2)
                 : make_float2(-(a.x + a.y) * RSQRT2, (a.x - a.y) * RSQRT2);
}

template <int DIR>
__device__ __forceinline__ void dft8(float2* v) {
  float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  float2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4<DIR>(e0, e1, e2, e3);
  dft4<DIR>(o0, o1, o2, o3);
  o1 = mul_w8_1<DIR>(o1);
  o2 = mul_dir_i<DIR>(o2);
  o3 = mul_w8_3<DIR>(o3);
  v[0] = cadd(e0, o0); v[4] = csub(e0, o0);
  v[1] = cadd(e1, o1); v[5] = csub(e1, o1);
  v[2] = cadd(e2, o2); v[6] = csub(e2, o2);
  v[3] = cadd(e3, o3); v[7] = csub(e3, o3);
}

// 16 = 4 x 4: DFT4 over n1 for each n2, twiddle W16^(n2 k1), DFT4 over n2, transpose.
template <int DIR>
__device__ __forceinline__ void dft16(float2* v) {
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) dft4<DIR>(v[n2], v[4 + n2], v[8 + n2], v[12 + n2]);
  // slot n2 + 4 k1 holds Y[n2][k1]
  v[1 + 4 * 1] = mul_cs<DIR>(v[1 + 4 * 1], C8_1, S8_1);     // W^1
  v[1 + 4 * 2] = mul_w8_1<DIR>(v[1 + 4 * 2]);               // W^2
  v[1 + 4 * 3] = mul_cs<DIR>(v[1 + 4 * 3], S8_1, C8_1);     // W^3
  v[2 + 4 * 1] = mul_w8_1<DIR>(v[2 + 4 * 1]);               // W^2
  v[2 + 4 * 2] = mul_dir_i<DIR>(v[2 + 4 * 2]);              // W^4
  v[2 + 4 * 3] = mul_w8_3<DIR>(v[2 + 4 * 3]);               // W^6
  v[3 + 4 * 1] = mul_cs<DIR>(v[3 + 4 * 1], S8_1, C8_1);     // W^3
  v[3 + 4 * 2] = mul_w8_3<DIR>(v[3 + 4 * 2]);               // W^6
  v[3 + 4 * 3] = mul_cs<DIR>(v[3 + 4 * 3], -C8_1, -S8_1);   // W^9
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) dft4<DIR>(v[4 * k1], v[4 * k1 + 1], v[4 * k1 + 2], v[4 * k1 + 3]);
  // slot 4 k1 + k2 holds X[k1 + 4 k2]: transpose the 4 x 4 block
  float2 t[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) t[i] = v[i];
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1)
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = t[4 * k1 + k2];
}

// Order in which to load a thread's R values for the first DFT: the order its first butterflies
// consume them (radix 16: the four DFT4s over n2 take v[n2], v[4 + n2], v[8 + n2], v[12 + n2];
// radix 8: the even, then the odd DFT4). Memory waits are in issue order, so the first butterfly
// can start once its own inputs have arrived.
template <int R>
constexpr int dft_load_order(int i) {
  return R == 16 ? (i % 4) * 4 + i / 4 : R == 8 ? (i % 4) * 2 + i / 4 : i;
}

template <int R, int DIR>
__device__ __forceinline__ void dft(float2* v) {
  if constexpr (R == 2) dft2<DIR>(v[0], v[1]);
  else if constexpr (R == 4) dft4<DIR>(v[0], v[1], v[2], v[3]);
  else if constexpr (R == 8) dft8<DIR>(v);
  else if constexpr (R == 16) dft16<DIR>(v);
  else static_assert(R == 2, "unsupported radix");
}

// ---------------------------------------------------------------------------------------------
// radix plans. N must be 4^a or 2*4^a (src/gpu/fft.js:20-31); the first radix is the largest,
// so a thread holds PTS = R0 values and T = N / R0 threads cooperate on one line. 128 and 256
// take three passes with R0 = 8 rather than two with R0 = 16: their kernels are bound by one
// thread's chain of work, and twice the threads per line halve it (a 128^2 step 79 -> 57 us).
// ---------------------------------------------------------------------------------------------
template <int N> struct Plan;
template <> struct Plan<16>   { static constexpr int R0 = 4,  R1 = 4,  R2 = 1; };
template <> struct Plan<32>   { static constexpr int R0 = 8,  R1 = 4,  R2 = 1; };
template <> struct Plan<64>   { static constexpr int R0 = 8,  R1 = 8,  R2 = 1; };
template <> struct Plan<128>  { static constexpr int R0 = 8,  R1 = 4,  R2 = 4; };
template <> struct Plan<256>  { static constexpr int R0 = 8,  R1 = 8,  R2 = 4; };
template <> struct Plan<512>  { static constexpr int R0 = 8,  R1 = 8,  R2 = 8; };
template <> struct Plan<1024> { static constexpr int R0 = 16, R1 = 16, R2 = 4; };
template <> struct Plan<2048> { static constexpr int R0 = 16, R1 = 16, R2 = 8; };
template <> struct Plan<4096> { static constexpr int R0 = 16, R1 = 16, R2 = 16; };

// Twiddle bases of a radix-R pass. For R = 8 and 16 a thread loads W^{k}, W^{2k}, W^{4k}, W^{8k}
// (and W^{16k}) and forms the other W^{rk} with one or two complex products (r = a + b or
// r = a - b of two bases; a few FP32 ulp): 5 loads per radix-16 pass instead of 15. The twiddle
// loads are vector memory instructions, and the column kernel is bound by its memory-instruction
// throughput.
template <int R>
struct TwBase {
  static constexpr int NB = R == 16 ? 5 : (R == 8 ? 4 : R - 1);
  static constexpr int exponent(int j) { return R >= 8 ? 1 << j : j + 1; }
};

template <int N>
struct Line {
  static constexpr int R0 = Plan<N>::R0, R1 = Plan<N>::R1, R2 = Plan<N>::R2;
  static constexpr int PTS = R0;       // values per thread
  static constexpr int T = N / R0;     // threads per line
  static constexpr int NPASS = R2 > 1 ? 3 : 2;
  static constexpr int LDS = N + N / 32;                     // padded LDS float2 per line
  // Twiddle tables: pass p (radix R, span NS) stores W_{NS R}^{e k} for k < NS and the exponents
  // e of TwBase<R> (R >= 8: the powers of two up to R; else every r = 1..R-1).
  static constexpr int TW1 = 0;                                   // pass-1 table offset
  static constexpr int TW2 = TwBase<R1>::NB * R0;                 // pass-2 table offset
  static constexpr int TWSIZE = TW2 + (R2 > 1 ? TwBase<R2>::NB * R0 * R1 : 0);
};

// LDS slot of line element i: one padding slot per 32 elements. The first exchange writes with
// stride R0 = 16 and the later passes access runs of consecutive elements; with this padding a
// 16- or 32-lane group of 64-bit accesses mostly lands on distinct bank pairs (32 or 64 banks),
// and the slot stays affine in the element index, so a thread's accesses share one address
// register and differ by immediate offsets.
__device__ __forceinline__ int swz(int i) { return i + (i >> 5); }

// Twiddle bases of the butterflies a thread computes in the two passes of a line transform
// (radix R1 span R0, radix R2 span R0 R1); they depend on the thread only, so a kernel can load
// them once (load_bases) and pass them to every line_fft.
template <int N>
struct Bases {
  using L = Line<N>;
  static constexpr int R2E = L::NPASS == 3 ? L::R2 : 2;
  static constexpr int NB1 = N / (L::R1 * L::T), NB2 = N / (R2E * L::T);
  static constexpr int NT1 = TwBase<L::R1>::NB, NT2 = L::NPASS == 3 ? TwBase<R2E>::NB : 0;
  static constexpr int SIZE = NB1 * NT1 + NB2 * NT2;
};

template <int N>
__device__ __forceinline__ void load_bases(float2 (&bs)[Bases<N>::SIZE], const float2* __restrict__ tw,
                                           int t) {
  using L = Line<N>;
  using B = Bases<N>;
#pragma unroll
  for (int i = 0; i < B::NB1; ++i)
#pragma unroll
    for (int j = 0; j < B::NT1; ++j) bs[i * B::NT1 + j] = tw[L::TW1 + j * L::R0 + (t + i * L::T) % L::R0];
#pragma unroll
  for (int i = 0; i < B::NB2; ++i)
#pragma unroll
    for (int j = 0; j < B::NT2; ++j)
      bs[B::NB1 * B::NT1 + i * B::NT2 + j] =
          tw[L::TW2 + j * L::R0 * L::R1 + (t + i * L::T) % (L::R0 * L::R1)];
}


// Twiddle bases of one Stockham pass (radix R, span NS) for this thread's NB butterflies.
template <int N, int R, int NS>
struct PassBases {
  static constexpr int NB = N / (R * Line<N>::T);  // butterflies per thread
  static constexpr int PER = TwBase<R>::NB;        // bases per butterfly
  static constexpr int SIZE = NB * PER;
};

// Load this thread's bases of one pass from its table, which starts `off` values into the
// buffer `tb` (pass table: base j of span offset k at j * NS + k). Buffer loads count in vmcnt
// only; a generic (flat) load also counts in lgkmcnt, so the wait for an exchange's LDS writes
// before its barrier would wait for the twiddle loads in flight as well.
template <int N, int R, int NS, int S>
__device__ __forceinline__ void load_pass_bases(float2 (&bs)[S], Buf tb, int off, int t) {
  using P = PassBases<N, R, NS>;
  static_assert(S == P::SIZE, "bases array of this pass");
#pragma unroll
  for (int i = 0; i < P::NB; ++i)
#pragma unroll
    for (int j = 0; j < P::PER; ++j)
      bs[i * P::PER + j] = ld2(tb, (uint32_t)((t + i * Line<N>::T) % NS) * 8u,
                               (uint32_t)(off + j * NS) * 8u);
}

// One butterfly of a middle/last Stockham pass (radix R, span NS): its R inputs u (read from
// LDS) twiddled with its bases `bi` (TwBase<R>::NB values, see load_pass_bases), then the DFT in
// registers.
template <int R, int DIR>
__device__ __forceinline__ void butterfly_twiddle_dft(float2 (&u)[R], const float2* bi) {
  float2 w[R];  // w[r] = W_{NS R}^{r k}, r = 1..R-1, k = b mod NS
  if constexpr (R >= 8) {
    w[1] = bi[0];
    w[2] = bi[1];
    w[4] = bi[2];
    w[3] = cmul(bi[1], bi[0]);
    w[5] = cmul(bi[2], bi[0]);
    w[6] = cmul(bi[2], bi[1]);
    w[7] = cmul(bi[3], cconj(bi[0]));
    if constexpr (R == 16) {
      w[8] = bi[3];
      w[9] = cmul(bi[3], bi[0]);
      w[10] = cmul(bi[3], bi[1]);
      w[11] = cmul(bi[3], w[3]);
      w[12] = cmul(bi[3], bi[2]);
      w[13] = cmul(bi[4], cconj(w[3]));
      w[14] = cmul(bi[4], cconj(bi[1]));
      w[15] = cmul(bi[4], cconj(bi[0]));
    }
  } else {
#pragma unroll
    for (int r = 1; r < R; ++r) w[r] = bi[r - 1];
  }
#pragma unroll
  for (int r = 1; r < R; ++r) u[r] = cmul(u[r], DIR > 0 ? cconj(w[r]) : w[r]);
  dft<R, DIR>(u);
}

// Middle/last Stockham pass, compute part: NB butterflies of radix R per thread, span NS, inputs
// read from LDS, twiddled with the bases `bs` (NB x TwBase<R>::NB, see load_pass_bases), DFT in
// registers. The results stay in u for pass_store.
template <int N, int R, int NS, int DIR>
__device__ __forceinline__ void pass_compute(float2 (&u)[N / (R * Line<N>::T)][R], const float2* lds,
                                             int t, const float2* bs) {
  constexpr int T = Line<N>::T;
  constexpr int NB = N / (R * T);
  constexpr int STRIDE = N / R;
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int b = t + i * T;
#pragma unroll
    for (int r = 0; r < R; ++r) u[i][r] = lds[swz(b + r * STRIDE)];
    const float2* bi = bs + i * TwBase<R>::NB;
    float2 w[R];  // w[r] = W_{NS R}^{r k}, r = 1..R-1, k = b mod NS
    if constexpr (R >= 8) {
      w[1] = bi[0];
      w[2] = bi[1];
      w[4] = bi[2];
      w[3] = cmul(bi[1], bi[0]);
      w[5] = cmul(bi[2], bi[0]);
      w[6] = cmul(bi[2], bi[1]);
      w[7] = cmul(bi[3], cconj(bi[0]));
      if constexpr (R == 16) {
        w[8] = bi[3];
        w[9] = cmul(bi[3], bi[0]);
        w[10] = cmul(bi[3], bi[1]);
        w[11] = cmul(bi[3], w[3]);
        w[12] = cmul(bi[3], bi[2]);
        w[13] = cmul(bi[4], cconj(w[3]));
        w[14] = cmul(bi[4], cconj(bi[1]));
        w[15] = cmul(bi[4], cconj(bi[0]));
      }
    } else {
#pragma unroll
      for (int r = 1; r < R; ++r) w[r] = bi[r - 1];
    }
#pragma unroll
    for (int r = 1; r < R; ++r) u[i][r] = cmul(u[i][r], DIR > 0 ? cconj(w[r]) : w[r]);
    dft<R, DIR>(u[i]);
  }
}

// Middle/last Stockham pass, store part: the last pass leaves X[t + q T] in v; a middle pass
// writes its outputs to LDS in Stockham order for the next pass (barrier, writes, barrier).
template <int N, int R, int NS, bool LAST>
__device__ __forceinline__ void pass_store(float2 (&u)[N / (R * Line<N>::T)][R],
                                           float2 (&v)[Line<N>::PTS], float2* lds, int t) {
  constexpr int T = Line<N>::T;
  constexpr int NB = N / (R * T);
  if constexpr (LAST) {
#pragma unroll
    for (int i = 0; i < NB; ++i)
#pragma unroll
      for (int r = 0; r < R; ++r) v[i + r * NB] = u[i][r];
  } else {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const int b = t + i * T;
      const int k = b % NS;
      const int base = (b / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) lds[swz(base + r * NS)] = u[i][r];
    }
    __syncthreads();
  }
}

// The passes of a line transform after the first DFT. Each pass's twiddle bases come from
// `fetch(pass, array)` (pass = std::integral_constant 1 or 2; a table load or a copy of
// registers), issued before the exchange that precedes the pass so that a table load's latency
// overlaps the barrier and the LDS round trip.
template <int N, int DIR, class Fetch>
__device__ __forceinline__ void line_passes(float2 (&v)[Line<N>::PTS], float2* lds, int t,
                                            Fetch&& fetch) {
  using L = Line<N>;
  using P1 = PassBases<N, L::R1, L::R0>;
  float2 b1[P1::SIZE];
  fetch(std::integral_constant<int, 1>{}, b1);
  __syncthreads();
#pragma unroll
  for (int r = 0; r < L::R0; ++r) lds[swz(t * L::R0 + r)] = v[r];
  __syncthreads();
  float2 u1[P1::NB][L::R1];
  pass_compute<N, L::R1, L::R0, DIR>(u1, lds, t, b1);
  if constexpr (L::NPASS == 2) {
    pass_store<N, L::R1, L::R0, true>(u1, v, lds, t);
  } else {
    using P2 = PassBases<N, L::R2, L::R0 * L::R1>;
    float2 b2[P2::SIZE];
    fetch(std::integral_constant<int, 2>{}, b2);
    pass_store<N, L::R1, L::R0, false>(u1, v, lds, t);
    float2 u2[P2::NB][L::R2];
    pass_compute<N, L::R2, L::R0 * L::R1, DIR>(u2, lds, t, b2);
    pass_store<N, L::R2, L::R0 * L::R1, true>(u2, v, lds, t);
  }
}

// Full line transform. On entry v[q] = x[t + q*T]; on exit v[q] = X[t + q*T].
// Every thread of the workgroup must call this the same number of times (it contains barriers);
// threads whose line is unused still call it on whatever data they hold.
template <int N, int DIR>
__device__ __forceinline__ void line_fft(float2 (&v)[Line<N>::PTS], float2* lds, int t,
                                         const float2* __restrict__ tw) {
  using L = Line<N>;
  // Opaque copy of the table pointer: keeps the compiler from hoisting every pass's twiddle loads
  // out of the callers' field loops, which would pin ~60 VGPRs for the whole kernel.
  const float2* twl = tw;
  asm volatile("" : "+s"(twl));
  const Buf tb = make_buf(twl, (uint32_t)(L::TWSIZE * 8));
  dft<L::R0, DIR>(v);
  line_passes<N, DIR>(v, lds, t, [&](auto pass, auto& bs) {
    if constexpr (decltype(pass)::value == 1)
      load_pass_bases<N, L::R1, L::R0>(bs, tb, L::TW1, t);
    else
      load_pass_bases<N, L::R2, L::R0 * L::R1>(bs, tb, L::TW2, t);
  });
}

// The same transform, with every Stockham exchange made in two rounds through a buffer of
// Line<N>::LDS floats, half the LDS of the full exchange: the real parts, then the imaginary
// parts (four barriers per exchange instead of two). The arithmetic is line_fft's, so the result
// is the same bit for bit. For a workgroup that must hold other data in LDS across a transform.
template <int N, int DIR>
__device__ __forceinline__ void line_fft_half(float2 (&v)[Line<N>::PTS], float* ldsf, int t,
                                              const float2* __restrict__ tw) {
  using L = Line<N>;
  constexpr int T = L::T;
  static_assert(L::NPASS == 3, "three-pass plans only");
  const float2* twl = tw;
  asm volatile("" : "+s"(twl));
  const Buf tb = make_buf(twl, (uint32_t)(L::TWSIZE * 8));
  dft<L::R0, DIR>(v);
  using P1 = PassBases<N, L::R1, L::R0>;
  using P2 = PassBases<N, L::R2, L::R0 * L::R1>;
  float2 b1[P1::SIZE];
  load_pass_bases<N, L::R1, L::R0>(b1, tb, L::TW1, t);
  // exchange 1: thread t's R0 outputs at t R0 + r; pass 1 reads b + r N / R1
  float2 u1[P1::NB][L::R1];
#pragma unroll
  for (int half = 0; half < 2; ++half) {
    __syncthreads();
#pragma unroll
    for (int r = 0; r < L::R0; ++r) ldsf[swz(t * L::R0 + r)] = half ? v[r].y : v[r].x;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < P1::NB; ++i)
#pragma unroll
      for (int r = 0; r < L::R1; ++r) {
        const float x = ldsf[swz(t + i * T + r * (N / L::R1))];
        if (half) u1[i][r].y = x;
        else u1[i][r].x = x;
      }
  }
#pragma unroll
  for (int i = 0; i < P1::NB; ++i) butterfly_twiddle_dft<L::R1, DIR>(u1[i], b1 + i * P1::PER);
  float2 b2[P2::SIZE];
  load_pass_bases<N, L::R2, L::R0 * L::R1>(b2, tb, L::TW2, t);
  // exchange 2: pass 1's outputs in Stockham order (span NS = R0); pass 2 reads b + r N / R2
  constexpr int NS = L::R0;
  float2 u2[P2::NB][L::R2];
#pragma unroll
  for (int half = 0; half < 2; ++half) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < P1::NB; ++i) {
      const int b = t + i * T;
      const int base = (b / NS) * NS * L::R1 + b % NS;
#pragma unroll
      for (int r = 0; r < L::R1; ++r) ldsf[swz(base + r * NS)] = half ? u1[i][r].y : u1[i][r].x;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < P2::NB; ++i)
#pragma unroll
      for (int r = 0; r < L::R2; ++r) {
        const float x = ldsf[swz(t + i * T + r * (N / L::R2))];
        if (half) u2[i][r].y = x;
        else u2[i][r].x = x;
      }
  }
#pragma unroll
  for (int i = 0; i < P2::NB; ++i) butterfly_twiddle_dft<L::R2, DIR>(u2[i], b2 + i * P2::PER);
#pragma unroll
  for (int i = 0; i < P2::NB; ++i)
#pragma unroll
    for (int r = 0; r < L::R2; ++r) v[i + r * P2::NB] = u2[i][r];
}

// The twiddles W^{r k}, r = 1..15, of a radix-16 butterfly from its bases (the products of
// butterfly_twiddle_dft): w[r - 1].
__device__ __forceinline__ void twiddles16(float2 (&w)[15], const float2* bi) {
  w[0] = bi[0];
  w[1] = bi[1];
  w[3] = bi[2];
  w[2] = cmul(bi[1], bi[0]);
  w[4] = cmul(bi[2], bi[0]);
  w[5] = cmul(bi[2], bi[1]);
  w[6] = cmul(bi[3], cconj(bi[0]));
  w[7] = bi[3];
  w[8] = cmul(bi[3], bi[0]);
  w[9] = cmul(bi[3], bi[1]);
  w[10] = cmul(bi[3], w[2]);
  w[11] = cmul(bi[3], bi[2]);
  w[12] = cmul(bi[4], cconj(w[2]));
  w[13] = cmul(bi[4], cconj(bi[1]));
  w[14] = cmul(bi[4], cconj(bi[0]));
}

// The transform for three-pass plans whose first two passes are radix 16 with one pass-1
// butterfly per thread, with the pass-1 twiddles of this thread (twiddles16 of its pass-1 bases)
// and the pass-2 bases already in registers: a kernel that runs several transforms forms the
// pass-1 twiddles once rather than per transform. The arithmetic is line_fft's, so the result is
// the same bit for bit.
template <int N>
using Pass2Bases = PassBases<N, Line<N>::R2, Line<N>::R0 * Line<N>::R1>;
template <int N, int DIR>
__device__ __forceinline__ void line_fft_w1(float2 (&v)[Line<N>::PTS], float2* lds, int t,
                                            const float2 (&w1)[15],
                                            const float2 (&b2)[Pass2Bases<N>::SIZE]) {
  using L = Line<N>;
  using P2 = Pass2Bases<N>;
  static_assert(L::R0 == 16 && L::R1 == 16 && L::NPASS == 3 && L::T * 16 == N,
                "radix-16 first passes, one pass-1 butterfly per thread");
  dft<16, DIR>(v);
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 16; ++r) lds[swz(t * 16 + r)] = v[r];
  __syncthreads();
  float2 u1[1][16];
#pragma unroll
  for (int r = 0; r < 16; ++r) u1[0][r] = lds[swz(t + r * (N / 16))];
#pragma unroll
  for (int r = 1; r < 16; ++r) {
    // opaque copies: the products stay here instead of being hoisted out of the caller's loops
    float x = w1[r - 1].x, y = w1[r - 1].y;
    asm volatile("" : "+v"(x), "+v"(y));
    const float2 w = make_float2(x, y);
    u1[0][r] = cmul(u1[0][r], DIR > 0 ? cconj(w) : w);
  }
  dft<16, DIR>(u1[0]);
  float2 b[P2::SIZE];
#pragma unroll
  for (int i = 0; i < P2::SIZE; ++i) {
    float x = b2[i].x, y = b2[i].y;
    asm volatile("" : "+v"(x), "+v"(y));
    b[i] = make_float2(x, y);
  }
  pass_store<N, 16, 16, false>(u1, v, lds, t);
  float2 u2[P2::NB][L::R2];
  pass_compute<N, L::R2, 256, DIR>(u2, lds, t, b);
  pass_store<N, L::R2, 256, true>(u2, v, lds, t);
}

// The same transform with the twiddle bases already in registers (load_bases).
template <int N, int DIR>
__device__ __forceinline__ void line_fft(float2 (&v)[Line<N>::PTS], float2* lds, int t,
                                         const float2 (&bases)[Bases<N>::SIZE]) {
  using L = Line<N>;
  using B = Bases<N>;
  dft<L::R0, DIR>(v);
  // opaque copies of the bases: the twiddles derived from them stay local to this transform
  // instead of being hoisted (all 30 of them) to the caller's scope
  line_passes<N, DIR>(v, lds, t, [&](auto pass, auto& bs) {
    constexpr int n = sizeof(bs) / sizeof(float2);
    constexpr int off = decltype(pass)::value == 1 ? 0 : B::NB1 * B::NT1;
#pragma unroll
    for (int i = 0; i < n; ++i) {
      float x = bases[off + i].x, y = bases[off + i].y;
      asm volatile("" : "+v"(x), "+v"(y));
      bs[i] = make_float2(x, y);
    }
  });
}

}  // namespace s2t
