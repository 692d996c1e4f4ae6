// Fused step kernels for gfx950, specialised on the grid size N.
//
// One right-hand-side evaluation plus its RK stage is two launches (the reference spends ten
// dispatches on it, src/gpu/solver.js:496-517):
//
//   k_rows  (row kernel)   for a pair of rows y0, y0+1: inverse row FFTs of the packed half
//                          spectra -> physical omega, u, v; products u omega, v omega
//                          (nl_pack, solver.wgsl:42-54); forward row FFTs; Hermitian unpack along
//                          x into P' = FFT_x(u omega), Q' = FFT_x(v omega) for kx = 0..K. Stage 0
//                          also reduces max |u|+|v| for the CFL step (reduce.wgsl) and, when
//                          asked, writes the dye velocity and transforms the pointer forcing
//                          (force.wgsl) alongside.
//   k_cols  (column kernel) for C adjacent columns kx: forward column FFTs of P', Q' ->
//                          N_hat = i kx P_hat + i ky Q_hat, rhs = (f_hat - N_hat) masked +
//                          stochastic forcing (assemble, solver.wgsl:56-95); the IF-RK3 update
//                          (rk3, solver.wgsl:103-122) written to the state; then the next
//                          stage's velocity pack (velocity_pack, solver.wgsl:21-40) and inverse
//                          column FFTs of omega_hat, u_hat, v_hat -> omega', u', v'.
//
// Both kernels stage each line in LDS only for the Stockham exchanges of line_fft. The mixed
// fields between them are 4-row tiles (Geo::tile): the row kernel reads and writes a row pair as
// 16-byte elements at a stride of 32 bytes, the two pairs of a tile running back to back on one
// XCD; a column reads and writes one 32-byte run per tile of 4 rows, and neighbouring column
// groups on one XCD share the cache lines through L2.
#include <hip/hip_runtime.h>

#include <cmath>
#include <stdexcept>
#include <string>

#include "bufio.hpp"
#include "core.hpp"
#include "physics.hpp"
#include "solver.hpp"

namespace s2t {

enum RowFlags { ROW_SPEED = 1, ROW_VEL = 2, ROW_FORCE = 4 };

// Buffers of instance i start at base + i * stride (blockIdx.y = i); `inst` holds the
// per-instance parameters and scalars.
struct RowArgs {
  const float2* __restrict__ mu;  // omega', u', v' tiles (F = 3)
  float2* __restrict__ mz;        // P', Q' tiles (F = 2)
  float2* __restrict__ mf;        // F' tiles (F = 1)
  float2* __restrict__ vel;       // physical (u, v), row-major N x N
  const float2* __restrict__ tw;
  Inst* inst;
  size_t s_mu, s_mz, s_mf, s_vel;  // per-instance strides (elements)
  Splats splats;
  int stage;
  int flags;
};

struct ColArgs {
  const float2* __restrict__ mz;   // P', Q' tiles
  const float2* __restrict__ mf;   // F' tiles (stage 0 with pointer forcing)
  float2* __restrict__ mu;         // omega', u', v' tiles (out)
  const float2* __restrict__ w0;   // state at the start of the step
  const float2* __restrict__ wcur; // state entering this stage
  float2* __restrict__ wout;       // state leaving this stage
  float2* __restrict__ fk;         // pointer-force spectrum (compact layout)
  float2* __restrict__ rhs_out;    // compact rhs (COLS_RHS)
  const float2* __restrict__ tw;
  Inst* inst;
  size_t s_mz, s_mf, s_mu, s_state;  // per-instance strides (elements)
  int stage;
  int force;
  int sd;  // the products arrive as S', D' (sum / difference form: the stage-0 row kernels)
  uint32_t step_override;
  double dt_override;
};

template <int N>
struct RowCfg {
  static constexpr int G = Line<N>::T >= 64 ? 1 : 64 / Line<N>::T;  // row blocks per workgroup
  // row pairs per row block: one, so a 4096^2 stage is 2048 workgroups over the 768 resident
  // slots (3 per CU), 2.67 rounds; two pairs per block gave 1.33 rounds and left two thirds of
  // the CUs idle in the second (measured: 125 -> 115 us per launch)
  static constexpr int PAIRS = 1;
  static constexpr int NB = (N / (2 * PAIRS) + G - 1) / G;  // workgroups per instance
};
template <int N>
struct ColCfg {
  // default columns per workgroup, the fastest single instance measured per line length: 2 at
  // 256 threads (4096: 512-thread workgroups, two per CU at 64 KB of LDS each), 1 at 64..128
  // (512, 1024, 2048), 4 at 16..32, else one 64-lane wave
  static constexpr int C = Line<N>::T >= 256  ? 2
                           : Line<N>::T >= 64  ? 1
                           : Line<N>::T >= 16  ? 4
                                               : 64 / Line<N>::T;
  // lanes of a wave take columns in runs of RUN consecutive lanes (one column per 16- or 32-lane
  // LDS bank group for the long lines); the column buffers of the short lines are staggered
  static constexpr int RUN = Line<N>::T < 32 ? Line<N>::T : 32;
  static constexpr int LSTRIDE =
      Line<N>::LDS + (N >= 512 ? 0 : N == 256 ? 2 : N >= 64 ? 8 : 4);
  // column and line thread of a workgroup thread, for a workgroup of CW columns
  template <int CW>
  static __device__ __forceinline__ int col(int tid) { return (tid / RUN) % CW; }
  template <int CW>
  static __device__ __forceinline__ int row(int tid) { return (tid / (RUN * CW)) * RUN + tid % RUN; }
};

// Column-group widths (columns per workgroup) compiled for the step's column kernel, ColCfg's
// default first; s2t_autotune picks among them per solver. Each column's arithmetic is the same
// source at every width, and the compiled instances give bit-identical results (tested for every
// width); the widths differ in workgroup size and count, in how many columns share a workgroup's
// loads of a 128-byte line, and in LDS per workgroup (at most 68 KB, two workgroups per CU).
// 4096 keeps its default only: a one-column instance fused two of the W8 products into the
// following additions differently (results equal to FP32 rounding, not bit for bit), and it is
// the slower width there.
template <int N>
struct ColWidths {
  static constexpr int n = 1;
  static constexpr int w[4] = {ColCfg<N>::C, 0, 0, 0};
};
// sizes with the field-parallel column kernel (k_cols_fp), which is then their default, and its
// number of thread groups per column. One instance steps in 54 / 100 us at 512^2 / 1024^2 with
// three groups against 64 / 122 us with the best column-group width (two groups: k_cols 12.1 /
// 22.9 us per launch against 11.5 / 19.4 with three). At 2048^2 two groups take 30.9 us per
// launch against 33.3 us for column groups of one column (three groups: 38.2 us). At 4096^2 the
// column groups of two columns are the fastest (104 us; field-parallel 128 us with two groups,
// 140 us with three): the field-parallel form runs one column per workgroup, 2 columns per CU at
// a time against 4.
template <int N>
struct ColFp {
  static constexpr int groups = N == 512 || N == 1024 ? 3 : N == 2048 ? 2 : 0;
  static constexpr bool on = groups > 0;
  static constexpr int width = -groups;  // its code among the column-group widths
};
template <> struct ColWidths<2048> { static constexpr int n = 3; static constexpr int w[4] = {1, 2, 4, 0}; };
template <> struct ColWidths<1024> { static constexpr int n = 4; static constexpr int w[4] = {1, 2, 4, 8}; };
template <> struct ColWidths<512>  { static constexpr int n = 4; static constexpr int w[4] = {1, 2, 4, 8}; };
template <> struct ColWidths<256>  { static constexpr int n = 3; static constexpr int w[4] = {4, 8, 16, 0}; };
template <> struct ColWidths<128>  { static constexpr int n = 3; static constexpr int w[4] = {4, 8, 16, 0}; };

// every width: at most 1024 threads and 68 KB of LDS, and the thread mapping col<CW> / row<CW>
// covers each (column, line thread) pair once
template <int N>
constexpr bool widths_fit() {
  for (int i = 0; i < ColWidths<N>::n; ++i) {
    const int cw = ColWidths<N>::w[i], T = Line<N>::T, run = ColCfg<N>::RUN;
    if (cw < 1 || cw * T > 1024 || cw * ColCfg<N>::LSTRIDE * 8 > 68 * 1024) return false;
    if (T % run != 0) return false;
    // the thread of each (c, t), mapped back: a bijection on cw * T threads
    for (int c = 0; c < cw; ++c)
      for (int t = 0; t < T; ++t) {
        const int tid = (t / run) * run * cw + c * run + t % run;
        if (tid >= cw * T || (tid / run) % cw != c || (tid / (run * cw)) * run + tid % run != t)
          return false;
      }
  }
  return ColWidths<N>::w[0] == ColCfg<N>::C;
}
static_assert(widths_fit<128>() && widths_fit<256>() && widths_fit<512>() && widths_fit<1024>() &&
                  widths_fit<2048>() && widths_fit<4096>(),
              "column-group widths: at most 1024 threads and 68 KB of LDS, the default first");

// ---------------------------------------------------------------------------------------------
// row kernel
// ---------------------------------------------------------------------------------------------
// Blocks of a line (thread t, block q holds element i = t + q T) that lie out of band for every t:
// K < i < N - K for all of them, a run in the middle (q = 6..9 of 16 at N = 4096). A forward row
// transform's outputs there are never read by the Hermitian unpack; the column kernel's mode loop
// takes blocks in pairs and skips the pairs inside this run.
template <int N>
struct ColBand {
  static constexpr int T = Line<N>::T, PTS = Line<N>::PTS, K = Geo<N>::K;
  static constexpr bool out(int q) { return q * T > K && (q + 1) * T - 1 < N - K; }
  static constexpr int skip_first() {
    for (int q = 0; q + 1 < PTS; q += 2)
      if (out(q) && out(q + 1)) return q;
    return PTS;
  }
  static constexpr int skip_count() {
    int n = 0;
    for (int q = skip_first(); q + 1 < PTS && out(q) && out(q + 1); q += 2) n += 2;
    return n;
  }
  // the run of blocks out of band for every lane, block by block (PTS and 0 when there is none)
  static constexpr int out_first() {
    for (int q = 0; q < PTS; ++q)
      if (out(q)) return q;
    return PTS;
  }
  static constexpr int out_count() {
    int n = 0;
    for (int q = out_first(); q < PTS && out(q); ++q) ++n;
    return n;
  }
};
static_assert(ColBand<4096>::skip_first() == 6 && ColBand<4096>::skip_count() == 4,
              "4096: blocks 6..9 are out of band");

// Hermitian unpack along x of a forward row transform z held by the line (thread t holds
// z[t + q T]): the two real fields packed as re / im, for kx = 0..K, into the tiled planes
// (Geo::off).
// `voff` is the lane's byte offset of (row pair, kx = t). PAIR: the fields are rows 2m and 2m+1
// of one plane, stored together as one 16-byte element; otherwise they are one row of two planes
// (uniform byte offsets soff0 / soff1, which include the row's 8-byte slot in the pair).
// The two halves: row_unpack_stage puts the line in LDS (every thread of the workgroup calls it:
// it holds the barriers), row_unpack_emit stores the unpacked values of one line.
template <int N>
__device__ __forceinline__ void row_unpack_stage(float2 (&z)[Line<N>::PTS], float2* lds, int t) {
  using L = Line<N>;
  __syncthreads();
#pragma unroll
  for (int q = 0; q < L::PTS; ++q)
    if (!ColBand<N>::out(q)) lds[swz(t + q * L::T)] = z[q];  // the others are never read
  __syncthreads();
}

template <int N, bool PAIR>
__device__ __forceinline__ void row_unpack_emit(const float2* lds, int t, bool live, Buf out,
                                                uint32_t voff, uint32_t soff0, uint32_t soff1) {
  using L = Line<N>;
  using Ge = Geo<N>;
  if (!live) return;
#pragma unroll
  for (int q = 0; q < L::PTS; ++q) {
    const int k = t + q * L::T;
    if (q * L::T <= Ge::K) {
      // lanes past kx = K store through an out-of-range offset (dropped) rather than branching
      const float2 zk = lds[swz(k)];
      const float2 zm = lds[swz((N - k) & (N - 1))];
      float2 p, r;
      herm_unpack(zk, zm, p, r);
      const uint32_t so = (uint32_t)(q * L::T) * Ge::KXB, v = k <= Ge::K ? voff : BUF_OOB;
      if constexpr (PAIR) {
        st4(out, p, r, v, soff0 + so);
      } else {
        st2(out, p, v, soff0 + so);
        st2(out, r, v, soff1 + so);
      }
    }
  }
}

template <int N, bool PAIR>
__device__ __forceinline__ void row_unpack_store(float2 (&z)[Line<N>::PTS], float2* lds, int t,
                                                 bool live, Buf out, uint32_t voff,
                                                 uint32_t soff0, uint32_t soff1) {
  row_unpack_stage<N>(z, lds, t);
  row_unpack_emit<N, PAIR>(lds, t, live, out, voff, soff0, soff1);
}

// Inverse row transform input, Hermitian-extended to the full row: thread t, value q holds
// frequency k = t + q T. vlo / vhi are the lane's byte offsets of (row pair, kx = t) and
// (row pair, kx = T - t) in a tiled plane. VPAIR: x = a(y0) + i a(y0+1) for one real field
// a (plane offset soffa; one 16-byte load per value); otherwise x = a(y) + i b(y) for two real
// fields of one row (uniform offsets soffa / soffb include the row's 8-byte slot).
template <int N, bool VPAIR>
__device__ __forceinline__ void row_load(float2 (&x)[Line<N>::PTS], Buf mu, int t, bool live,
                                         uint32_t vlo, uint32_t vhi, uint32_t soffa,
                                         uint32_t soffb) {
  using L = Line<N>;
  using Ge = Geo<N>;
  constexpr int T = L::T, PTS = L::PTS, K = Ge::K;
  // Lanes whose k is out of band load through an out-of-range offset (zero, no memory access)
  // instead of branching around the load: a load under a per-lane branch makes the compiler wait
  // for it before the branch merges, which serialised the PTS loads of a line.
  auto pair = [&](uint32_t v, uint32_t so, float2& va, float2& vb) {
    if constexpr (VPAIR) {
      ld4(mu, v, soffa + so, va, vb);
    } else {
      va = ld2(mu, v, soffa + so);
      vb = ld2(mu, v, soffb + so);
    }
  };
#pragma unroll
  for (int i = 0; i < PTS; ++i) {
    const int q = dft_load_order<PTS>(i);  // the first DFT's order of use
    const int k = t + q * T;
    float2 v = make_float2(0.f, 0.f);
    if (2 * q < PTS) {
      if (q * T <= K) {
        float2 va, vb;
        pair(live && k <= K ? vlo : BUF_OOB, (uint32_t)(q * T) * Ge::KXB, va, vb);
        // k = 0: the real parts (the DC column of a real field)
        v = k == 0 ? make_float2(va.x, vb.x) : make_float2(va.x - vb.y, va.y + vb.x);
      }
    } else if ((q + 1) * T - 1 >= N - K) {
      // m = N - k = (T - t) + (PTS - 1 - q) T
      float2 va, vb;
      pair(live && k >= N - K ? vhi : BUF_OOB, (uint32_t)((PTS - 1 - q) * T) * Ge::KXB, va, vb);
      v = make_float2(va.x + vb.y, vb.x - va.y);
    }
    x[q] = v;
  }
}

// Row kernel variants. The stage-0 planes hold omega', s' = u' + v' and d' = u' - v' (the column
// kernel writes that form at stage 2 and in the prime); after stages 0 and 1 they hold omega', u'
// and v'. Stage 0 needs the CFL maximum of |u| + |v| = max(|s|, |d|), and it stores the products
// s omega, d omega as S', D' (the column kernel recovers N_hat = i kx P_hat + i ky Q_hat from them
// by linearity).
// ROWS_PAIRED_SD (stage 0) transforms the omega, s and d pairs of two rows at once, like
// ROWS_PAIRED; the maximum is max(max|s|, max|d|), each from one transform.
// ROWS_SINGLE (stage 0 with dye) pairs the transforms per row (B' = d'(y0) + i d'(y1) once per row
// pair, A' = omega' + i s' per row) so that s and d of a row are in registers together, for the dye
// velocity u = (s + d) / 2, v = (s - d) / 2; ROWS_SINGLE_FORCE adds the pointer forcing.
// ROWS_PAIRED (stages 1 and 2) transforms omega, u and v each for the two rows of a pair at once
// and packs the products across the pair (u0 omega0 + i u1 omega1, v0 omega0 + i v1 omega1), so
// every load is one 16-byte row-pair element. Every variant does five line FFTs per row pair.
// (The stage-0 maximum with u', v' planes needed the u pair held through the v transform, ~210
// VGPRs, or s and d each formed from both planes, a second read of u' and v'.)
enum RowKind { ROWS_SINGLE = 0, ROWS_SINGLE_FORCE = 1, ROWS_PAIRED = 2, ROWS_PAIRED_SD = 3 };

// LDS of the row kernel: G exchange lines; one real row of each pair (v(y0+1), or omega(y0+1) in
// ROWS_PAIRED), parked by each thread at its own indices while the other fields are transformed
// (holding it in registers instead would cost 16 VGPRs across two transforms); the wave maxima
template <int N, int G>
struct RowSmem {
  float2 lds[G * Line<N>::LDS];
  float vpark[G * N];
  float wmax[G * Line<N>::T / 64];
};

// One workgroup of the row kernel: workgroup bx (of the instance's grid of row blocks) of
// instance iy. Launched on its own by k_rows, or as an item of the grouped launch k_rows_group.
template <int N, int G, int KIND>
__device__ __forceinline__ void rows_body(const RowArgs& a, int bx, int iy, RowSmem<N, G>& sm) {
  using L = Line<N>;
  using Ge = Geo<N>;
  constexpr int T = L::T, PTS = L::PTS;
  constexpr uint32_t PLANE_B = (uint32_t)(Ge::PLANE * 8);
  float2* lds_all = sm.lds;
  float* vpark_all = sm.vpark;
  Inst* I = a.inst + iy;
  if (!I->st.active) return;  // uniform over the block
  const size_t inst = iy;
  constexpr bool FORCE = KIND == ROWS_SINGLE_FORCE;
  constexpr bool do_force = FORCE;
  const bool do_vel = a.flags & ROW_VEL;
  const Buf bu = make_buf(a.mu + inst * a.s_mu, 3 * PLANE_B);
  const Buf bz = make_buf(a.mz + inst * a.s_mz, 2 * PLANE_B);
  const Buf bf = make_buf(do_force ? a.mf + inst * a.s_mf : a.mf, do_force ? PLANE_B : 0u);
  const Buf bv = make_buf(do_vel ? a.vel + inst * a.s_vel : a.vel,
                          do_vel ? (uint32_t)((size_t)N * N * 8) : 0u);
  const int g = threadIdx.x / T, t = threadIdx.x - g * T;
  constexpr int PAIRS = RowCfg<N>::PAIRS;
  // XCD-aware mapping: consecutive workgroup ids go round-robin over the 8 XCDs; each XCD gets a
  // contiguous range of row blocks, so the row pairs of one tile run back to back on one XCD
  constexpr int NB = (N / (2 * PAIRS) + G - 1) / G, NB8 = (NB + 7) / 8 * 8;
  const int lb = (bx % 8) * (NB8 / 8) + bx / 8;
  if (lb >= NB) return;  // uniform over the block
  const int yb = lb * G + g;
  const bool live = yb < N / (2 * PAIRS);
  float2* lds = lds_all + g * L::LDS;
  float* vpark = vpark_all + g * N;
  // Twiddle bases of this thread, loaded once for all transforms of the kernel: only where the
  // LDS limit (three workgroups per CU at N = 4096) leaves the registers free, in the paired
  // variants (164 VGPRs at stage 0, held to the 168 of three waves per SIMD by the launch bound,
  // without spills); smaller N would lose occupancy. Stages 1-2 also form the 15 first-exchange
  // twiddles once for their five transforms (158 VGPRs; the issue-bound kernel saves 44 VALU per
  // transform). At 2048, where the kernel waits on its chain rather than on issue, the same took
  // 150 VGPRs and no time off.
  constexpr bool TWFULL = KIND == ROWS_PAIRED && N == 4096;
  constexpr bool TWREG = (KIND == ROWS_PAIRED || KIND == ROWS_PAIRED_SD) && N == 4096;
  float2 tb[Bases<N>::SIZE];
  if constexpr (TWREG) load_bases<N>(tb, a.tw, t);
  constexpr int NB2 = TWFULL ? Pass2Bases<N>::SIZE : 1;
  float2 w1[15], b2[NB2];
  if constexpr (TWFULL) {
    static_assert(Bases<N>::SIZE == TwBase<16>::NB + NB2, "one radix-16 butterfly in pass 1");
    twiddles16(w1, tb);
#pragma unroll
    for (int i = 0; i < NB2; ++i) b2[i] = tb[TwBase<16>::NB + i];
  }
  auto fft_inv = [&](float2 (&x)[PTS], int tt) {
    if constexpr (TWFULL) line_fft_w1<N, +1>(x, lds, tt, w1, b2);
    else if constexpr (TWREG) line_fft<N, +1>(x, lds, tt, tb);
    else line_fft<N, +1>(x, lds, tt, a.tw);
  };
  auto fft_fwd = [&](float2 (&x)[PTS], int tt) {
    if constexpr (TWFULL) line_fft_w1<N, -1>(x, lds, tt, w1, b2);
    else if constexpr (TWREG) line_fft<N, -1>(x, lds, tt, tb);
    else line_fft<N, -1>(x, lds, tt, a.tw);
  };

  float smax = 0.f;
  bool bad = false;
  if constexpr (KIND == ROWS_PAIRED || KIND == ROWS_PAIRED_SD) {
    constexpr bool SD = KIND == ROWS_PAIRED_SD;
#pragma unroll 1
    for (int pair = 0; pair < PAIRS; ++pair) {
      const int tp = launder(t);
      const int y0 = (yb * PAIRS + pair) * 2;
      const uint32_t vlo = Ge::off(y0, tp);
      const uint32_t vhi = Ge::off(y0, T - tp);
      // omega of rows y0, y0+1: omega0 stays in registers, omega1 is parked
      float2 X[PTS];
      row_load<N, true>(X, bu, tp, live, vlo, vhi, 0u, 0u);
      fft_inv(X, tp);
      float om0[PTS];
#pragma unroll
      for (int q = 0; q < PTS; ++q) {
        om0[q] = X[q].x;
        vpark[tp + q * T] = X[q].y;
      }
      // f = 0: u pair -> P'(y0), P'(y1); f = 1: v pair -> Q'(y0), Q'(y1)
      // (SD: f = 0: s pair -> S', f = 1: d pair -> D', with the CFL maximum of |s|, |d|)
#pragma unroll 1
      for (int f = 0; f < 2; ++f) {
        const int tr = launder(tp);
        const uint32_t vlr = Ge::off(y0, tr);
        const uint32_t vhr = Ge::off(y0, T - tr);
        row_load<N, true>(X, bu, tr, live, vlr, vhr, (uint32_t)(1 + f) * PLANE_B, 0u);
        fft_inv(X, tr);
        if constexpr (SD) {
#pragma unroll
          for (int q = 0; q < PTS; ++q) {
            const float m0 = fabsf(X[q].x), m1 = fabsf(X[q].y);
            bad |= !(m0 <= 3.4e38f) || !(m1 <= 3.4e38f);
            smax = fmaxf(smax, fmaxf(m0, m1));
          }
        }
#pragma unroll
        for (int q = 0; q < PTS; ++q)
          X[q] = make_float2(X[q].x * om0[q], X[q].y * vpark[tr + q * T]);
        fft_fwd(X, tr);
        row_unpack_store<N, true>(X, lds, tr, live, bz, vlr, (uint32_t)f * PLANE_B, 0u);
      }
    }
  } else {
#pragma unroll 1
  for (int pair = 0; pair < PAIRS; ++pair) {
    const int tp = launder(t);
    const int y0 = (yb * PAIRS + pair) * 2;
    // lane byte offsets of (row pair y0 / 2, kx = t) and (row pair, kx = T - t) in a plane
    const uint32_t vlo = Ge::off(y0, tp);
    const uint32_t vhi = Ge::off(y0, T - tp);
    // d of rows y0, y0+1 from one complex transform: B' = D0 + i D1, Hermitian-extended
    float2 B[PTS];
    row_load<N, true>(B, bu, tp, live, vlo, vhi, 2 * PLANE_B, 0u);
    fft_inv(B, tp);
    float d0[PTS];
#pragma unroll
    for (int q = 0; q < PTS; ++q) {
      d0[q] = B[q].x;
      vpark[tp + q * T] = B[q].y;
    }

#pragma unroll 1
    for (int r = 0; r < 2; ++r) {
      const int tr = launder(tp);
      const int y = y0 + r;
      const uint32_t vlr = Ge::off(y0, tr);
      const uint32_t vhr = Ge::off(y0, T - tr);
      // A' = omega' + i s' of row y (its 8-byte slot r in the row pair)
      float2 A[PTS];
      row_load<N, false>(A, bu, tr, live, vlr, vhr, (uint32_t)r * 8u, PLANE_B + (uint32_t)r * 8u);
      fft_inv(A, tr);
      // physical omega = Re A, s = Im A, d = Re/Im B; packed product Z = s omega + i d omega
      const uint32_t vvel = ((uint32_t)y * N + (uint32_t)tr) * 8u;
#pragma unroll
      for (int q = 0; q < PTS; ++q) {
        const float om = A[q].x, sv = A[q].y;
        const float dv = r == 0 ? d0[q] : vpark[tr + q * T];
        if (a.flags & ROW_SPEED) {
          const float sp = fmaxf(fabsf(sv), fabsf(dv));  // = |u| + |v|
          bad |= !(sp <= 3.4e38f);
          smax = fmaxf(smax, sp);
        }
        if (do_vel && live)
          st2(bv, make_float2(0.5f * (sv + dv), 0.5f * (sv - dv)), vvel, (uint32_t)(q * T) * 8u);
        A[q] = make_float2(sv * om, dv * om);
      }
      fft_fwd(A, tr);
      row_unpack_store<N, false>(A, lds, tr, live, bz, vlr, (uint32_t)r * 8u,
                                 PLANE_B + (uint32_t)r * 8u);
    }

    if constexpr (FORCE) {
      // pointer forcing of rows y0, y0+1 at cell centres, packed as re/im of one transform
      const float yu0 = ((float)y0 + 0.5f) / (float)N;
      const float yu1 = ((float)(y0 + 1) + 0.5f) / (float)N;
#pragma unroll
      for (int q = 0; q < PTS; ++q) {
        const float xu = ((float)(tp + q * T) + 0.5f) / (float)N;
        B[q] = live ? make_float2(pointer_curl(xu, yu0, a.splats), pointer_curl(xu, yu1, a.splats))
                    : make_float2(0.f, 0.f);
      }
      fft_fwd(B, tp);
      row_unpack_store<N, true>(B, lds, tp, live, bf, vlo, 0u, 0u);
    }
  }
  }

  if (a.flags & ROW_SPEED) {
    // wave max, workgroup max through LDS, one atomic per workgroup into the block's slot
    float* wmax = sm.wmax;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) smax = fmaxf(smax, __shfl_xor(smax, off));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = smax;
    __syncthreads();
    if (threadIdx.x == 0) {
      float m = wmax[0];
      for (int w = 1; w < G * L::T / 64; ++w) m = fmaxf(m, wmax[w]);
      atomicMax(&I->st.speed_bits[bx % DevState::SPEED_SLOTS], __float_as_uint(m));
    }
    if (bad) atomicOr(&I->st.nonfinite, 1u);
  }
}

template <int N, int G, int KIND>
__global__ void __launch_bounds__(G * Line<N>::T,
                                  N == 4096 && KIND != ROWS_SINGLE_FORCE ? 3 : 1) k_rows(RowArgs a) {
  __shared__ RowSmem<N, G> sm;
  rows_body<N, G, KIND>(a, (int)blockIdx.x, (int)blockIdx.y, sm);
}

// Row kernel, field-parallel form (512^2 and 1024^2, where one 64-thread group per row pair
// leaves most SIMDs idle: 512 one-wave workgroups at 1024^2). One row pair per workgroup and three
// thread groups of T threads: the three inverse transforms side by side, then the two forward
// transforms side by side while the third group transforms the pointer forcing (or zeros, which
// keeps the barriers in step). Paired kinds: groups 0, 1, 2 take the omega, u, v pairs (omega, s, d
// at stage 0), omega goes to LDS, groups 1 and 2 form and transform the P' and Q' (S', D') pairs.
// Single kinds (stage 0 with dye velocity): group 0 takes the d pair, groups 1 and 2 the
// omega + i s rows y0 and y0 + 1, the d pair goes to LDS, group 1 + r forms the products of row
// y0 + r. Every value goes through the same device code as in k_rows, so the state is the same bit
// for bit.
// Sizes with the field-parallel row kernel (the default there): 7.7 / 14.0 us per launch at
// 512^2 / 1024^2 against 9.2 / 16.7 us with one group per row pair; at 2048^2 (two waves per row
// pair already, 67 KB of LDS per three-group workgroup) it took 41.3 us against 24.3 us.
template <int N>
struct RowFp {
  static constexpr bool on = N == 512 || N == 1024;
};
constexpr int ROWFP_GROUPS = 3;

template <int N, int KIND>
__global__ void __launch_bounds__(ROWFP_GROUPS * Line<N>::T) k_rows_fp(RowArgs a) {
  using L = Line<N>;
  using Ge = Geo<N>;
  constexpr int T = L::T, PTS = L::PTS;
  constexpr uint32_t PLANE_B = (uint32_t)(Ge::PLANE * 8);
  constexpr bool PAIRED = KIND == ROWS_PAIRED || KIND == ROWS_PAIRED_SD;
  constexpr bool SD = KIND == ROWS_PAIRED_SD;
  constexpr bool FORCE = KIND == ROWS_SINGLE_FORCE;
  static_assert(T % 64 == 0, "thread groups of whole waves");
  __shared__ float2 lds_all[ROWFP_GROUPS * L::LDS];
  __shared__ float2 share[N];  // the omega pair (paired kinds) or the d pair (single kinds)
  __shared__ float wmax[ROWFP_GROUPS * T / 64];
  Inst* I = a.inst + blockIdx.y;
  if (!I->st.active) return;  // uniform over the block
  const size_t inst = blockIdx.y;
  const bool do_vel = a.flags & ROW_VEL;
  const Buf bu = make_buf(a.mu + inst * a.s_mu, 3 * PLANE_B);
  const Buf bz = make_buf(a.mz + inst * a.s_mz, 2 * PLANE_B);
  const Buf bf = make_buf(FORCE ? a.mf + inst * a.s_mf : a.mf, FORCE ? PLANE_B : 0u);
  const Buf bv = make_buf(do_vel ? a.vel + inst * a.s_vel : a.vel,
                          do_vel ? (uint32_t)((size_t)N * N * 8) : 0u);
  constexpr int NB = N / 2, NB8 = (NB + 7) / 8 * 8;
  const int bx = (int)blockIdx.x;
  const int lb = (bx % 8) * (NB8 / 8) + bx / 8;  // XCD-aware, as in k_rows
  if (lb >= NB) return;  // uniform over the block
  const int g = threadIdx.x / T, t = threadIdx.x - g * T;
  const int y0 = 2 * lb;
  float2* lds = lds_all + g * L::LDS;
  const uint32_t vlo = Ge::off(y0, t);
  const uint32_t vhi = Ge::off(y0, T - t);
  float smax = 0.f;
  bool bad = false;

  // inverse transforms, one field per group
  float2 X[PTS];
  if constexpr (PAIRED) {
    row_load<N, true>(X, bu, t, true, vlo, vhi, (uint32_t)g * PLANE_B, 0u);
  } else {
    // group 0: d pair; group 1 + r: omega + i s of row y0 + r
    if (g == 0) {
      row_load<N, true>(X, bu, t, true, vlo, vhi, 2 * PLANE_B, 0u);
    } else {
      const uint32_t r = (uint32_t)(g - 1);
      row_load<N, false>(X, bu, t, true, vlo, vhi, r * 8u, PLANE_B + r * 8u);
    }
  }
  line_fft<N, +1>(X, lds, t, a.tw);
  if (g == 0) {
#pragma unroll
    for (int q = 0; q < PTS; ++q) share[t + q * T] = X[q];
  }
  if constexpr (SD) {
    if (g > 0) {
#pragma unroll
      for (int q = 0; q < PTS; ++q) {
        const float m0 = fabsf(X[q].x), m1 = fabsf(X[q].y);
        bad |= !(m0 <= 3.4e38f) || !(m1 <= 3.4e38f);
        smax = fmaxf(smax, fmaxf(m0, m1));
      }
    }
  }
  __syncthreads();

  // products and forward transforms (groups 1, 2); pointer forcing or zeros (group 0)
  if (g > 0) {
    if constexpr (PAIRED) {
#pragma unroll
      for (int q = 0; q < PTS; ++q) {
        const float2 om = share[t + q * T];
        X[q] = make_float2(X[q].x * om.x, X[q].y * om.y);
      }
    } else {
      const int r = g - 1;
      const uint32_t vvel = ((uint32_t)(y0 + r) * N + (uint32_t)t) * 8u;
#pragma unroll
      for (int q = 0; q < PTS; ++q) {
        const float om = X[q].x, sv = X[q].y;
        const float2 dp = share[t + q * T];
        const float dv = r == 0 ? dp.x : dp.y;
        if (a.flags & ROW_SPEED) {
          const float sp = fmaxf(fabsf(sv), fabsf(dv));  // = |u| + |v|
          bad |= !(sp <= 3.4e38f);
          smax = fmaxf(smax, sp);
        }
        if (do_vel)
          st2(bv, make_float2(0.5f * (sv + dv), 0.5f * (sv - dv)), vvel, (uint32_t)(q * T) * 8u);
        X[q] = make_float2(sv * om, dv * om);
      }
    }
  } else {
    if constexpr (FORCE) {
      const float yu0 = ((float)y0 + 0.5f) / (float)N;
      const float yu1 = ((float)(y0 + 1) + 0.5f) / (float)N;
#pragma unroll
      for (int q = 0; q < PTS; ++q) {
        const float xu = ((float)(t + q * T) + 0.5f) / (float)N;
        X[q] = make_float2(pointer_curl(xu, yu0, a.splats), pointer_curl(xu, yu1, a.splats));
      }
    } else {
#pragma unroll
      for (int q = 0; q < PTS; ++q) X[q] = make_float2(0.f, 0.f);
    }
  }
  line_fft<N, -1>(X, lds, t, a.tw);
  row_unpack_stage<N>(X, lds, t);
  if (g > 0) {
    if constexpr (PAIRED) {
      row_unpack_emit<N, true>(lds, t, true, bz, vlo, (uint32_t)(g - 1) * PLANE_B, 0u);
    } else {
      const uint32_t r = (uint32_t)(g - 1);
      row_unpack_emit<N, false>(lds, t, true, bz, vlo, r * 8u, PLANE_B + r * 8u);
    }
  } else if (FORCE) {
    row_unpack_emit<N, true>(lds, t, true, bf, vlo, 0u, 0u);
  }

  if (a.flags & ROW_SPEED) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) smax = fmaxf(smax, __shfl_xor(smax, off));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = smax;
    __syncthreads();
    if (threadIdx.x == 0) {
      float m = wmax[0];
      for (int w = 1; w < ROWFP_GROUPS * T / 64; ++w) m = fmaxf(m, wmax[w]);
      atomicMax(&I->st.speed_bits[bx % DevState::SPEED_SLOTS], __float_as_uint(m));
    }
    if (bad) atomicOr(&I->st.nonfinite, 1u);
  }
}

// ---------------------------------------------------------------------------------------------
// column kernel
// ---------------------------------------------------------------------------------------------

// P_hat waits for Q_hat in LDS instead of in HBM: in column groups of two at 4096, in the groups
// of the batch's grouped launches (8 columns at 256, 4 at 512 and 1024) and in groups of four at
// 128 to 2048 (the batch classes' and serving chunks' width once they fill the chip). A column's
// in-band row blocks of a thread (at 4096 blocks 0..5 and 10..15; 6..9 are out of band for every
// lane) take NSLOT slot blocks of T values; the two partly in-band blocks around the out-of-band run
// (QA: t <= K - QA T, QB: t >= N - K - QB T) share one where their lanes are disjoint (4096,
// 1024). The park (C NSLOT T float2) and a half-size exchange buffer for the Q' transform
// (C LSTRIDE floats) hold the full-size exchanges of the other transforms, which overlap both. At
// 4096 they fit in 78.8 KB, so two workgroups still share a CU; the grouped launches' largest
// (1024) takes 39.4 KB, four workgroups per CU as before; four columns at 2048 take 78.8 KB, two
// 512-thread workgroups per CU as before.
template <int N, int C>
struct ColPark {
  static constexpr bool on = (N == 4096 && C == 2) || (N == 256 && C == 8) ||
                             (C == 4 && N >= 128 && N <= 2048);
  static constexpr int T = Line<N>::T, PTS = Line<N>::PTS, K = Geo<N>::K;
  static constexpr int O0 = ColBand<N>::out_first(), ON = ColBand<N>::out_count();
  static constexpr int QA = O0 - 1, QB = O0 + ON;  // the blocks around the out-of-band run
  static constexpr bool SHARE = K - QA * T < (N - K) - QB * T;
  static constexpr int NSLOT = PTS - ON - (SHARE ? 1 : 0);
  // slot block of row block q (q not out of band)
  static __device__ __forceinline__ int slot(int q) {
    return q < O0 ? q : SHARE ? (q == QB ? QA : q - ON - 1) : q - ON;
  }
  static constexpr int LDS2 = C * NSLOT * T + C * ColCfg<N>::LSTRIDE / 2;  // float2
  static constexpr bool fits() {
    return !on || (ON > 0 && ColCfg<N>::LSTRIDE % 2 == 0 && Line<N>::NPASS == 3 &&
                   C * ColCfg<N>::LSTRIDE <= LDS2);
  }
};
static_assert(ColPark<4096, 2>::QA == 5 && ColPark<4096, 2>::QB == 10 &&
                  ColPark<4096, 2>::SHARE && ColPark<4096, 2>::NSLOT == 11 &&
                  2 * (ColPark<4096, 2>::LDS2 * 8 + 16) <= 160 * 1024,
              "4096: blocks 5 and 10 in band for disjoint lanes; two workgroups per CU");
static_assert(ColPark<1024, 4>::SHARE && ColPark<1024, 4>::NSLOT == 11 &&
                  !ColPark<512, 4>::SHARE && ColPark<512, 4>::NSLOT == 6 &&
                  !ColPark<256, 8>::SHARE && ColPark<256, 8>::NSLOT == 6 &&
                  4 * (ColPark<1024, 4>::LDS2 * 8 + 16) <= 160 * 1024,
              "grouped classes: slot blocks; four workgroups of the largest per CU");
static_assert(ColPark<4096, 2>::fits() && ColPark<2048, 4>::fits() && ColPark<1024, 4>::fits() &&
                  ColPark<512, 4>::fits() && ColPark<256, 8>::fits() && ColPark<256, 4>::fits() &&
                  ColPark<128, 4>::fits() && 2 * (ColPark<2048, 4>::LDS2 * 8 + 16) <= 160 * 1024,
              "park and half exchange hold the full exchanges; two 2048 workgroups per CU");

template <int N, int C>
struct ColSmem {
  float2 lds[ColPark<N, C>::on ? ColPark<N, C>::LDS2 : C * ColCfg<N>::LSTRIDE];
  float speed;
};

// One workgroup of the column kernel: workgroup bx (of the instance's grid of column groups) of
// instance iy. Launched on its own by k_cols, or as an item of the grouped launch k_cols_group.
template <int N, int C, int MODE, bool FORCE>
__device__ __forceinline__ void cols_body(const ColArgs& a, int bx, int iy, ColSmem<N, C>& sm) {
  using L = Line<N>;
  using Ge = Geo<N>;
  constexpr int T = L::T, PTS = L::PTS, K = Ge::K, KC = Ge::KC, NJ = Ge::NJ;
  constexpr int LSTRIDE = ColCfg<N>::LSTRIDE;
  constexpr int GROUPS = (KC + C - 1) / C;
  constexpr int G8 = (GROUPS + 7) / 8 * 8;
  constexpr uint32_t PLANE_B = (uint32_t)(Ge::PLANE * 8), STATE_B = (uint32_t)(Ge::STATE * 8);
  constexpr uint32_t ROWS_B = (uint32_t)T * Ge::KP * 8;  // row block q -> q + 1 (T / 2 row pairs)
  float2* lds_all = sm.lds;

  // XCD-aware mapping: consecutive workgroups go round-robin over the 8 XCDs, so give each XCD
  // a contiguous range of column groups (neighbouring groups share L2 lines of the planes).
  const int bid = bx;
  const int cg = (bid % 8) * (G8 / 8) + bid / 8;
  if (cg >= GROUPS) return;
  Inst* I = a.inst + iy;
  if (!I->st.active) return;  // uniform over the block
  // a copy taken before the kernel's first store: the loads are known unclobbered, so they are
  // scalar loads and the parameters stay in SGPRs (a reference would be re-read from memory after
  // every buffer store, each time with a wait for all loads in flight)
  const Phys ph = I->ph;
  const size_t inst = iy;
  const Buf bz = make_buf(a.mz + inst * a.s_mz, 2 * PLANE_B);
  const Buf bu = make_buf(a.mu + inst * a.s_mu, 3 * PLANE_B);
  const Buf bw0 = make_buf(a.w0 + inst * a.s_state, STATE_B);
  const Buf bwc = make_buf(a.wcur + inst * a.s_state, STATE_B);
  // state leaving the stage (the right-hand side in COLS_RHS). Without the LDS park (PARK) P_hat
  // is parked here first, each thread at the indices it later overwrites with the new state: a
  // write and a read of the in-band P_hat in HBM.
  constexpr bool PARK = ColPark<N, C>::on;
  using PK = ColPark<N, C>;
  const Buf bwo = MODE == COLS_RHS ? make_buf(a.rhs_out, STATE_B)
                                   : make_buf(a.wout + inst * a.s_state, STATE_B);
  const Buf bf = make_buf(FORCE ? a.mf + inst * a.s_mf : a.mf, FORCE ? PLANE_B : 0u);
  const Buf bfk = make_buf(FORCE ? a.fk + inst * a.s_state : a.fk, FORCE ? STATE_B : 0u);
  const int tid = threadIdx.x;
  const int c = ColCfg<N>::template col<C>(tid), t = ColCfg<N>::template row<C>(tid);
  const int kx = cg * C + c;
  const bool active = kx < KC;
  float2* lds = lds_all + c * LSTRIDE;
  // PARK: the column's slot blocks (slot s of thread t at park[s T + t]) and its half-size
  // exchange buffer, after the C columns' parks
  float2* park = lds_all + c * PK::NSLOT * T;
  float* ldsh = reinterpret_cast<float*>(lds_all + C * PK::NSLOT * T) + c * LSTRIDE;
  const float fkx = (float)kx;
  // per-lane byte offsets: row t of column kx in a tiled mixed plane / in the compact state
  // (row block q adds q T rows: q T / 2 row pairs, T is even)
  const uint32_t vm = Ge::off(t, kx);
  // the same, out of range for a lane past the last column (its loads read zero, with no branch)
  const uint32_t vmp = active ? vm : BUF_OOB;
  const uint32_t vs = ((uint32_t)kx * Ge::PITCH + (uint32_t)t) * 8u;
  // row block q holds rows t + q T: the low half are the compact rows j = i, the high half
  // j = i - (N - NJ) (T divides N / 2, so a block never spans both halves)
  auto soff_s = [](int q) -> uint32_t {
    return (uint32_t)(q * T - (2 * q >= PTS ? N - NJ : 0)) * 8u;
  };
  auto inband = [&](int q) -> bool {
    const int i = t + q * T;
    return active && (2 * q < PTS ? i <= K : i >= N - K);
  };
  // row blocks out of band for every lane: neither parked, computed nor transformed
  auto qout = [](int q) constexpr { return ColBand<N>::out(q); };

  if constexpr (MODE != COLS_PRIME) {
    double dt;
    uint32_t step;
    if constexpr (MODE == COLS_RHS) {
      dt = a.dt_override;
      step = a.step_override;
    } else {
      step = I->st.step;
      if (a.stage == 0) {
        // the step's speed maximum: max over the row kernel's slots
        if (tid < 64) {
          float v = __uint_as_float(I->st.speed_bits[tid % DevState::SPEED_SLOTS]);
#pragma unroll
          for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
          if (tid == 0) sm.speed = v;
        }
        __syncthreads();
        const float speed = sm.speed;
        dt = ph.fixed_dt > 0.0 ? ph.fixed_dt : cfl_dt(speed, ph);
        if (bid == 0 && tid == 0) {
          I->st.dt = dt;
          I->st.speed = speed;
        }
      } else {
        dt = I->st.dt;
      }
    }
    // pointer-forcing spectrum: at stage 0 transform F' and keep the in-band modes in fk
    if constexpr (FORCE) {
      if (MODE == COLS_RHS || a.stage == 0) {
        float2 F[PTS];
#pragma unroll
        for (int i = 0; i < PTS; ++i) {
          const int q = dft_load_order<PTS>(i);
          F[q] = ld2(bf, vmp, q * ROWS_B);
        }
        line_fft<N, -1>(F, lds, t, a.tw);
#pragma unroll
        for (int q = 0; q < PTS; ++q)
          if (inband(q)) st2(bfk, F[q], vs, soff_s(q));
      }
    }
    float2 Q[PTS];
    auto load_q = [&] {
#pragma unroll
      for (int i = 0; i < PTS; ++i) {
        const int q = dft_load_order<PTS>(i);
        Q[q] = ld2(bz, vmp, PLANE_B + q * ROWS_B);
      }
    };
    {
      float2 P[PTS];
#pragma unroll
      for (int i = 0; i < PTS; ++i) {
        const int q = dft_load_order<PTS>(i);
        P[q] = ld2(bz, vmp, q * ROWS_B);
      }
      line_fft<N, -1>(P, lds, t, a.tw);
      if constexpr (PARK) {
        // Q' is fetched while the workgroup meets at the barrier and P_hat goes to the park
        // (0.7% of a step; not with pointer forcing, whose step kernel would spill 15 VGPRs)
        if constexpr (!FORCE) load_q();
        __syncthreads();  // the last exchange is read (the park overlaps it)
#pragma unroll
        for (int q = 0; q < PTS; ++q)
          if (!qout(q) && inband(q)) park[PK::slot(q) * T + t] = P[q];
      } else {
#pragma unroll
        for (int q = 0; q < PTS; ++q)
          if (!qout(q) && inband(q)) st2(bwo, P[q], vs, soff_s(q));
      }
    }
    {
      if constexpr (!PARK || FORCE) load_q();
      if constexpr (PARK) {
        // the half-size exchanges leave the park in place; then N_hat replaces P_hat in the
        // thread's own slots (no other thread reads them: no barrier)
        line_fft_half<N, -1>(Q, ldsh, t, a.tw);
#pragma unroll
        for (int q = 0; q < PTS; ++q) {
          if (qout(q) || !inband(q)) continue;
          const int idx = t + q * T;
          const float fky = (float)wn(idx, N);
          float2& pq = park[PK::slot(q) * T + t];
          pq = a.sd ? nhat(pq, Q[q], 0.5f * (fkx + fky), 0.5f * (fkx - fky))
                    : nhat(pq, Q[q], fkx, fky);
        }
      } else {
        // Q_hat parks in the exchange buffer, each thread at its own indices
        line_fft<N, -1>(Q, lds, t, a.tw);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < PTS; ++q)
          if (!qout(q)) lds[swz(t + q * T)] = Q[q];
      }
    }

    // per-mode right-hand side and RK stage, one row block at a time (rolled, so only one
    // block's values are live). The loop body covers two blocks with two operand sets, A for even
    // and B for odd q: the operands of block q + 1 are fetched before block q is computed, into
    // the set block q does not use, so no register copies wait on loads in flight.
    static_assert(PTS % 2 == 0, "the mode loop takes row blocks in pairs");
    const float fdt = (float)dt;
    const bool need_wc = MODE == COLS_STEP && a.stage != 0;
    const float2 zero = make_float2(0.f, 0.f);
    // Loads and stores of lanes out of band take an offset out of range (no branch around them).
    float2 pA = zero, w0A = zero, wcA = zero, fkA = zero;
    float2 pB = zero, w0B = zero, wcB = zero, fkB = zero;
    auto fetch = [&](int q, float2& p_, float2& w0_, float2& wc_, float2& fk_) {
      const uint32_t so = soff_s(q), v = q < PTS && inband(q) ? vs : BUF_OOB;
      if constexpr (!PARK) p_ = ld2(bwo, v, so);
      if (MODE == COLS_STEP) {
        w0_ = ld2(bw0, v, so);
        // issued at stage 0 too (reading zero): a load under a run-time condition makes the
        // compiler merge the untaken case into the register at once, waiting for the other loads
        wc_ = ld2(bwc, need_wc ? v : BUF_OOB, so);
      }
      if (FORCE) fk_ = ld2(bfk, v, so);
    };
    auto block = [&](int q, const float2 pv, const float2 w0v, const float2 wcv, const float2 fkv) {
      float2 wv = zero, rhs = zero;
      const bool inb = inband(q);
      if (inb) {
        const int idx = t + q * T;
        const float fky = (float)wn(idx, N);
        const float k2 = fkx * fkx + fky * fky;
        if (k2 > 0.f) {
          // N_hat = i kx P_hat + i ky Q_hat; from S_hat = P_hat + Q_hat, D_hat = P_hat - Q_hat it
          // is i ((kx + ky) S_hat + (kx - ky) D_hat) / 2 (PARK: already in the slot)
          const float2 nh =
              PARK ? park[PK::slot(q) * T + t]
                   : (a.sd ? nhat(pv, lds[swz(idx)], 0.5f * (fkx + fky), 0.5f * (fkx - fky))
                           : nhat(pv, lds[swz(idx)], fkx, fky));
          rhs = make_float2(-nh.x, -nh.y);
          if constexpr (FORCE) rhs = cadd(rhs, fkv);
          if (ph.force_amp > 0.f)
            rhs = cadd(rhs, stoch_force(idx, kx, fkx, fky, k2, ph, step, fdt, N));
          if constexpr (MODE == COLS_STEP) {
            const IFactor e = ifactor_half((double)kx, (double)k2, ph, dt);
            wv = rk_update(a.stage, e, w0v, a.stage == 0 ? w0v : wcv, rhs, dt);
          }
        }
      }
      st2(bwo, MODE == COLS_RHS ? rhs : wv, inb ? vs : BUF_OOB, soff_s(q));
      // the new state replaces Q_hat in this thread's own slot of the exchange buffer (PARK:
      // N_hat in its slot, in band only, as a shared slot block holds the other block's lanes)
      if constexpr (MODE == COLS_STEP) {
        if constexpr (PARK) {
          if (inb) park[PK::slot(q) * T + t] = wv;
        } else {
          lds[swz(t + q * T)] = wv;
        }
      }
    };
    // pairs of blocks, skipping the pairs out of band (pair pi starts at block qof(pi))
    constexpr int SK0 = ColBand<N>::skip_first(), SKN = ColBand<N>::skip_count();
    constexpr int NPI = (PTS - SKN) / 2;
    auto qof = [](int pi) { return 2 * pi + (2 * pi >= SK0 ? SKN : 0); };
    fetch(qof(0), pA, w0A, wcA, fkA);
#pragma unroll 1
    for (int pi = 0; pi < NPI; ++pi) {
      const int q = qof(pi);
      fetch(q + 1, pB, w0B, wcB, fkB);
      block(q, pA, w0A, wcA, fkA);
      fetch(qof(pi + 1), pA, w0A, wcA, fkA);  // past the last block: reads nothing (no branch)
      block(q + 1, pB, w0B, wcB, fkB);
    }
    if constexpr (MODE == COLS_RHS) return;
  }

  // velocity pack of the new state and inverse column transforms; the 1/N per inverse pass of
  // both directions (src/gpu/fft.js:59) is applied here as one exact 1/N^2
  constexpr float S = 1.0f / ((float)N * (float)N);
  // the new state (zero out of band), scaled, in registers for the three transforms
  float2 W[PTS];
#pragma unroll
  for (int q = 0; q < PTS; ++q) {
    float2 w = make_float2(0.f, 0.f);
    if (!qout(q)) {
      if constexpr (MODE == COLS_PRIME)
        w = inband(q) ? ld2(bwc, vs, soff_s(q)) : make_float2(0.f, 0.f);
      else if constexpr (PARK)
        w = inband(q) ? park[PK::slot(q) * T + t] : make_float2(0.f, 0.f);
      else
        w = lds[swz(t + q * T)];
    }
    W[q] = cscale(w, S);
  }
  // After stage 2 and in the prime the next row kernel is a stage-0 one: it takes s' = u' + v'
  // and d' = u' - v' (s_hat = i (ky - kx) w_hat / k^2, d_hat = i (ky + kx) w_hat / k^2) in the
  // planes of u' and v'.
  const bool sdo = MODE == COLS_PRIME || a.stage == 2;
#pragma unroll 1
  for (int f = 0; f < 3; ++f) {
    // re-derive the lane's indices each field so nothing per-value stays live across transforms
    const int tl = launder(t);
    const int kxl = launder(kx);
    const uint32_t vml = Ge::off(tl, kxl);
    const float fkxl = (float)kxl;
    // field f = 1, 2: g = (cy ky + cx kx) / k^2, the transform input i g w_hat
    const float cy = f == 1 || sdo ? 1.f : 0.f;
    const float cx = f == 1 ? (sdo ? -1.f : 0.f) : (sdo ? 1.f : -1.f);
    float2 X[PTS];
#pragma unroll
    for (int q = 0; q < PTS; ++q) {
      if (qout(q)) {
        X[q] = make_float2(0.f, 0.f);
        continue;
      }
      const int idx = tl + q * T;
      const float2 w = W[q];
      const float fky = (float)wn(idx, N);
      const float k2 = fkxl * fkxl + fky * fky;
      // hardware reciprocal (1 ulp; the IEEE quotient takes a dozen instructions per value)
      const float inv = k2 > 0.f ? __builtin_amdgcn_rcpf(k2) : 0.f;
      // u_hat = i ky / k^2 omega_hat, v_hat = -i kx / k^2 omega_hat
      const float g = (cy * fky + cx * fkxl) * inv;
      X[q] = f == 0 ? w : cscale(cmuli(w), g);
    }
    line_fft<N, +1>(X, lds, tl, a.tw);
    if (kxl < KC) {
#pragma unroll
      for (int q = 0; q < PTS; ++q) st2(bu, X[q], vml, (uint32_t)f * PLANE_B + q * ROWS_B);
    }
  }
  if constexpr (MODE == COLS_STEP) {
    if (a.stage == 2) {
      // The last column workgroup of the instance to finish runs the per-step epilogue. It only
      // has to come after every workgroup's reads of the instance scalars (step, dt, active),
      // which each wave consumed at its start, so the counter is a relaxed atomic with no fence:
      // an agent-scope fence per workgroup writes back the XCD's L2 and cost about 30 us per
      // launch at 4096^2. The next launch sees the epilogue's stores through the kernel boundary.
      __syncthreads();
      if (tid < 64) {  // wave 0
        unsigned last = 0u;
        if (tid == 0)
          last = atomicAdd(&I->st.cols_done, 1u) == (unsigned)(GROUPS - 1) ? 1u : 0u;
        last = __shfl(last, 0);
        if (last) {
          if (tid < DevState::SPEED_SLOTS)
            __hip_atomic_store(&I->st.speed_bits[tid], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (tid == 0) {
            step_epilogue(I->st);
            __hip_atomic_store(&I->st.cols_done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          }
        }
      }
    }
  }
}

// With the LDS park the step kernels are held to four waves per SIMD (128 VGPRs; hip-clang's
// second launch bound is waves per SIMD): with 512 threads (4096, 2048) two workgroups per CU, as
// many as the LDS allows (left to itself the compiler gave the pointer-forcing step at 4096 129
// VGPRs, three waves, which at 512 threads is one workgroup per CU). The pointer-forcing kernels
// of 256 threads stay unbound: bound they would spill, and three waves per SIMD are still three
// workgroups.
template <int N, int C, int MODE, bool FORCE>
__global__ void __launch_bounds__(C * Line<N>::T,
                                  (ColPark<N, C>::on && MODE == COLS_STEP &&
                                           (!FORCE || C * Line<N>::T >= 512)
                                       ? 4
                                       : 1))
    k_cols(ColArgs a) {
  __shared__ ColSmem<N, C> sm;
  cols_body<N, C, MODE, FORCE>(a, (int)blockIdx.x, (int)blockIdx.y, sm);
}

// ---------------------------------------------------------------------------------------------
// column kernel, field-parallel form (grids of 512^2 to 2048^2, where one thread group per
// column leaves SIMDs idle: 342 one-wave workgroups at 1024^2)
// ---------------------------------------------------------------------------------------------
// One column per workgroup and GR = 2 or 3 thread groups of T threads side by side: groups 0 and 1
// transform P' and Q' (with three groups, group 2 transforms the pointer forcing at stage 0, else
// zeros, which keeps the barriers in step; with two, the force transform runs first, group 1
// transforming zeros); P_hat and Q_hat both stay in LDS (no round trip through the output state);
// the per-mode loop is split over the groups by pairs of row blocks; then the inverse transforms
// of omega', u', v' (s', d' after stage 2) are dealt to the groups. Every value goes through the
// same device code as in k_cols, so the state is the same bit for bit.
template <int N, int GR, bool FORCE>
__global__ void __launch_bounds__(GR * Line<N>::T) k_cols_fp(ColArgs a) {
  using L = Line<N>;
  using Ge = Geo<N>;
  constexpr int T = L::T, PTS = L::PTS, K = Ge::K, KC = Ge::KC, NJ = Ge::NJ;
  constexpr int LSTRIDE = ColCfg<N>::LSTRIDE;
  constexpr int G8 = (KC + 7) / 8 * 8;
  constexpr uint32_t PLANE_B = (uint32_t)(Ge::PLANE * 8), STATE_B = (uint32_t)(Ge::STATE * 8);
  constexpr uint32_t ROWS_B = (uint32_t)T * Ge::KP * 8;
  static_assert(T % 64 == 0, "thread groups of whole waves");
  static_assert(GR == 2 || GR == 3, "two or three thread groups");
  __shared__ float2 lds_all[GR * LSTRIDE];

  const int bid = blockIdx.x;
  const int kx = (bid % 8) * (G8 / 8) + bid / 8;  // XCD-aware: a contiguous range of columns per XCD
  if (kx >= KC) return;
  Inst* I = a.inst + blockIdx.y;
  if (!I->st.active) return;  // uniform over the block
  const Phys ph = I->ph;
  const size_t inst = blockIdx.y;
  const Buf bz = make_buf(a.mz + inst * a.s_mz, 2 * PLANE_B);
  const Buf bu = make_buf(a.mu + inst * a.s_mu, 3 * PLANE_B);
  const Buf bw0 = make_buf(a.w0 + inst * a.s_state, STATE_B);
  const Buf bwc = make_buf(a.wcur + inst * a.s_state, STATE_B);
  const Buf bwo = make_buf(a.wout + inst * a.s_state, STATE_B);
  const Buf bf = make_buf(FORCE ? a.mf + inst * a.s_mf : a.mf, FORCE ? PLANE_B : 0u);
  const Buf bfk = make_buf(FORCE ? a.fk + inst * a.s_state : a.fk, FORCE ? STATE_B : 0u);
  const int tid = threadIdx.x;
  const int g = tid / T, t = tid - g * T;
  float2* ldsP = lds_all;
  float2* ldsQ = lds_all + LSTRIDE;
  float2* lds = lds_all + g * LSTRIDE;  // this group's exchange buffer
  const float fkx = (float)kx;
  const uint32_t vm = Ge::off(t, kx);
  const uint32_t vs = ((uint32_t)kx * Ge::PITCH + (uint32_t)t) * 8u;
  auto soff_s = [](int q) -> uint32_t {
    return (uint32_t)(q * T - (2 * q >= PTS ? N - NJ : 0)) * 8u;
  };
  auto inband = [&](int q) -> bool {
    const int i = t + q * T;
    return 2 * q < PTS ? i <= K : i >= N - K;
  };
  auto qout = [](int q) constexpr { return ColBand<N>::out(q); };

  double dt;
  const uint32_t step = I->st.step;
  if (a.stage == 0) {
    __shared__ float speed_s;
    if (tid < 64) {
      float v = __uint_as_float(I->st.speed_bits[tid % DevState::SPEED_SLOTS]);
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
      if (tid == 0) speed_s = v;
    }
    __syncthreads();
    const float speed = speed_s;
    dt = ph.fixed_dt > 0.0 ? ph.fixed_dt : cfl_dt(speed, ph);
    if (bid == 0 && tid == 0) {
      I->st.dt = dt;
      I->st.speed = speed;
    }
  } else {
    dt = I->st.dt;
  }

  // forward column transforms side by side: P' (group 0), Q' (group 1), F' (group 2, stage 0 with
  // pointer forcing; otherwise group 2 transforms zeros, so that every group meets the same
  // barriers). Two groups transform F' first: group 0 F', group 1 zeros.
  const bool force_now = FORCE && a.stage == 0;
  if (GR == 2 && force_now) {
    float2 X[PTS];
#pragma unroll
    for (int i = 0; i < PTS; ++i) {
      const int q = dft_load_order<PTS>(i);
      X[q] = ld2(bf, g == 0 ? vm : BUF_OOB, q * ROWS_B);
    }
    line_fft<N, -1>(X, lds, t, a.tw);
    if (g == 0) {
#pragma unroll
      for (int q = 0; q < PTS; ++q)
        if (inband(q)) st2(bfk, X[q], vs, soff_s(q));
    }
  }
  {
    const Buf src = g == 2 ? bf : bz;
    const uint32_t base = g == 1 ? PLANE_B : 0u;
    const uint32_t v = (g < 2 || force_now) ? vm : BUF_OOB;
    float2 X[PTS];
#pragma unroll
    for (int i = 0; i < PTS; ++i) {
      const int q = dft_load_order<PTS>(i);
      X[q] = ld2(src, v, base + q * ROWS_B);
    }
    line_fft<N, -1>(X, lds, t, a.tw);
    __syncthreads();
    if (g < 2) {
#pragma unroll
      for (int q = 0; q < PTS; ++q)
        if (!qout(q)) lds[swz(t + q * T)] = X[q];
    } else if (force_now) {
#pragma unroll
      for (int q = 0; q < PTS; ++q)
        if (inband(q)) st2(bfk, X[q], vs, soff_s(q));
    }
  }
  __syncthreads();

  // per-mode right-hand side and RK stage (as in k_cols), pairs of row blocks dealt to the groups
  static_assert(PTS % 2 == 0, "the mode loop takes row blocks in pairs");
  constexpr int SK0 = ColBand<N>::skip_first(), SKN = ColBand<N>::skip_count();
  constexpr int NPI = (PTS - SKN) / 2;
  auto qof = [](int pi) { return 2 * pi + (2 * pi >= SK0 ? SKN : 0); };
  {
    const float fdt = (float)dt;
    const bool need_wc = a.stage != 0;
    const float2 zero = make_float2(0.f, 0.f);
    auto block = [&](int q) {
      const bool inb = inband(q);
      const uint32_t so = soff_s(q), v = inb ? vs : BUF_OOB;
      const float2 w0v = ld2(bw0, v, so);
      const float2 wcv = ld2(bwc, need_wc ? v : BUF_OOB, so);
      const float2 fkv = FORCE ? ld2(bfk, v, so) : zero;
      const int idx = t + q * T;
      const float2 pv = ldsP[swz(idx)];
      float2 wv = zero;
      if (inb) {
        const float fky = (float)wn(idx, N);
        const float k2 = fkx * fkx + fky * fky;
        if (k2 > 0.f) {
          const float2 nh = a.sd ? nhat(pv, ldsQ[swz(idx)], 0.5f * (fkx + fky), 0.5f * (fkx - fky))
                                 : nhat(pv, ldsQ[swz(idx)], fkx, fky);
          float2 rhs = make_float2(-nh.x, -nh.y);
          if constexpr (FORCE) rhs = cadd(rhs, fkv);
          if (ph.force_amp > 0.f)
            rhs = cadd(rhs, stoch_force(idx, kx, fkx, fky, k2, ph, step, fdt, N));
          const IFactor e = ifactor_half((double)kx, (double)k2, ph, dt);
          wv = rk_update(a.stage, e, w0v, a.stage == 0 ? w0v : wcv, rhs, dt);
        }
      }
      st2(bwo, wv, v, so);
      // the new state replaces P_hat in its slot (read above by this thread only)
      ldsP[swz(idx)] = wv;
    };
#pragma unroll 1
    for (int pi = g; pi < NPI; pi += GR) {
      const int q = qof(pi);
      block(q);
      block(q + 1);
    }
  }
  __syncthreads();

  // velocity pack of the new state and the inverse column transforms of the three fields, dealt
  // to the groups (two groups: group 0 omega' and v', group 1 u' and a transform of zeros, so that
  // both meet the same barriers)
  constexpr float S = 1.0f / ((float)N * (float)N);
  constexpr int ROUNDS = (3 + GR - 1) / GR;
  const bool sdo = a.stage == 2;
  // with more than one round the new state is held in registers for all of them
  float2 W[PTS];
  if constexpr (ROUNDS > 1) {
#pragma unroll
    for (int q = 0; q < PTS; ++q)
      W[q] = qout(q) ? make_float2(0.f, 0.f) : cscale(ldsP[swz(t + q * T)], S);
    __syncthreads();  // every group has read the new state before group 0 exchanges over it
  }
#pragma unroll 1
  for (int r = 0; r < ROUNDS; ++r) {
    const int f = g + r * GR;
    const int tl = launder(t);
    const float cy = f == 1 || sdo ? 1.f : 0.f;
    const float cx = f == 1 ? (sdo ? -1.f : 0.f) : (sdo ? 1.f : -1.f);
    float2 X[PTS];
#pragma unroll
    for (int q = 0; q < PTS; ++q) {
      if (qout(q) || f > 2) {
        X[q] = make_float2(0.f, 0.f);
        continue;
      }
      const int idx = tl + q * T;
      const float2 w = ROUNDS > 1 ? W[q] : cscale(ldsP[swz(idx)], S);
      const float fky = (float)wn(idx, N);
      const float k2 = fkx * fkx + fky * fky;
      const float inv = k2 > 0.f ? __builtin_amdgcn_rcpf(k2) : 0.f;
      const float gg = (cy * fky + cx * fkx) * inv;
      X[q] = f == 0 ? w : cscale(cmuli(w), gg);
    }
    if constexpr (ROUNDS == 1) __syncthreads();  // every group has read the new state
    line_fft<N, +1>(X, lds, tl, a.tw);
    if (f <= 2) {
#pragma unroll
      for (int q = 0; q < PTS; ++q) st2(bu, X[q], vm, (uint32_t)f * PLANE_B + q * ROWS_B);
    }
  }

  if (a.stage == 2) {
    // per-step epilogue by the last workgroup of the instance to finish (as in k_cols)
    __syncthreads();
    if (tid < 64) {
      unsigned last = 0u;
      if (tid == 0) last = atomicAdd(&I->st.cols_done, 1u) == (unsigned)(KC - 1) ? 1u : 0u;
      last = __shfl(last, 0);
      if (last) {
        if (tid < DevState::SPEED_SLOTS)
          __hip_atomic_store(&I->st.speed_bits[tid], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (tid == 0) {
          step_epilogue(I->st);
          __hip_atomic_store(&I->st.cols_done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------------------------
// full-field helpers: plain 2-D FFT (Fft2D.run, fft.js:81-92), state pack / unpack
// ---------------------------------------------------------------------------------------------
template <int N, int G, int DIR>
__global__ void __launch_bounds__(G * Line<N>::T) k_fft_rows(float2* data, const float2* tw,
                                                             float scale) {
  using L = Line<N>;
  __shared__ float2 lds_all[G * L::LDS];
  const int g = threadIdx.x / L::T, t = threadIdx.x - g * L::T;
  const int row = blockIdx.x * G + g;
  const bool live = row < N;
  float2* lds = lds_all + g * L::LDS;
  float2 v[L::PTS];
#pragma unroll
  for (int q = 0; q < L::PTS; ++q)
    v[q] = live ? data[(size_t)row * N + t + q * L::T] : make_float2(0.f, 0.f);
  line_fft<N, DIR>(v, lds, t, tw);
  if (live) {
#pragma unroll
    for (int q = 0; q < L::PTS; ++q) data[(size_t)row * N + t + q * L::T] = cscale(v[q], scale);
  }
}

template <int N, int C, int DIR>
__global__ void __launch_bounds__(C * Line<N>::T) k_fft_cols(float2* data, const float2* tw,
                                                             float scale) {
  using L = Line<N>;
  constexpr int LSTRIDE = ColCfg<N>::LSTRIDE;
  __shared__ float2 lds_all[C * LSTRIDE];
  const int c = ColCfg<N>::template col<C>(threadIdx.x);
  const int t = ColCfg<N>::template row<C>(threadIdx.x);
  const int colx = blockIdx.x * C + c;
  const bool live = colx < N;
  float2* lds = lds_all + c * LSTRIDE;
  float2 v[L::PTS];
#pragma unroll
  for (int q = 0; q < L::PTS; ++q)
    v[q] = live ? data[(size_t)(t + q * L::T) * N + colx] : make_float2(0.f, 0.f);
  line_fft<N, DIR>(v, lds, t, tw);
  if (live) {
#pragma unroll
    for (int q = 0; q < L::PTS; ++q) data[(size_t)(t + q * L::T) * N + colx] = cscale(v[q], scale);
  }
}

// full N x N spectrum -> compact masked half spectrum, k = 0 zeroed (seed.wgsl:4-17)
template <int N>
__global__ void k_pack(const float2* __restrict__ full, float2* __restrict__ st) {
  using Ge = Geo<N>;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Ge::STATE) return;
  const int kx = (int)(i / Ge::PITCH), j = (int)(i % Ge::PITCH);
  float2 v = make_float2(0.f, 0.f);
  if (j < Ge::NJ && !(kx == 0 && j == 0)) {
    const int row = j <= Ge::K ? j : j + (N - Ge::NJ);
    v = full[(size_t)row * N + kx];
  }
  st[i] = v;
}

// compact half spectrum -> full N x N spectrum (Hermitian extension, zero out of band)
template <int N>
__global__ void k_unpack(const float2* __restrict__ st, float2* __restrict__ full) {
  using Ge = Geo<N>;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)N * N) return;
  const int row = (int)(i / N), colx = (int)(i % N);
  float2 v = make_float2(0.f, 0.f);
  if (colx <= Ge::K) {
    const int j = Ge::jrow(row);
    if (j >= 0) v = st[(size_t)colx * Ge::PITCH + j];
  } else if (colx >= N - Ge::K) {
    const int j = Ge::jrow((N - row) % N);
    if (j >= 0) v = cconj(st[(size_t)(N - colx) * Ge::PITCH + j]);
  }
  full[i] = v;
}

// ---------------------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------------------
template <int N>
struct Impl {
  using L = Line<N>;
  using Ge = Geo<N>;
  static constexpr int G = RowCfg<N>::G;
  static constexpr int C = ColCfg<N>::C;

  // the twiddle tables of Line / TwBase, FP64 angles rounded once
  template <int R>
  static void put_pass(float2* h, int off, int ns) {
    for (int j = 0; j < TwBase<R>::NB; ++j)
      for (int k = 0; k < ns; ++k) {
        const long period = (long)ns * R;
        const long m = ((long)TwBase<R>::exponent(j) * k) % period;
        const double ang = -2.0 * M_PI * (double)m / (double)period;
        h[off + j * ns + k] = make_float2((float)std::cos(ang), (float)std::sin(ang));
      }
  }
  static void twiddles(float2* h) {
    put_pass<L::R1>(h, L::TW1, L::R0);
    if constexpr (L::NPASS == 3) put_pass<L::R2>(h, L::TW2, L::R0 * L::R1);
  }

  // Returns true when the products were written in the sum / difference form (S', D'): the
  // stage-0 kinds, which read the omega', s', d' planes. vel and force apply at stage 0 only.
  static void launch_rows_fp(Solver* s, const RowArgs& a, bool sd, bool vel, bool force) {
    if constexpr (RowFp<N>::on) {
      constexpr int FG = (N / 2 + 7) / 8 * 8;
      const dim3 grid(FG, s->batch), block(ROWFP_GROUPS * L::T);
      if (sd && force)
        hipLaunchKernelGGL((k_rows_fp<N, ROWS_SINGLE_FORCE>), grid, block, 0, s->stream, a);
      else if (sd && vel)
        hipLaunchKernelGGL((k_rows_fp<N, ROWS_SINGLE>), grid, block, 0, s->stream, a);
      else if (sd)
        hipLaunchKernelGGL((k_rows_fp<N, ROWS_PAIRED_SD>), grid, block, 0, s->stream, a);
      else
        hipLaunchKernelGGL((k_rows_fp<N, ROWS_PAIRED>), grid, block, 0, s->stream, a);
    }
  }

  static bool rows(Solver* s, int stage, bool speed, bool vel, bool force) {
    RowArgs a{};
    a.mu = s->mixed_u;
    a.mz = s->mixed_z;
    a.mf = s->mixed_f;
    a.vel = s->vel;
    a.tw = s->tw;
    a.inst = s->inst;
    a.s_mu = 3 * Ge::PLANE;
    a.s_mz = 2 * Ge::PLANE;
    a.s_mf = Ge::PLANE;
    a.s_vel = (size_t)N * N;
    a.splats = s->splats;
    a.stage = stage;
    a.flags = (speed ? ROW_SPEED : 0) | (vel ? ROW_VEL : 0) | (force ? ROW_FORCE : 0);
    const int grid = (RowCfg<N>::NB + 7) / 8 * 8;
    prof_begin(s, S2T_PROF_ROWS);
    const bool sd = stage == 0;
    if (!sd && (speed || vel || force))
      throw std::logic_error("row kernel: speed, dye velocity and pointer forcing are stage-0 work");
    if (RowFp<N>::on && s->rows_fp) {
      launch_rows_fp(s, a, sd, vel, force);
    } else if (sd && force) {
      hipLaunchKernelGGL((k_rows<N, G, ROWS_SINGLE_FORCE>), dim3(grid, s->batch), dim3(G * L::T), 0,
                         s->stream, a);
    } else if (sd && vel) {
      hipLaunchKernelGGL((k_rows<N, G, ROWS_SINGLE>), dim3(grid, s->batch), dim3(G * L::T), 0,
                         s->stream, a);
    } else if (sd) {
      hipLaunchKernelGGL((k_rows<N, G, ROWS_PAIRED_SD>), dim3(grid, s->batch), dim3(G * L::T), 0,
                         s->stream, a);
    } else {
      hipLaunchKernelGGL((k_rows<N, G, ROWS_PAIRED>), dim3(grid, s->batch), dim3(G * L::T), 0,
                         s->stream, a);
    }
    check_launch("k_rows");
    prof_end(s);
    return sd;
  }

  static ColArgs col_args(Solver* s, int stage) {
    ColArgs a{};
    a.mz = s->mixed_z;
    a.mf = s->mixed_f;
    a.mu = s->mixed_u;
    const int p = s->phase;
    // 3-buffer rotation of src/gpu/solver.js:193-212
    a.w0 = s->state[p];
    a.wcur = s->state[(p + stage) % 3];
    a.wout = s->state[(p + (stage == 1 ? 2 : 1)) % 3];
    a.fk = s->force_k;
    a.tw = s->tw;
    a.inst = s->inst;
    a.s_mz = 2 * Ge::PLANE;
    a.s_mf = Ge::PLANE;
    a.s_mu = 3 * Ge::PLANE;
    a.s_state = Ge::STATE;
    a.stage = stage;
    return a;
  }

  template <int CW>
  static void launch_cols(Solver* s, const ColArgs& a, bool force) {
    constexpr int GROUPS = (Ge::KC + CW - 1) / CW;
    constexpr int G8 = (GROUPS + 7) / 8 * 8;
    const dim3 grid(G8, s->batch);
    if (force)
      hipLaunchKernelGGL((k_cols<N, CW, COLS_STEP, true>), grid, dim3(CW * L::T), 0, s->stream, a);
    else
      hipLaunchKernelGGL((k_cols<N, CW, COLS_STEP, false>), grid, dim3(CW * L::T), 0, s->stream,
                         a);
  }
  static void launch_cols_fp(Solver* s, const ColArgs& a, bool force) {
    if constexpr (ColFp<N>::on) {
      constexpr int G8 = (Ge::KC + 7) / 8 * 8;
      constexpr int GR = ColFp<N>::groups;
      const dim3 grid(G8, s->batch), block(GR * L::T);
      if (force)
        hipLaunchKernelGGL((k_cols_fp<N, GR, true>), grid, block, 0, s->stream, a);
      else
        hipLaunchKernelGGL((k_cols_fp<N, GR, false>), grid, block, 0, s->stream, a);
    }
  }
  template <int I = 0>
  static void launch_cols_width(Solver* s, const ColArgs& a, bool force, int w) {
    if (ColFp<N>::on && w == ColFp<N>::width) return launch_cols_fp(s, a, force);
    if constexpr (I < ColWidths<N>::n) {
      constexpr int CW = ColWidths<N>::w[I];
      if (w == CW) return launch_cols<CW>(s, a, force);
      return launch_cols_width<I + 1>(s, a, force, w);
    } else {
      throw Fail(S2T_EINVAL, "column-group width " + std::to_string(w) +
                                 " is not compiled for this grid size");
    }
  }

  static void cols(Solver* s, int stage, bool force, bool sd) {
    ColArgs a = col_args(s, stage);
    a.force = force ? 1 : 0;
    a.sd = sd ? 1 : 0;
    prof_begin(s, S2T_PROF_COLS);
    int w0[MAX_COL_WIDTHS];
    col_widths(w0);
    launch_cols_width(s, a, force, s->cols_c != 0 ? s->cols_c : w0[0]);
    check_launch("k_cols");
    prof_end(s);
  }

  // the compiled widths, the default first: the field-parallel kernel where it is compiled,
  // then the column-group widths
  static int col_widths(int* out) {
    int k = 0;
    if (ColFp<N>::on) out[k++] = ColFp<N>::width;
    for (int i = 0; i < ColWidths<N>::n; ++i) out[k++] = ColWidths<N>::w[i];
    return k;
  }

  // velocity prime (encodePrime, solver.js:475-486) of instances i0 .. i0+ni-1 from state[phase]
  static void prime(Solver* s, int i0, int ni) {
    ColArgs a = col_args(s, 0);
    a.wcur = s->state[s->phase] + (size_t)i0 * Ge::STATE;
    a.mu = s->mixed_u + (size_t)i0 * 3 * Ge::PLANE;
    a.inst = s->inst + i0;
    constexpr int GROUPS = (Ge::KC + C - 1) / C;
    constexpr int G8 = (GROUPS + 7) / 8 * 8;
    prof_begin(s, S2T_PROF_OTHER);
    hipLaunchKernelGGL((k_cols<N, C, COLS_PRIME, false>), dim3(G8, ni), dim3(C * L::T), 0,
                       s->stream, a);
    check_launch("k_cols(prime)");
    prof_end(s);
  }

  // fused right-hand side of instance 0 with an explicit step index and dt (test hook)
  static void rhs(Solver* s, bool force, bool sd, float2* rhs_out, uint32_t step, double dt) {
    ColArgs a = col_args(s, 0);
    a.rhs_out = rhs_out;
    a.force = force ? 1 : 0;
    a.sd = sd ? 1 : 0;
    a.step_override = step;
    a.dt_override = dt;
    constexpr int GROUPS = (Ge::KC + C - 1) / C;
    constexpr int G8 = (GROUPS + 7) / 8 * 8;
    if (force)
      hipLaunchKernelGGL((k_cols<N, C, COLS_RHS, true>), dim3(G8, 1), dim3(C * L::T), 0,
                         s->stream, a);
    else
      hipLaunchKernelGGL((k_cols<N, C, COLS_RHS, false>), dim3(G8, 1), dim3(C * L::T), 0,
                         s->stream, a);
    check_launch("k_cols(rhs)");
  }

  static void fft2(Solver* s, float2* data, int inverse) {
    const float sc = inverse ? 1.0f / (float)N : 1.0f;
    const int grow = (N + G - 1) / G;
    const int gcol = (N + C - 1) / C;
    if (inverse) {
      hipLaunchKernelGGL((k_fft_rows<N, G, +1>), dim3(grow), dim3(G * L::T), 0, s->stream, data,
                         s->tw, sc);
      check_launch("k_fft_rows");
      hipLaunchKernelGGL((k_fft_cols<N, C, +1>), dim3(gcol), dim3(C * L::T), 0, s->stream, data,
                         s->tw, sc);
      check_launch("k_fft_cols");
    } else {
      hipLaunchKernelGGL((k_fft_rows<N, G, -1>), dim3(grow), dim3(G * L::T), 0, s->stream, data,
                         s->tw, sc);
      check_launch("k_fft_rows");
      hipLaunchKernelGGL((k_fft_cols<N, C, -1>), dim3(gcol), dim3(C * L::T), 0, s->stream, data,
                         s->tw, sc);
      check_launch("k_fft_cols");
    }
  }

  static void pack(Solver* s, const float2* full, float2* st) {
    const size_t n = Ge::STATE;
    hipLaunchKernelGGL((k_pack<N>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s->stream,
                       full, st);
    check_launch("k_pack");
  }

  static void unpack(Solver* s, const float2* st, float2* full) {
    const size_t n = (size_t)N * N;
    hipLaunchKernelGGL((k_unpack<N>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s->stream,
                       st, full);
    check_launch("k_unpack");
  }

  static const Ops* ops() {
    static const Ops o = {Ge::STATE, Ge::PLANE, L::TWSIZE, &twiddles, &rows,
                          &cols,     &prime,    &rhs,      &fft2,     &pack,
                          &unpack,   &col_widths};
    return &o;
  }
};

// ---------------------------------------------------------------------------------------------
// grouped launches: one launch per pass over the instances of several grid sizes
// ---------------------------------------------------------------------------------------------
// The workgroups of a launch form a list of (size class, instance, block) items: class c owns
// workgroups first[c] .. first[c] + nb[c] * instances - 1, instance-major, and each workgroup runs
// the row or column body of its class's grid size with 256 threads (several lines or columns of
// the smaller sizes per workgroup). Every class's range starts at a multiple of 8, so the XCD of a
// workgroup is that of its block index within the instance, as in the per-class launches.
constexpr int GROUP_THREADS = 256;
constexpr int GROUP_MAX = 4;

template <int N>
struct GroupCfg {
  static constexpr int G = GROUP_THREADS / Line<N>::T;  // row blocks per row workgroup
  static constexpr int C = GROUP_THREADS / Line<N>::T;  // columns per column workgroup
  static constexpr int ROW_NB8 = ((N / 2 + G - 1) / G + 7) / 8 * 8;
  static constexpr int COL_NB8 = ((Geo<N>::KC + C - 1) / C + 7) / 8 * 8;
};
static_assert(GroupCfg<256>::C == 8 && GroupCfg<512>::C == 4 && GroupCfg<1024>::C == 4,
              "grouped column workgroups use compiled column-group widths");

struct GroupClass {
  float2* mu;                      // omega', u', v' tiles
  float2* mz;                      // P', Q' tiles
  float2* w0, *wcur, *wout;        // states of the stage (column launches)
  const float2* tw;
  Inst* inst;
  int n, first, nb;  // grid size; first workgroup, workgroups per instance in this launch
};
struct GroupArgs {
  GroupClass c[GROUP_MAX];
  int count, stage;
};

union RowGroupSmem {
  RowSmem<256, GroupCfg<256>::G> n256;
  RowSmem<512, GroupCfg<512>::G> n512;
  RowSmem<1024, GroupCfg<1024>::G> n1024;
};
union ColGroupSmem {
  ColSmem<256, GroupCfg<256>::C> n256;
  ColSmem<512, GroupCfg<512>::C> n512;
  ColSmem<1024, GroupCfg<1024>::C> n1024;
};

// the class of workgroup b, and its block and instance within the class
__device__ __forceinline__ const GroupClass& group_item(const GroupArgs& g, int& bx, int& iy) {
  const int b = (int)blockIdx.x;
  int c = 0;
  while (c + 1 < g.count && b >= g.c[c + 1].first) ++c;
  const GroupClass& k = g.c[c];
  const int local = b - k.first;
  iy = local / k.nb;
  bx = local - iy * k.nb;
  return k;
}

template <int N, int KIND>
__device__ __forceinline__ void group_rows(const GroupClass& k, int stage, int bx, int iy,
                                           RowSmem<N, GroupCfg<N>::G>& sm) {
  RowArgs a{};
  a.mu = k.mu;
  a.mz = k.mz;
  a.tw = k.tw;
  a.inst = k.inst;
  a.s_mu = 3 * Geo<N>::PLANE;
  a.s_mz = 2 * Geo<N>::PLANE;
  a.stage = stage;
  a.flags = KIND == ROWS_PAIRED_SD ? ROW_SPEED : 0;
  rows_body<N, GroupCfg<N>::G, KIND>(a, bx, iy, sm);
}

template <int KIND>
__global__ void __launch_bounds__(GROUP_THREADS) k_rows_group(GroupArgs g) {
  __shared__ RowGroupSmem sm;
  int bx, iy;
  const GroupClass& k = group_item(g, bx, iy);
  switch (k.n) {
    case 256: group_rows<256, KIND>(k, g.stage, bx, iy, sm.n256); break;
    case 512: group_rows<512, KIND>(k, g.stage, bx, iy, sm.n512); break;
    case 1024: group_rows<1024, KIND>(k, g.stage, bx, iy, sm.n1024); break;
  }
}

template <int N>
__device__ __forceinline__ void group_cols(const GroupClass& k, int stage, int bx, int iy,
                                           ColSmem<N, GroupCfg<N>::C>& sm) {
  ColArgs a{};
  a.mz = k.mz;
  a.mu = k.mu;
  a.w0 = k.w0;
  a.wcur = k.wcur;
  a.wout = k.wout;
  a.tw = k.tw;
  a.inst = k.inst;
  a.s_mz = 2 * Geo<N>::PLANE;
  a.s_mu = 3 * Geo<N>::PLANE;
  a.s_state = Geo<N>::STATE;
  a.stage = stage;
  a.sd = stage == 0 ? 1 : 0;  // the stage-0 row launch writes the sum / difference form
  cols_body<N, GroupCfg<N>::C, COLS_STEP, false>(a, bx, iy, sm);
}

// four waves per SIMD (the LDS allows four workgroups per CU; unbound the compiler takes 141 VGPRs)
__global__ void __launch_bounds__(GROUP_THREADS, 4) k_cols_group(GroupArgs g) {
  __shared__ ColGroupSmem sm;
  int bx, iy;
  const GroupClass& k = group_item(g, bx, iy);
  switch (k.n) {
    case 256: group_cols<256>(k, g.stage, bx, iy, sm.n256); break;
    case 512: group_cols<512>(k, g.stage, bx, iy, sm.n512); break;
    case 1024: group_cols<1024>(k, g.stage, bx, iy, sm.n1024); break;
  }
}

bool group_size(int n) { return n == 256 || n == 512 || n == 1024; }

// One step of the classes (each of a grouped size, with no pointer forcing and no dye, all at the
// same buffer phase and primed) in six launches on `stream`.
void group_step(Solver* const* cls, int count, hipStream_t stream) {
  if (count < 1 || count > GROUP_MAX) throw Fail(S2T_EINVAL, "grouped step: 1 to 4 size classes");
  auto nb8 = [](int n, bool rows) {
    switch (n) {
      case 256: return rows ? GroupCfg<256>::ROW_NB8 : GroupCfg<256>::COL_NB8;
      case 512: return rows ? GroupCfg<512>::ROW_NB8 : GroupCfg<512>::COL_NB8;
      case 1024: return rows ? GroupCfg<1024>::ROW_NB8 : GroupCfg<1024>::COL_NB8;
    }
    throw Fail(S2T_EINVAL, "grouped step: unsupported grid size " + std::to_string(n));
  };
  for (int stage = 0; stage < 3; ++stage) {
    for (int rows = 1; rows >= 0; --rows) {
      GroupArgs g{};
      g.count = count;
      g.stage = stage;
      int total = 0;
      for (int c = 0; c < count; ++c) {
        const Solver* s = cls[c];
        if (s->splats.count > 0) throw Fail(S2T_EINVAL, "grouped step: no pointer forcing");
        GroupClass& k = g.c[c];
        const int p = s->phase;
        k.mu = s->mixed_u;
        k.mz = s->mixed_z;
        k.w0 = s->state[p];
        k.wcur = s->state[(p + stage) % 3];
        k.wout = s->state[(p + (stage == 1 ? 2 : 1)) % 3];
        k.tw = s->tw;
        k.inst = s->inst;
        k.n = s->n;
        k.nb = nb8(s->n, rows != 0);
        k.first = total;
        total += k.nb * s->batch;
      }
      if (rows) {
        if (stage == 0)
          hipLaunchKernelGGL((k_rows_group<ROWS_PAIRED_SD>), dim3(total), dim3(GROUP_THREADS), 0,
                             stream, g);
        else
          hipLaunchKernelGGL((k_rows_group<ROWS_PAIRED>), dim3(total), dim3(GROUP_THREADS), 0,
                             stream, g);
        check_launch("k_rows_group");
      } else {
        hipLaunchKernelGGL(k_cols_group, dim3(total), dim3(GROUP_THREADS), 0, stream, g);
        check_launch("k_cols_group");
      }
    }
  }
  for (int c = 0; c < count; ++c) cls[c]->phase = (cls[c]->phase + 1) % 3;
}

const Ops* ops_for(int n) {
  switch (n) {
    case 16: return Impl<16>::ops();
    case 32: return Impl<32>::ops();
    case 64: return Impl<64>::ops();
    case 128: return Impl<128>::ops();
    case 256: return Impl<256>::ops();
    case 512: return Impl<512>::ops();
    case 1024: return Impl<1024>::ops();
    case 2048: return Impl<2048>::ops();
    case 4096: return Impl<4096>::ops();
  }
  return nullptr;
}

}  // namespace s2t
