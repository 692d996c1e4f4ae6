/*
 * Known answers for the hash compositions of the reference's GPU forcing and its spectrum seeding,
 * computed with plain C uint32 arithmetic, independently of the NumPy oracle. The printed numbers
 * are the constants FORCE_KAT and FILL_KAT of tests/test_oracle_pins.py, which also rebuilds this
 * file with gcc and checks that they still agree.
 *
 *   hash1(x): src/shaders/common.wgsl:76-80 (hash32 of src/gpu/solver.js:43-47 is the same).
 *   stochastic forcing phase: 2 pi hashf(hr*N + hc, stepIndex), hashf(a, b) =
 *     f32(hash1(a ^ hash1(b))) * 2^-32, in f32 (solver.wgsl:87, common.wgsl:82-84); the mode's
 *     canonical index is its own when ky > 0, or ky = 0 and kx > 0, else its mirror's
 *     (solver.wgsl:80-86).
 *   fillSpectrum phase: 2 pi hash32((hr*n + hc) ^ hash32(seed)) / 2^32 in FP64
 *     (src/gpu/solver.js:387, 404).
 *
 * Build and run: gcc -O0 -o hash_kat hash_kat.c -lm && ./hash_kat
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>

static uint32_t hash1(uint32_t x) {
  uint32_t v = x * 747796405u + 2891336453u;
  v = ((v >> ((v >> 28u) + 4u)) ^ v) * 277803737u;
  return (v >> 22u) ^ v;
}

static int wn(int i, int n) { return i < n / 2 ? i : i - n; }

int main(void) {
  /* forcing: N = 64, stepIndex = 3; modes as (row, col) */
  {
    const int n = 64;
    const uint32_t step = 3;
    const int modes[][2] = {{2, 9}, {2, 55}, {62, 9}, {0, 10}, {0, 54}, {5, 61}};
    printf("FORCE_KAT = [  # (row, col, sign, f32 phase): N = 64, stepIndex = 3\n");
    for (unsigned i = 0; i < sizeof modes / sizeof modes[0]; ++i) {
      const int row = modes[i][0], col = modes[i][1];
      const int kx = wn(col, n), ky = wn(row, n);
      int hr = row, hc = col, sgn = 1;
      if (!(ky > 0 || (ky == 0 && kx > 0))) {
        hr = (n - row) % n;
        hc = (n - col) % n;
        sgn = -1;
      }
      const uint32_t h = hash1((uint32_t)(hr * n + hc) ^ hash1(step));
      const float u = (float)h * 2.3283064365386963e-10f;
      const float ph = 6.28318530717958647f * u;
      printf("    (%d, %d, %d, %.9g),\n", row, col, sgn, (double)ph);
    }
    printf("]\n");
  }
  /* fillSpectrum: n = 32, seed = 5 */
  {
    const int n = 32;
    const uint32_t salt = hash1(5u);
    const int modes[][2] = {{3, 4}, {3, 28}, {29, 4}, {1, 2}, {0, 7}, {0, 25}};
    printf("FILL_KAT = [  # (row, col, sign, FP64 phase): n = 32, seed = 5\n");
    for (unsigned i = 0; i < sizeof modes / sizeof modes[0]; ++i) {
      const int row = modes[i][0], col = modes[i][1];
      const int kx = wn(col, n), ky = wn(row, n);
      int hr = row, hc = col, sgn = 1;
      if (!(ky > 0 || (ky == 0 && kx > 0))) {
        hr = (n - row) % n;
        hc = (n - col) % n;
        sgn = -1;
      }
      const uint32_t h = hash1((uint32_t)(hr * n + hc) ^ salt);
      const double ph = 2.0 * M_PI * (double)h / 4294967296.0;
      printf("    (%d, %d, %d, %.17g),\n", row, col, sgn, ph);
    }
    printf("]\n");
  }
  return 0;
}
