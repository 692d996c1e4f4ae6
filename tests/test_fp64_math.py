"""The FP64 exp / sincos of the integrating factor (csrc/physics.hpp) against libm.

The device functions are __host__ __device__; this builds a host-only program from the same
header with hipcc and compares every result with the C library to a few double ulp, so the FP32
factor rounded from them is the correctly rounded value.
"""

import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

SRC = r"""
#include <cstdio>
#include <cmath>
#include "physics.hpp"
int main() {
  double worst_e = 0, worst_s = 0, worst_c = 0;
  for (int i = 0; i <= 200000; ++i) {
    double y = -740.0 + 1449.0 * i / 200000.0;
    double a = s2t::exp_f64(y), b = std::exp(y);
    if (b > 1e-300) worst_e = std::fmax(worst_e, std::fabs(a - b) / b);
    double x = -2.0e4 + 4.0e4 * i / 200000.0 + 1e-7 * i;
    double s, c;
    s2t::sincos_f64(x, &s, &c);
    worst_s = std::fmax(worst_s, std::fabs(s - std::sin(x)));
    worst_c = std::fmax(worst_c, std::fabs(c - std::cos(x)));
  }
  for (int i = 0; i <= 100000; ++i) {  // small arguments, relative error of sin
    double x = 1e-8 + 0.8 * i / 100000.0;
    double s, c;
    s2t::sincos_f64(x, &s, &c);
    worst_s = std::fmax(worst_s, std::fabs(s - std::sin(x)) / std::sin(x));
  }
  std::printf("%.3e %.3e %.3e\n", worst_e, worst_s, worst_c);
  return 0;
}
"""


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc missing")
def test_exp_sincos_f64(tmp_path):
    src = tmp_path / "fp64.cpp"
    src.write_text(SRC)
    exe = tmp_path / "fp64"
    inc = os.path.join(ROOT, "spectral-2d-turbulence_amd", "csrc")
    r = subprocess.run([HIPCC, "-O2", "-std=c++17", "-x", "hip", "--offload-arch=gfx950",
                        "-I" + inc, str(src), "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    e, s, c = map(float, out.stdout.split())
    assert e < 1e-15 and s < 1e-15 and c < 1e-15, out.stdout
