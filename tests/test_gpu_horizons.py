"""The decaying and forced workloads at their own horizons, checked by what the physics fixes.

Pointwise comparison with the FP64 oracle means nothing over thousands of chaotic steps (FP32 and
FP64 trajectories part exponentially; the short-horizon pointwise errors are in
tests/test_gpu_config_errors.py). These runs go the full length on the device and check:

* decaying turbulence at 1024^2 with dye on 2048^2, 2000 CFL steps: enstrophy never grows beyond
  FP32 rounding of the state and falls by at least 5% (reference/spectral2d.py:214-216), energy is
  conserved to the hyperviscous loss, the energy spectrum's slope over the inertial range lies in
  the reference's band (-4.2, -2.2) (spectral2d.py:204-213), and the dye stays a bounded colour
  field (the limiter's max principle);
* forced beta-plane turbulence from rest at 2048^2 and 4096^2, 5000 CFL steps: the energy budget
  closes, E(T) = eps T - integral of (2 mu E + 2 nu_h sum k^(2p) E(k)) dt with eps = 1e-3 the
  closed-form injection rate of the forcing amplitude (beta and the nonlinear term conserve
  energy), and the enstrophy-cascade range of E(k) has a slope in the same band.

The workloads are bench.py's (`--workload horizons` runs the same functions and writes the full
series, the final zonal-mean u(y) and E(k)).
"""

import importlib.util
import os

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

_spec = importlib.util.spec_from_file_location("s2t_bench", os.path.join(ROOT, "bench.py"))
bench = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(bench)


def test_decaying_1024_to_2000_steps(s2t):
    r = bench.run_decaying_horizon(s2t, 2000)
    z = np.asarray(r["series"]["enstrophy"])
    e = np.asarray(r["series"]["energy"])
    print("decaying 1024^2, 2000 steps: t %.3f, enstrophy decay %.4f, energy decay %.2e, slope "
          "%.2f, dye [%.4f, %.4f]" % (r["time"], r["enstrophy_decay"], r["energy_decay"],
                                      r["spectrum_slope_k32_k160"], r["dye_min"], r["dye_max"]))
    # the nonlinear term conserves Z and the hyperviscosity removes it; the state is FP32, so a
    # sample may exceed the previous one by its rounding (about 1e-7 relative), never more
    assert np.all(z[1:] <= z[:-1] * (1 + 1e-6))
    assert r["enstrophy_decay"] >= 0.05
    assert np.all(e[1:] <= e[:-1] * (1 + 1e-6)) and 0 <= r["energy_decay"] < 1e-3
    assert -4.2 < r["spectrum_slope_k32_k160"] < -2.2
    assert r["dye_min"] >= -1e-3 and r["dye_max"] <= 1.0 + 1e-3 and r["dye_alpha_one"]


@pytest.mark.parametrize("n", [2048, 4096])
def test_forced_beta_plane_to_5000_steps(s2t, n):
    r = bench.run_forced_horizon(s2t, n, 5000)
    b = r["budget"]
    print("forced %d^2, 5000 steps: t %.3f, E %.6e, injected %.6e, dissipated %.6e, residual "
          "%.2e of the injection, slope %.2f, <u^2>/<v^2> %.2f"
          % (n, r["time"], b["energy"], b["injected"], b["dissipated"],
             b["residual_over_injected"], r["spectrum_slope_k64_k256"], r["anisotropy_u2_v2"]))
    assert np.all(np.isfinite(r["series"]["energy"]))
    # the random work of the forcing against the flow makes the injection fluctuate about eps T
    # (0.04% and 0.1% of it in these runs, well inside 2%); a forcing amplitude off by sqrt 2
    # doubles the injection, and an injection that scaled with dt would follow the CFL step
    assert abs(b["residual_over_injected"]) < 0.02
    assert -4.2 < r["spectrum_slope_k64_k256"] < -2.2
    assert len(r["zonal_mean_u"]) == n and len(r["energy_spectrum"]) == n // 2 + 1
