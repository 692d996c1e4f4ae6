"""CPU checks of bench.py's contract: one GPU only, the forced-turbulence workload's forcing
amplitude, and the roofline traffic read from the committed counter summary."""

import importlib.util
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import gpusem as gs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
bench = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(bench)


def test_more_than_one_gpu_is_refused():
    proc = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                          stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert proc.returncode == 2
    assert "one GPU" in proc.stderr
    assert proc.stdout == ""


@pytest.mark.parametrize("n", [256, 4096])
def test_forcing_amplitude_gives_the_stated_injection_rate(n):
    """The workload's forceAmp is the closed-form inverse of the mean injection rate
    eps = 1/2 (0.01 A)^2 sum env^2 / k^2 at eps = 1e-3, k_f = 32 +- 2, as the oracle states it."""
    p = bench.workload_params(n)
    assert p["force_k"] == 32.0 and p["force_width"] == 2.0
    eps = gs.injection_rate(n, p["force_amp"], 32.0, 2.0)
    assert abs(eps / 1e-3 - 1.0) < 1e-12
    assert abs(p["force_amp"] / gs.amp_for_rate(1e-3, n, 32.0, 2.0) - 1.0) < 1e-12


def test_roofline_traffic_comes_from_the_committed_profile():
    cols = bench.measured_traffic("k_cols", 4096)
    rows = bench.measured_traffic("k_rows", 4096)
    # HBM bytes per launch: at or above the algorithmic bytes (303.4 and 223.8 MB) and well
    # below the reference's unfused traffic
    assert cols is not None and 3.0e8 < cols < 6.0e8
    assert rows is not None and 2.2e8 < rows < 4.0e8
    assert bench.measured_traffic("k_cols", 12345) is None
