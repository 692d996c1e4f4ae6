"""Errors of the FP32 device path against the FP64 oracle, stated per workload over many steps.

Each test runs one of the solver's workloads on the device through the C ABI and advances the FP64
oracle (oracle/spectral.py, the scheme of reference/spectral2d.py:73-90) with the same dt and step
index (the device's dt comes from the CFL reduction of src/gpu/reduce.wgsl:22-63; the oracle is
given that dt so that both integrate the same times). It prints and bounds:
* the relative error of the vorticity spectrum as max |error| over max |value| across the
  interleaved real and imaginary parts (the metric of the reference harness, tests/index.html:89-97)
* the relative errors of the energy and the enstrophy (spectral2d.py:92-97)
* for the forced beta-plane runs, the zonal-mean u profile (spectral2d.py:106-108) and the radial
  energy spectrum (spectral2d.py:98-104), each as max |error| over the largest |value|

Horizons follow the error statements of the workloads: 20 and 100 steps (3 and 20 at 4096^2);
the FP64 oracle runs on the host cores (its transforms and pointwise stages threaded). Pointwise
errors over the workloads' own thousands of chaotic steps are not meaningful (FP32 and FP64
trajectories part exponentially) and are not stated; the long-horizon statistics are checked in
tests/test_gpu_horizons.py.
"""

import numpy as np
import pytest

from conftest import rel_err
from oracle import gpusem as gs
from oracle.spectral import Scheme, nu_from_rate

pytestmark = pytest.mark.gpu



def oracle_for(params, n):
    return Scheme(n, beta=params["beta"], mu=params["drag"],
                  nu=nu_from_rate(params["viscosity"], n, params["hyper_p"]),
                  p=int(params["hyper_p"]), cfl=params["cfl"], dt_max=params["dt_max"])


def errors(d, o):
    """Device diagnostics (FP64 reductions of the FP32 state) vs the FP64 oracle's."""
    out = dict(energy=abs(d["energy"] - o.energy()) / o.energy(),
               enstrophy=abs(d["enstrophy"] - o.enstrophy()) / o.enstrophy())
    zon = o.zonal_mean_u()
    out["zonal"] = float(np.max(np.abs(d["zonal"] - zon)) / np.max(np.abs(zon)))
    spec = o.spectrum()
    out["spectrum"] = float(np.max(np.abs(d["spectrum"] - spec)) / np.max(spec))
    return out


def blob_dye(m, seed, count=12):
    """Seeded Gaussian colour blobs on the m x m dye grid (rgba16f, alpha 1, colours in [0, 1])."""
    rng = np.random.default_rng(seed)
    x = (np.arange(m) + 0.5) / m
    dye = np.zeros((m, m, 4))
    dye[..., 3] = 1.0
    for _ in range(count):
        cx, cy, r = rng.uniform(0, 1), rng.uniform(0, 1), rng.uniform(0.03, 0.1)
        dx = (x - cx + 0.5) % 1.0 - 0.5
        dy = (x - cy + 0.5) % 1.0 - 0.5
        g = np.exp(-(dy[:, None] ** 2 + dx[None, :] ** 2) / r ** 2)
        dye[..., :3] += g[..., None] * rng.uniform(0, 1, 3)
    dye[..., :3] = np.clip(dye[..., :3], 0.0, 1.0)
    return dye.astype(np.float16)


def oracle_steps(o, dts, forcing=None):
    for k, dt in enumerate(dts):
        o.step(dt=dt, forcing=None if forcing is None else forcing(k, dt))
        if (k + 1) % 10 == 0:
            print("  oracle step %d of %d" % (k + 1, len(dts)), flush=True)


def report(name, steps, err_w, errs):
    print("%s, %d steps: vorticity %.3e, energy %.3e, enstrophy %.3e, zonal u %.3e, "
          "spectrum %.3e" % (name, steps, err_w, errs["energy"], errs["enstrophy"],
                             errs["zonal"], errs["spectrum"]))


@pytest.mark.parametrize("steps", [20, 100])
def test_decaying_turbulence_1024_with_dye(s2t, steps):
    """Decaying turbulence at 1024^2: Gaussian amplitudes with uniform random phases band-limited
    to |k| in [8, 16], seed 2, omega_rms 6; hyperviscosity p = 4 with the damping rate at the
    cut-off equal to omega_rms (the dissipation scale at kmax); CFL 0.5 adaptive dt; seeded
    Gaussian dye blobs advected on the 2048^2 refined grid every step. Against FP64: vorticity,
    energy, enstrophy and E(k) within 1e-5 relative."""
    n = 1024
    params = dict(beta=0.0, drag=0.0, viscosity=6.0, hyper_p=4.0, force_amp=0.0, force_k=32.0,
                  force_width=2.0, cfl=0.5, dt_min=1e-5, dt_max=0.02, dye_fade=1.0, fixed_dt=0.0)
    o = oracle_for(params, n)
    o.w = gs.bandlimited_spectrum(n, 8.0, 16.0, 6.0, seed=2).astype(np.complex128)
    dts = []
    with s2t.Solver(n, dye_scale=2) as s:
        assert s.dye_size() == 2 * n
        s.configure(**params)
        s.seed_bandlimited(8.0, 16.0, 6.0, seed=2)
        s.write_dye(blob_dye(2 * n, seed=2))
        for _ in range(steps):
            s.step(1, dye=True)
            dts.append(s.diagnostics()["dt"])
        got = s.read_state()
        d = s.diagnostics()
        dye = s.read_dye().astype(np.float64)
    oracle_steps(o, dts)
    err_w = rel_err(got, o.w)
    e = errors(d, o)
    report("decaying 1024^2 + dye 2048^2", steps, err_w, e)
    assert d["nonfinite"] == 0
    assert err_w < 1e-5
    assert e["energy"] < 1e-5 and e["enstrophy"] < 1e-5 and e["spectrum"] < 1e-5
    # the dye stays a bounded colour field (BFECC limiter; pinned against the oracle elsewhere)
    assert np.all(np.isfinite(dye)) and dye.min() >= -1e-3 and dye.max() <= 1.0 + 1e-3
    assert np.all(dye[..., 3] == 1.0)


@pytest.mark.parametrize("n,steps", [(2048, 20), (4096, 3), (2048, 100), (4096, 20)])
def test_forced_beta_plane_from_rest(s2t, n, steps):
    """Forced beta-plane turbulence from rest: beta = 10, drag 0.01, hyperviscosity p = 4 at rate 1,
    stochastic forcing white in time at k_f = 32 +- 2 with the amplitude of the closed-form
    injection rate 1e-3, CFL 0.5 adaptive dt. Against FP64 after `steps` steps: vorticity,
    energy and enstrophy within 1e-5 relative, the zonal-mean u profile and the radial energy
    spectrum within 1e-5 of their largest values."""
    amp = gs.amp_for_rate(1e-3, n, 32.0, 2.0)
    params = dict(beta=10.0, drag=0.01, viscosity=1.0, hyper_p=4.0, force_amp=amp, force_k=32.0,
                  force_width=2.0, cfl=0.5, dt_min=1e-5, dt_max=0.02, dye_fade=1.0, fixed_dt=0.0)
    o = oracle_for(params, n)
    o.w = np.zeros((n, n), np.complex128)
    dts = []
    with s2t.Solver(n) as s:
        s.configure(**params)
        s.write_state(o.w)
        for _ in range(steps):
            s.step(1)
            dts.append(s.diagnostics()["dt"])
        got = s.read_state()
        d = s.diagnostics()
    oracle_steps(o, dts, lambda k, dt: gs.stochastic_forcing(n, amp, 32.0, 2.0, k, dt))
    err_w = rel_err(got, o.w)
    e = errors(d, o)
    report("forced beta-plane %d^2 from rest" % n, steps, err_w, e)
    assert d["nonfinite"] == 0
    assert err_w < 1e-5
    assert e["energy"] < 1e-5 and e["enstrophy"] < 1e-5
    assert e["zonal"] < 1e-5 and e["spectrum"] < 1e-5
