"""The batched-ensemble and serving workloads at their stated shapes, on the device through the C
ABI, checked member by member against the plain solver (bit for bit) and sampled members against
the FP64 oracle (oracle/spectral.py; oracle/gpusem.py for the forcing, the seeding and the
scenarios).

* Ensemble: 64 members, 32 x 256^2 + 24 x 512^2 + 8 x 1024^2 (16,777,216 points), Laplacian
  viscosity nu log-uniform in [1e-5, 1e-3] (p = 1), beta ~ U[0, 20], stochastic forcing at
  k_f in {16, 32} injecting 1e-3, CFL 0.5; members stepped together by the batch path (one launch
  per kernel and size class, a hipGraph per step).
* Serving: a Poisson burst of requests over N in {128, 256, 512, 1024, 2048}^2 and the four
  scenarios (Taylor-Green, dipole, merger, seeded turbulence; src/scenarios.js:82-145,
  src/gpu/solver.js:384-418) into the continuous-batching server, with CFL steps.

Each error is max |error| over max |value| across the interleaved real and imaginary parts
(tests/index.html:89-97); the bound is 1e-5.
"""

import importlib.util
import os
import time

import numpy as np
import pytest

from conftest import ROOT, rel_err
from oracle import gpusem as gs
from oracle.spectral import Scheme, nu_from_rate

pytestmark = pytest.mark.gpu

_spec = importlib.util.spec_from_file_location("s2t_bench", os.path.join(ROOT, "bench.py"))
bench = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(bench)


def _oracle(params, n):
    return Scheme(n, beta=params["beta"], mu=params["drag"],
                  nu=nu_from_rate(params["viscosity"], n, int(params["hyper_p"])),
                  p=int(params["hyper_p"]), cfl=params["cfl"], dt_max=params["dt_max"])


def test_ensemble_64_members(s2t):
    """Every member of the 64-member mix equals its plain solver bit for bit after 10 graph-replayed
    steps; one member of each size class follows the FP64 oracle (given the member's device dt
    per step) within 1e-5."""
    sizes = bench.ENSEMBLE_SIZES
    assert len(sizes) == 64 and sum(n * n for n in sizes) == 4096 * 4096
    steps = 10
    sample = [0, 32, 56]  # one 256^2, one 512^2, one 1024^2 member
    params = [bench.ensemble_params(i, n) for i, n in enumerate(sizes)]
    with s2t.Batch(sizes) as b:
        b.use_graphs(True)
        for i, n in enumerate(sizes):
            b.configure(i, **params[i])
            b.seed_spectrum(i, k0=14.0, amplitude=6.0, seed=1000 + i)
        w0 = {i: b.read_state(i) for i in sample}
        dts = {i: [] for i in sample}
        for k in range(steps):
            b.step(1)
            for i in sample:
                dts[i].append(b.diagnostics(i)["dt"])
        print("ensemble: %d steps done" % steps, flush=True)
        got = [b.read_state(i) for i in range(len(sizes))]
        diags = [b.diagnostics(i) for i in range(len(sizes))]
    for i, n in enumerate(sizes):
        with s2t.Solver(n) as s:
            s.configure(**params[i])
            s.seed_spectrum(k0=14.0, amplitude=6.0, seed=1000 + i)
            s.step(steps)
            assert np.array_equal(got[i], s.read_state()), "member %d (%d^2) differs" % (i, n)
        assert diags[i]["stepIndex"] == steps and diags[i]["nonfinite"] == 0
    print("ensemble: every member equals its plain solver", flush=True)
    for i in sample:
        n, p = sizes[i], params[i]
        o = _oracle(p, n)
        o.w = w0[i].astype(np.complex128)
        for k, dt in enumerate(dts[i]):
            o.step(dt=dt, forcing=gs.stochastic_forcing(n, p["force_amp"], p["force_k"],
                                                        p["force_width"], k, dt))
        err = rel_err(got[i], o.w)
        print("ensemble member %d (%d^2, nu %.2e, beta %.1f, k_f %d): %d steps vs FP64 %.3e"
              % (i, n, p["viscosity"] / (n / 3.0) ** 2, p["beta"], p["force_k"], steps, err),
              flush=True)
        assert err < 1e-5


def test_ensemble_per_class_launches(s2t):
    """The same mix stepped by per-class launches on graph branches (the classes at their default
    widths: four columns per workgroup with P_hat kept in LDS, many instances per launch): every
    member equals its plain solver bit for bit."""
    sizes = bench.ENSEMBLE_SIZES
    steps = 4
    params = [bench.ensemble_params(i, n) for i, n in enumerate(sizes)]
    with s2t.Batch(sizes) as b:
        b.use_graphs(True)
        b.use_grouped(False)
        assert b.column_widths() == {256: 4, 512: 4, 1024: 4}
        for i, n in enumerate(sizes):
            b.configure(i, **params[i])
            b.seed_spectrum(i, k0=14.0, amplitude=6.0, seed=1000 + i)
        b.step(steps)
        got = [b.read_state(i) for i in range(len(sizes))]
    for i, n in enumerate(sizes):
        with s2t.Solver(n) as s:
            s.configure(**params[i])
            s.seed_spectrum(k0=14.0, amplitude=6.0, seed=1000 + i)
            s.step(steps)
            assert np.array_equal(got[i], s.read_state()), "member %d (%d^2) differs" % (i, n)


SERVE_PHYS = dict(beta=5.0, drag=0.01, viscosity=1.0, hyper_p=4.0, cfl=0.5, dt_min=1e-5,
                  dt_max=0.02)
SERVE_SLOTS = {128: 4, 256: 4, 512: 2, 1024: 2, 2048: 1}


def _burst():
    """15 requests: each size three times, the four scenarios in turn, 12..30 steps."""
    scen = ["taylor-green", "dipole", "merger", "turbulence"]
    reqs = []
    for k in range(15):
        n = [128, 256, 512, 1024, 2048][k % 5]
        sc = scen[k % 4]
        reqs.append(dict(n=n, steps=12 + (7 * k) % 19, scenario=sc,
                         scenario_param=0.5 + 0.07 * k if sc == "merger" else 0.0,
                         k0=max(4.0, n / 40.0), amplitude=6.0, seed=20 + k, keep_state=True))
    return reqs


def _oracle_initial(r):
    n = r["n"]
    o = _oracle(dict(SERVE_PHYS), n)
    if r["scenario"] == "turbulence":
        o.w = gs.fill_spectrum(n, k0=r["k0"], amplitude=r["amplitude"],
                               seed=r["seed"]).astype(np.complex128)
    else:
        o.set_physical(gs.scenario_field(r["scenario"], n, {"mergerSep": r["scenario_param"]}))
    return o


def test_serving_poisson_burst(s2t):
    """Requests arriving as a Poisson burst (100 per second) into the server: every result equals
    the same request served alone, bit for bit; the first request of each size follows the FP64
    oracle (its own CFL steps) within 1e-5 and reaches the same simulated time."""
    reqs = _burst()
    assert {r["scenario"] for r in reqs} == {"taylor-green", "dipole", "merger", "turbulence"}
    gaps = np.random.default_rng(3).exponential(1.0 / 100.0, size=len(reqs))
    with s2t.Server(SERVE_SLOTS, budget_bytes=200e9) as sv:
        ids = []
        for r, g in zip(reqs, gaps):
            time.sleep(g)
            ids.append(sv.submit(**r, **SERVE_PHYS))
        sv.drain(300)
        mixed = [sv.read_state(i) for i in ids]
        mdiag = [sv.poll(i)[1] for i in ids]
        stats = sv.stats()
    assert stats["completed"] == len(reqs)
    print("serve burst: %d done, p50 %.3f s, p99 %.3f s" % (stats["completed"],
                                                          stats["p50_latency_s"],
                                                          stats["p99_latency_s"]), flush=True)
    with s2t.Server(SERVE_SLOTS, budget_bytes=200e9) as sv:
        for k, r in enumerate(reqs):
            rid = sv.submit(**r, **SERVE_PHYS)
            sv.drain(120)
            assert np.array_equal(sv.read_state(rid), mixed[k]), "request %d differs alone" % k
    print("serve burst: every request equals its run alone", flush=True)
    seen = set()
    for k, r in enumerate(reqs):
        if r["n"] in seen:
            continue
        seen.add(r["n"])
        o = _oracle_initial(r)
        for _ in range(r["steps"]):
            o.step()
        err = rel_err(mixed[k], o.w)
        print("serve request %d (%d^2, %s, %d steps) vs FP64 %.3e, time %.6f vs %.6f"
              % (k, r["n"], r["scenario"], r["steps"], err, mdiag[k]["time"], o.t), flush=True)
        assert mdiag[k]["stepIndex"] == r["steps"]
        assert err < 1e-5
        assert abs(mdiag[k]["time"] - o.t) < 1e-6 * o.t
