"""Throughput of the spectral 2-D turbulence step on one MI355X.

Workload (forced beta-plane turbulence, 4096^2 FP32): beta = 10, drag mu = 0.01, hyperviscosity
order 4 at rate 1 at the cut-off, stochastic forcing in the annulus k_f = 32 +- 2 with the
amplitude that injects energy at 1e-3 per unit time, CFL 0.5 adaptive step, from rest. A step is
one full IF-RK3 step (three right-hand-side evaluations, forcing, CFL reduction) on the device.

Prints one JSON line. `value` is Mpts*RK3-steps/s with the state resident in HBM; `roofline`
describes the dominant kernel, timed with HIP events on the solver's stream in one timed step of
four (all of its stages); `cpu_baseline` times the FP64 NumPy oracle step on the host cores.
Unless --no-workloads, the line also carries the batched ensemble, the serving loop and the
1024^2 decaying workload with dye (`ensemble`, `serve`, `decaying`).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 4096] [--no-cpu-baseline]
    python bench.py --workload ensemble [--steps K] [--plain-steps P]
    python bench.py --workload serve [--rates r1,r2,...] [--requests R]
    python bench.py --workload decaying [--steps K]
"""

import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0  # MI355X datasheet HBM3E bandwidth
B_CANONICAL = 208.0    # bytes/pt/step of the fused full-complex schedule used for the target
PROFILE_EVERY = 4      # the dominant kernel is timed in one step of this many
SETTLE_S = 1.5         # seconds of untimed stepping after the warm-up, before the timed steps


def measured_traffic(kernel, n):
    """HBM bytes per launch of `kernel` at grid n from the committed rocprofv3 counter summary.

    The summary (profiles/step_kernels_<n>_forced_turbulence.md) holds FETCH_SIZE and WRITE_SIZE
    per dispatch of this workload, scaled by the copy-kernel calibration, as `TRAFFIC <name> <bytes>`
    lines. The step variant without pointer forcing is the one the bench runs. None when absent.
    """
    path = os.path.join(ROOT, "profiles", "step_kernels_%d_forced_turbulence.md" % n)
    if not os.path.exists(path):
        return None
    # the kernel's one-group or field-parallel form (k_cols<n,C,0,false>, k_cols_fp<n,G,false>;
    # k_rows<n,G,2>, k_rows_fp<n,2>)
    prefixes = ("s2t::%s<%d," % (kernel, n), "s2t::%s_fp<%d," % (kernel, n))
    suffixes = (",0,false>",) if kernel == "k_cols" else (",2>",)
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) != 3 or parts[0] != "TRAFFIC":
                continue
            name = parts[1]
            if name.startswith(prefixes[0]) and name.endswith(suffixes[0]):
                return float(parts[2])
            if name.startswith(prefixes[1]) and (kernel == "k_rows" and name.endswith(",2>")
                                                 or kernel == "k_cols" and name.endswith(",false>")):
                return float(parts[2])
    return None


def load_binding():
    spec = importlib.util.spec_from_file_location(
        "s2t_amd", os.path.join(ROOT, "spectral-2d-turbulence_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def force_amp_for_rate(eps, n, k_f, width):
    """Amplitude giving mean injection rate eps: eps = 1/2 (0.01 A)^2 sum_k env_k^2 / k^2.

    Sum over the in-band modes with env = exp(-((|k| - k_f)/width)^2) > 1e-4.
    """
    k = np.fft.fftfreq(n, 1.0 / n)
    kcut = n / 3.0
    kb = k[np.abs(k) < kcut]
    ky, kx = np.meshgrid(kb, kb, indexing="ij")
    k2 = kx * kx + ky * ky
    kr = np.sqrt(k2)
    env = np.exp(-(((kr - k_f) / width) ** 2))
    keep = (env > 1e-4) & (k2 > 0)
    s = float(np.sum(env[keep] ** 2 / k2[keep]))
    return float(np.sqrt(2.0 * eps / s) / 0.01)


def workload_params(n):
    return dict(beta=10.0, drag=0.01, viscosity=1.0, hyper_p=4.0,
                force_amp=force_amp_for_rate(1e-3, n, 32.0, 2.0), force_k=32.0, force_width=2.0,
                cfl=0.5, dt_min=1e-5, dt_max=0.02, dye_fade=1.0, fixed_dt=0.0)


def cpu_baseline(n, params, steps=1, threads=0):
    """Time the FP64 NumPy oracle step (same operation order as reference/spectral2d.py) on one
    thread (numpy.fft), and, with threads > 1, the same steps again with scipy.fft on that many
    workers and the pointwise stages over row bands (reported under "threaded")."""
    sys.path.insert(0, ROOT)
    from oracle import gpusem
    from oracle import spectral as sp
    from oracle.spectral import Scheme, nu_from_rate

    def run(workers):
        sp.set_fft_workers(workers)
        o = Scheme(n, beta=params["beta"], mu=params["drag"],
                   nu=nu_from_rate(params["viscosity"], n, int(params["hyper_p"])),
                   p=int(params["hyper_p"]), cfl=params["cfl"], dt_max=params["dt_max"])
        o.w = gpusem.fill_spectrum(n, k0=64.0, amplitude=1.0, seed=1).astype(np.complex128)
        t0 = time.perf_counter()
        for k in range(steps):
            dt = o.cfl_dt()
            f = gpusem.stochastic_forcing(n, params["force_amp"], params["force_k"],
                                          params["force_width"], k, dt)
            o.step(dt=dt, forcing=f)
        return time.perf_counter() - t0

    el = run(1)
    out = dict(value=n * n * steps / el / 1e6, unit="Mpts·steps/s", cores=1, kind="port",
               sample="%d IF-RK3 step(s) at %d^2 of the FP64 NumPy oracle (numpy.fft, one thread),"
                      " CFL dt and stochastic forcing included, %.1f s" % (steps, n, el))
    if threads > 1:
        el_t = run(threads)
        out["threaded"] = dict(value=n * n * steps / el_t / 1e6, cores=threads,
                               sample="the same steps with scipy.fft on %d workers and the pointwise"
                                      " stages over %d row bands, %.1f s" % (threads, threads, el_t))
    sp.set_fft_workers(1)
    return out


ENSEMBLE_SIZES = [256] * 32 + [512] * 24 + [1024] * 8  # 16,777,216 points = 4096^2


def ensemble_params(i, n):
    """Member i of the mixed-size ensemble: Laplacian viscosity nu log-uniform in [1e-5, 1e-3]
    (p = 1; the viscosity knob is the rate at the cut-off, nu * (n/3)^2), beta ~ U[0, 20],
    forcing at k_f in {16, 32} +- 2 injecting 1e-3, CFL 0.5."""
    rng = np.random.default_rng(1000 + i)
    nu = 10 ** rng.uniform(-5, -3)
    kf = float(rng.choice([16, 32]))
    return dict(beta=float(rng.uniform(0, 20)), drag=0.0, viscosity=nu * (n / 3.0) ** 2,
                hyper_p=1.0, force_amp=force_amp_for_rate(1e-3, n, kf, 2.0), force_k=kf,
                force_width=2.0, cfl=0.5, dt_min=1e-5, dt_max=0.02, dye_fade=1.0, fixed_dt=0.0)


def bench_ensemble(s2t, steps, warmup, plain_steps, tune=True, grouped_only=False, eager=False):
    """64 independent members (32 x 256^2, 24 x 512^2, 8 x 1024^2) advanced together by the
    batch path: grouped launches (the default; six launches per step, each over every member,
    hipGraph replay), then one launch per kernel and size class on concurrent graph branches
    (default column-group widths, and the widths the tuner picks per class), against a loop over
    the plain single-instance path on the same GPU. grouped_only: the grouped launches alone (the
    per-class launches, the tuner and the plain loop are skipped; for profiling them). eager: the
    launches are issued one by one rather than replayed from a captured graph (the same kernels)."""
    sizes = ENSEMBLE_SIZES
    pts = sum(n * n for n in sizes)

    def timed(b, settle=False):
        b.step(warmup)
        b.sync()
        # untimed steps until the GPU clocks have settled, as for the default workload
        t_settle = time.perf_counter()
        while settle and time.perf_counter() - t_settle < SETTLE_S:
            b.step(20)
            b.sync()
        t0 = time.perf_counter()
        b.step(steps)
        b.sync()
        return time.perf_counter() - t0

    with s2t.Batch(sizes) as b:
        b.use_graphs(not eager)
        for i, n in enumerate(sizes):
            b.configure(i, **ensemble_params(i, n))
            b.seed_spectrum(i, k0=14.0, amplitude=6.0, seed=1000 + i)
        el = timed(b, settle=True)
        if grouped_only:
            if not all(np.isfinite(b.diagnostics(i)["energy"]) for i in range(len(sizes))):
                raise RuntimeError("non-finite ensemble members")
            return {"metric": "ensemble Mpts·RK3-steps/s (grouped launches only)",
                    "value": round(pts * steps / el / 1e6, 2), "unit": "Mpts·steps/s",
                    "steps": steps, "warmup": warmup, "ms_per_step": round(el * 1e3 / steps, 4)}
        b.use_grouped(False)
        default_widths = b.column_widths()
        el_class = timed(b)
        per_class = {"value": round(pts * steps / el_class / 1e6, 2),
                     "ms_per_step": round(el_class * 1e3 / steps, 4),
                     "widths": {str(k): v for k, v in sorted(default_widths.items())}}
        if tune:
            # tuning runs on scratch solvers, outside the timed regions
            widths = b.autotune(steps=20)
            el_tuned = timed(b)
            per_class["tuned"] = {"value": round(pts * steps / el_tuned / 1e6, 2),
                                  "widths": {str(k): v for k, v in sorted(widths.items())}}
        bad = [i for i in range(len(sizes)) if not np.isfinite(b.diagnostics(i)["energy"])]
    if bad:
        raise RuntimeError("non-finite ensemble members %s" % bad)
    batch_rate = pts * steps / el / 1e6
    # plain path: each member on its own solver, stepped one after the other
    el_plain = 0.0
    for i, n in enumerate(sizes):
        with s2t.Solver(n) as s:
            s.configure(**ensemble_params(i, n))
            s.seed_spectrum(k0=14.0, amplitude=6.0, seed=1000 + i)
            s.step(warmup)
            s.sync()
            t0 = time.perf_counter()
            s.step(plain_steps)
            s.sync()
            el_plain += time.perf_counter() - t0
    plain_rate = pts * plain_steps / el_plain / 1e6
    return {"metric": "ensemble Mpts·RK3-steps/s (64 members, 32x256^2 + 24x512^2 + 8x1024^2)",
            "value": round(batch_rate, 2), "unit": "Mpts·steps/s", "n_gpus": 1, "steps": steps,
            "warmup": warmup, "ms_per_step": round(el * 1e3 / steps, 4), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic: forced turbulence members, seeds 1000+i",
            "config": {"workload": "batched ensemble, 64 members, 16,777,216 points",
                       "members": len(sizes), "points": pts,
                       "launches": "grouped: 6 per step over all members"},
            "per_class_graph": per_class,
            "plain_loop": {"value": round(plain_rate, 2), "steps_per_member": plain_steps,
                           "seconds": round(el_plain, 3)},
            "speedup_vs_plain_loop": round(batch_rate / plain_rate, 3)}


SERVE_SIZES = [128, 256, 512, 1024, 2048]
SERVE_WEIGHTS = [0.3, 0.3, 0.2, 0.15, 0.05]
SERVE_SLOTS = {128: 32, 256: 32, 512: 16, 1024: 8, 2048: 4}
SERVE_SCENARIOS = ["taylor-green", "dipole", "merger", "turbulence"]


def serve_requests(count, seed):
    """Request mix: N by weight, steps log-uniform in [100, 5000], scenario uniform, merger
    separation ~ U[0.5, 1.6]."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(rng.choice(SERVE_SIZES, p=SERVE_WEIGHTS))
        steps = int(round(10 ** rng.uniform(2, np.log10(5000))))
        sc = SERVE_SCENARIOS[int(rng.integers(4))]
        param = float(rng.uniform(0.5, 1.6)) if sc == "merger" else 0.0
        out.append(dict(n=n, steps=steps, scenario=sc, scenario_param=param, seed=k + 1))
    return out


SERVE_PHYS = dict(beta=5.0, drag=0.01, viscosity=1.0, hyper_p=4.0, cfl=0.5, dt_min=1e-5,
                  dt_max=0.02)


def run_served(s2t, reqs, gaps):
    """Submit `reqs` at the given gaps (s; None: all at once) into a server with a 200 GB budget
    and wait for them; returns the server's statistics and the makespan."""
    with s2t.Server(SERVE_SLOTS, budget_bytes=200e9) as sv:
        t0 = time.perf_counter()
        due = t0
        for k, r in enumerate(reqs):
            if gaps is not None:
                due += gaps[k]
                delay = due - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
            sv.submit(**r, **SERVE_PHYS)
        sv.drain(900)
        span = time.perf_counter() - t0
        st = sv.stats()
    return st, span


def run_plain_loop(s2t, reqs):
    """The same requests one after another on plain solvers of their sizes (seeding included,
    as the server's admission does): seconds in total."""
    total = 0.0
    solvers = {}
    try:
        for r in reqs:
            n = r["n"]
            if n not in solvers:
                solvers[n] = s2t.Solver(n)
            s = solvers[n]
            t0 = time.perf_counter()
            s.configure(**SERVE_PHYS)
            if r["scenario"] == "turbulence":
                s.seed_spectrum(k0=14.0, amplitude=6.0, seed=r["seed"])
            else:
                s.seed_physical(serve_field(r["scenario"], n, r["scenario_param"]))
            s.step(r["steps"])
            s.sync()
            total += time.perf_counter() - t0
    finally:
        for s in solvers.values():
            s.close()
    return total


def serve_field(scenario, n, param):
    """Physical vorticity of a server scenario on x_j = 2 pi j / N (src/scenarios.js:82-145)."""
    h = 2 * np.pi / n
    y, x = np.meshgrid(np.arange(n) * h, np.arange(n) * h, indexing="ij")
    if scenario == "taylor-green":
        return (6 * np.cos(2 * x) * np.cos(2 * y)).astype(np.float32)
    if scenario == "dipole":
        a = 0.3
        amp = 12 / (np.pi * a * a)
        cores = [(0.6 * np.pi, np.pi - 0.45, a, amp), (0.6 * np.pi, np.pi + 0.45, a, -amp)]
    else:
        d = param if param > 0 else 0.7
        cores = [(np.pi - d / 2, np.pi, 0.22, 30.0), (np.pi + d / 2, np.pi, 0.22, 30.0)]
    f = np.zeros((n, n))
    for cx, cy, a, amp in cores:
        dx = (x - cx + np.pi) % (2 * np.pi) - np.pi
        dy = (y - cy + np.pi) % (2 * np.pi) - np.pi
        r2 = dx * dx + dy * dy
        f += np.where(r2 < 25 * a * a, amp * np.exp(-r2 / (a * a)), 0.0)
    return f.astype(np.float32)


def bench_serve(s2t, rates, count, seed=5, plain=True):
    """Poisson arrivals at each rate (requests/s) into the continuous-batching server, then all at
    once (the saturation throughput); completion-latency percentiles and completed Mpts*steps/s
    per rate, and the same requests run one at a time on plain solvers."""
    reqs = serve_requests(count, seed)
    work = float(np.sum([r["n"] ** 2 * r["steps"] for r in reqs]))
    lines = []
    for rate in list(rates) + [None]:
        gaps = None if rate is None else np.random.default_rng(seed + 1).exponential(1.0 / rate,
                                                                                    size=count)
        st, span = run_served(s2t, reqs, gaps)
        lines.append({"rate_req_s": rate if rate is not None else "all at once",
                      "offered_mpts_steps_per_s": round(rate * work / count / 1e6, 1)
                      if rate is not None else None,
                      "requests": count, "completed": st["completed"],
                      "rejected": st["rejected"], "p50_latency_s": round(st["p50_latency_s"], 4),
                      "p99_latency_s": round(st["p99_latency_s"], 4),
                      "mean_latency_s": round(st["mean_latency_s"], 4),
                      "mpts_steps_per_s": round(st["mpts_steps_per_s"], 1),
                      "makespan_s": round(span, 3), "peak_gb": round(st["peak_bytes"] / 1e9, 3)})
        print(json.dumps(lines[-1]), file=sys.stderr, flush=True)
    saturation = lines[-1]["mpts_steps_per_s"]
    out = {"metric": "serving Mpts·RK3-steps/s under Poisson arrivals",
           "value": saturation, "unit": "Mpts·steps/s",
           "n_gpus": 1, "higher_is_better": True, "vs_baseline": None, "dtype": "f32",
           "data": "synthetic: scenario mix, N in {128..2048}^2, steps log-uniform 100..5000",
           "config": {"workload": "continuous-batching server", "slots_per_chunk": SERVE_SLOTS,
                      "requests": count, "work_pts_steps": work, "budget_bytes": 200e9},
           "saturation_mpts_steps_per_s": saturation, "sweep": lines}
    if plain:
        el = run_plain_loop(s2t, reqs)
        out["plain_loop"] = {"seconds": round(el, 3),
                             "mpts_steps_per_s": round(work / el / 1e6, 1)}
        out["served_vs_plain"] = round(saturation / (work / el / 1e6), 3)
    return out


def blob_dye(m, seed, count=12):
    """Seeded Gaussian colour blobs on the m x m dye grid (rgba16f, alpha 1, colours in [0, 1])."""
    rng = np.random.default_rng(seed)
    x = (np.arange(m) + 0.5) / m
    dye = np.zeros((m, m, 4), np.float32)
    dye[..., 3] = 1.0
    for _ in range(count):
        cx, cy, r = rng.uniform(0, 1), rng.uniform(0, 1), rng.uniform(0.03, 0.1)
        dx = (x - cx + 0.5) % 1.0 - 0.5
        dy = (x - cy + 0.5) % 1.0 - 0.5
        g = np.exp(-(dy[:, None] ** 2 + dx[None, :] ** 2) / r ** 2).astype(np.float32)
        dye[..., :3] += g[..., None] * rng.uniform(0, 1, 3).astype(np.float32)
    dye[..., :3] = np.clip(dye[..., :3], 0.0, 1.0)
    return dye.astype(np.float16)


def bench_decaying(s2t, steps, warmup):
    """Decaying turbulence at 1024^2 with dye on the 2048^2 refined grid: band-limited seed
    |k| in [8, 16] (seed 2, omega_rms 6), hyperviscosity p = 4 at rate omega_rms, CFL 0.5, seeded
    Gaussian dye blobs, BFECC dye advection every step. After the warm-up, untimed steps run for
    SETTLE_S seconds; the step rate is then timed without events, and the dye step and the step
    kernels with events in a separate run."""
    n = 1024
    params = dict(beta=0.0, drag=0.0, viscosity=6.0, hyper_p=4.0, force_amp=0.0, force_k=32.0,
                  force_width=2.0, cfl=0.5, dt_min=1e-5, dt_max=0.02, dye_fade=1.0, fixed_dt=0.0)
    with s2t.Solver(n, dye_scale=2) as s:
        m = s.dye_size()
        s.configure(**params)
        s.seed_bandlimited(8.0, 16.0, 6.0, seed=2)
        s.write_dye(blob_dye(m, seed=2))
        s.step(warmup, dye=True)
        s.sync()
        # untimed steps until the GPU clocks have settled, as for the default workload
        t_settle = time.perf_counter()
        while time.perf_counter() - t_settle < SETTLE_S:
            s.step(50, dye=True)
            s.sync()
        t0 = time.perf_counter()
        s.step(steps, dye=True)
        s.sync()
        el = time.perf_counter() - t0
        s.profile(True, kinds=[s2t.PROF_DYE, s2t.PROF_COLS, s2t.PROF_ROWS])
        s.step(max(10, steps // 4), dye=True)
        s.sync()
        s.profile(False)
        prof = {k: s.profile_read(k) for k in (s2t.PROF_ROWS, s2t.PROF_COLS, s2t.PROF_DYE)}
        dye_bytes = s.kernel_bytes(s2t.PROF_DYE)
        # the dye launch alone: explicit dye steps back to back on the solver's stream (they
        # advance the dye without the flow; after the timed steps)
        s.sync()
        reps = 100
        t0 = time.perf_counter()
        for _ in range(reps):
            s.dye_step()
        s.sync()
        dye_alone_us = (time.perf_counter() - t0) / reps * 1e6
        d = s.diagnostics()
        if not np.isfinite(d["energy"]) or d["nonfinite"]:
            raise RuntimeError("non-finite state after the timed steps")
    avg_us = {k: prof[k][1] / max(1, prof[k][0]) * 1e3 for k in prof}
    dye_gbs = dye_bytes / (dye_alone_us * 1e-6) / 1e9
    return {"metric": "Mpts·RK3-steps/s at 1024² FP32 with dye at 2048²",
            "value": round(n * n * steps / el / 1e6, 2), "unit": "Mpts·steps/s", "n_gpus": 1,
            "steps": steps, "warmup": warmup, "ms_per_step": round(el * 1e3 / steps, 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic: band-limited seed 2, seeded Gaussian dye blobs",
            "config": {"workload": "decaying turbulence 1024^2 FP32, dye 2048^2 rgba16f", "n": n,
                       "dye": m, "hyper_p": 4, "viscosity_rate": 6.0, "cfl": 0.5,
                       "final_dt": d["dt"], "final_energy": d["energy"]},
            "dye_step": {"bound": "hbm", "avg_us": round(dye_alone_us, 2),
                         "span_beside_steps_us": round(avg_us[s2t.PROF_DYE], 2),
                         "schedule": "in a step it runs on a stream of its own beside stages 1 "
                                     "and 2; avg_us (and achieved) is the launch alone, "
                                     "span_beside_steps_us its span there, which includes the "
                                     "time it shares the device with them",
                         "bytes": dye_bytes, "achieved": round(dye_gbs, 1),
                         "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(dye_gbs / HBM_PEAK_GBS, 4)},
            "kernels_us": {"k_rows": round(avg_us[s2t.PROF_ROWS], 2),
                           "k_cols": round(avg_us[s2t.PROF_COLS], 2),
                           "dye (fused, alone)": round(dye_alone_us, 2),
                           "dye (fused, beside the step)": round(avg_us[s2t.PROF_DYE], 2)}}


def dissipation_rate(spectrum, n, params):
    """Energy removed per unit time by drag and hyperviscosity, from the FP64 E(k) bins:
    sum_b 2 (mu + nu_h b^(2p)) E_b with nu_h = rate / (n/3)^(2p). The bins are round(|k|); over a
    bin of width one, |k|^(2p) averages to b^(2p) (1 + O(p^2 / b^2)), and the hyperviscous loss
    sits at b in the hundreds."""
    b = np.arange(spectrum.size, dtype=np.float64)
    p = params["hyper_p"]
    nu = params["viscosity"] / (n / 3.0) ** (2 * p) if params["viscosity"] > 0 else 0.0
    return float(np.sum(2.0 * (params["drag"] + nu * b ** (2 * p)) * spectrum))


def spectral_slope(spectrum, kmin, kmax):
    """Least-squares slope of log E(k) against log k over the bins kmin..kmax."""
    b = np.arange(spectrum.size)
    sel = (b >= kmin) & (b <= kmax) & (spectrum > 0)
    return float(np.polyfit(np.log(b[sel]), np.log(spectrum[sel]), 1)[0])


def anisotropy(s2t, w, n):
    """<u^2> / <v^2> of the physical velocity of the spectrum w (FP64 on the host)."""
    k = np.fft.fftfreq(n, 1.0 / n)
    ky, kx = np.meshgrid(k, k, indexing="ij")
    k2 = kx * kx + ky * ky
    inv = np.where(k2 > 0, 1.0 / np.where(k2 > 0, k2, 1.0), 0.0)
    u = np.fft.ifft2(1j * ky * inv * w).real
    v = np.fft.ifft2(-1j * kx * inv * w).real
    return float(np.mean(u * u) / np.mean(v * v))


def run_forced_horizon(s2t, n, steps, every=25):
    """The forced beta-plane workload from rest for `steps` CFL steps, sampling the FP64
    diagnostics every `every` steps. Returns the time series of E, Z and the dissipation rate,
    the energy budget (E(T) against the injected eps T minus the integrated dissipation,
    trapezoid in time over the samples), the final zonal-mean u(y), E(k), its slope over the
    enstrophy-cascade range and the velocity anisotropy <u^2>/<v^2>."""
    params = workload_params(n)
    eps = 1e-3
    ts, es, zs, ds = [0.0], [0.0], [0.0], [0.0]
    t0 = time.perf_counter()
    with s2t.Solver(n) as s:
        s.configure(**params)
        s.seed_spectrum(k0=1.0, amplitude=0.0, seed=1)  # from rest
        done = 0
        while done < steps:
            k = min(every, steps - done)
            s.step(k)
            done += k
            d = s.diagnostics()
            if d["nonfinite"]:
                raise RuntimeError("non-finite state at step %d" % done)
            ts.append(d["time"])
            es.append(d["energy"])
            zs.append(d["enstrophy"])
            ds.append(dissipation_rate(d["spectrum"], n, params))
        wall = time.perf_counter() - t0
        w = s.read_state()
    t = np.asarray(ts)
    lost = float(np.sum(0.5 * (np.asarray(ds[1:]) + np.asarray(ds[:-1])) * np.diff(t)))
    injected = eps * t[-1]
    return {"n": n, "steps": steps, "time": float(t[-1]), "wall_s": round(wall, 2),
            "final_dt": d["dt"], "energy": d["energy"], "enstrophy": d["enstrophy"],
            "budget": {"injected": injected, "dissipated": lost,
                       "predicted_energy": injected - lost, "energy": d["energy"],
                       "residual_over_injected": (d["energy"] - (injected - lost)) / injected},
            "spectrum_slope_k64_k256": spectral_slope(d["spectrum"], 64, 256),
            "anisotropy_u2_v2": anisotropy(s2t, w, n),
            "series": {"time": ts, "energy": es, "enstrophy": zs, "dissipation": ds},
            "zonal_mean_u": [float("%.6g" % x) for x in d["zonal"]],
            "energy_spectrum": [float("%.6g" % x) for x in d["spectrum"]]}


def run_decaying_horizon(s2t, steps, every=20):
    """The 1024^2 decaying workload with dye on 2048^2 for `steps` CFL steps: the enstrophy and
    energy series (every `every` steps), the E(k) slope over k in [32, 160] at the end and the
    dye's range."""
    n = 1024
    params = dict(beta=0.0, drag=0.0, viscosity=6.0, hyper_p=4.0, force_amp=0.0, force_k=32.0,
                  force_width=2.0, cfl=0.5, dt_min=1e-5, dt_max=0.02, dye_fade=1.0, fixed_dt=0.0)
    ts, es, zs = [], [], []
    t0 = time.perf_counter()
    with s2t.Solver(n, dye_scale=2) as s:
        m = s.dye_size()
        s.configure(**params)
        s.seed_bandlimited(8.0, 16.0, 6.0, seed=2)
        s.write_dye(blob_dye(m, seed=2))
        d = s.diagnostics()
        ts.append(d["time"]), es.append(d["energy"]), zs.append(d["enstrophy"])
        done = 0
        while done < steps:
            k = min(every, steps - done)
            s.step(k, dye=True)
            done += k
            d = s.diagnostics()
            if d["nonfinite"]:
                raise RuntimeError("non-finite state at step %d" % done)
            ts.append(d["time"]), es.append(d["energy"]), zs.append(d["enstrophy"])
        wall = time.perf_counter() - t0
        dye = s.read_dye().astype(np.float64)
    return {"n": n, "dye": m, "steps": steps, "time": ts[-1], "wall_s": round(wall, 2),
            "final_dt": d["dt"], "enstrophy_decay": 1.0 - zs[-1] / zs[0],
            "energy_decay": 1.0 - es[-1] / es[0],
            "spectrum_slope_k32_k160": spectral_slope(d["spectrum"], 32, 160),
            "dye_min": float(dye[..., :3].min()), "dye_max": float(dye[..., :3].max()),
            "dye_alpha_one": bool(np.all(dye[..., 3] == 1.0)),
            "series": {"time": ts, "energy": es, "enstrophy": zs},
            "energy_spectrum": [float("%.6g" % x) for x in d["spectrum"]]}


def bench_horizons(s2t, out=None):
    """C2 to 2000 steps and C3 to 5000 steps at 2048^2 and 4096^2: the workloads' own horizons
    and outputs (zonal-mean u and E(k)), with the statistics their physics fixes."""
    res = {"decaying_1024": run_decaying_horizon(s2t, 2000)}
    print("decaying 1024^2: t %.3f, slope %.2f" % (res["decaying_1024"]["time"],
          res["decaying_1024"]["spectrum_slope_k32_k160"]), file=sys.stderr, flush=True)
    for n in (2048, 4096):
        r = run_forced_horizon(s2t, n, 5000)
        res["forced_%d" % n] = r
        print("forced %d^2: t %.3f, budget residual %.4f, slope %.2f, anisotropy %.2f"
              % (n, r["time"], r["budget"]["residual_over_injected"],
                 r["spectrum_slope_k64_k256"], r["anisotropy_u2_v2"]), file=sys.stderr, flush=True)
    if out:
        with open(out, "w") as f:
            json.dump(res, f, indent=1)
    summary = {k: {kk: vv for kk, vv in v.items()
                   if kk not in ("series", "zonal_mean_u", "energy_spectrum")}
               for k, v in res.items()}
    return {"metric": "long-horizon statistics of the decaying and forced workloads",
            "config": {"workload": "decaying 1024^2 x 2000 steps with dye 2048^2; forced "
                                   "beta-plane 2048^2 and 4096^2 x 5000 steps from rest"},
            "results": summary}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-workloads", action="store_true",
                    help="step: leave out the ensemble and serving summaries")
    ap.add_argument("--workload", choices=["step", "ensemble", "serve", "decaying", "horizons"],
                    default="step",
                    help="step: one 4096^2 instance (default); ensemble: 64 mixed-size members; "
                         "serve: Poisson arrivals into the serving loop; decaying: 1024^2 "
                         "decaying turbulence with dye on the 2048^2 grid")
    ap.add_argument("--grouped-only", action="store_true",
                    help="ensemble: the grouped launches alone (no per-class launches, tuner or "
                         "plain loop)")
    ap.add_argument("--eager", action="store_true",
                    help="ensemble: issue the launches one by one instead of replaying a graph")
    ap.add_argument("--plain-steps", type=int, default=50,
                    help="ensemble: steps per member of the plain-loop baseline")
    ap.add_argument("--rates", type=str, default="",
                    help="serve: comma-separated arrival rates (requests/s)")
    ap.add_argument("--requests", type=int, default=120, help="serve: requests per rate")
    ap.add_argument("--out", type=str, default="",
                    help="horizons: also write the full series and profiles to this JSON file")
    ap.add_argument("--copy-probe", type=float, default=0.0,
                    help="only run a plain device copy of this many bytes (counter calibration)")
    args = ap.parse_args(argv)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: a single instance does not shard across GPUs "
              "in this build; use --gpus 1", file=sys.stderr)
        return 2

    s2t = load_binding()
    if args.copy_probe > 0:
        ms = s2t.copy_probe(args.copy_probe, reps=5)
        print(json.dumps({"copy_bytes": args.copy_probe, "ms": ms,
                          "GBps": 2 * args.copy_probe / ms / 1e6}))
        return 0
    if args.workload == "ensemble":
        print(json.dumps(bench_ensemble(s2t, args.steps, args.warmup, args.plain_steps,
                                        grouped_only=args.grouped_only, eager=args.eager)))
        return 0
    if args.workload == "horizons":
        print(json.dumps(bench_horizons(s2t, args.out)))
        return 0
    if args.workload == "decaying":
        print(json.dumps(bench_decaying(s2t, args.steps, args.warmup)))
        return 0
    if args.workload == "serve":
        rates = [float(r) for r in args.rates.split(",")] if args.rates else [10.0, 30.0, 60.0]
        print(json.dumps(bench_serve(s2t, rates, args.requests)))
        return 0
    n = args.n
    params = workload_params(n)
    solver = s2t.Solver(n)
    solver.configure(**params)
    solver.seed_spectrum(k0=1.0, amplitude=0.0, seed=1)  # from rest
    solver.step(args.warmup)
    solver.sync()
    # untimed steps until the GPU clocks have settled (at least SETTLE_S of back-to-back work
    # after the warm-up): a few warm-up steps alone leave the first timed steps at lower clocks
    t_settle = time.perf_counter()
    settle_steps = 0
    while time.perf_counter() - t_settle < SETTLE_S:
        solver.step(50)
        solver.sync()
        settle_steps += 50

    # timed region: HIP events only around the dominant kernel (the column kernel), in one step
    # of four (all three stages of that step); each event pair costs a few microseconds of stream
    # time, so timing every launch would depress `value`
    solver.profile(True, kinds=[s2t.PROF_COLS], every=PROFILE_EVERY)
    t0 = time.perf_counter()
    solver.step(args.steps)
    solver.sync()
    el = time.perf_counter() - t0
    solver.profile(False)
    prof = {k: solver.profile_read(k) for k in (s2t.PROF_ROWS, s2t.PROF_COLS, s2t.PROF_OTHER)}
    # the row kernel's time for the `kernels` record, from a short separate run
    solver.profile(True, kinds=[s2t.PROF_ROWS])
    solver.step(max(10, args.steps // 10))
    solver.sync()
    solver.profile(False)
    prof[s2t.PROF_ROWS] = solver.profile_read(s2t.PROF_ROWS)
    diag = solver.diagnostics()
    if not np.isfinite(diag["energy"]) or diag["nonfinite"]:
        print("non-finite state after the timed steps", file=sys.stderr)
        return 1

    value = n * n * args.steps / el / 1e6
    ms = el * 1e3 / args.steps
    # dominant kernel and its algorithmic bytes per launch (mean over the three stages)
    names = {s2t.PROF_ROWS: "k_rows", s2t.PROF_COLS: "k_cols"}
    # dominant kernel by device time per step (both run three times a step)
    avg = {k: prof[k][1] / max(1, prof[k][0]) for k in (s2t.PROF_ROWS, s2t.PROF_COLS)}
    dom = max(avg, key=avg.get)
    launches, tot_ms = prof[dom]
    avg_s = tot_ms / launches / 1e3
    bytes_launch = np.mean([solver.kernel_bytes(dom, st) for st in (0, 1, 2)])
    achieved = bytes_launch / avg_s / 1e9
    kernels = {names[k]: dict(launches=prof[k][0], avg_us=prof[k][1] / max(1, prof[k][0]) * 1e3,
                              bytes_per_launch=float(np.mean([solver.kernel_bytes(k, st)
                                                              for st in (0, 1, 2)])))
               for k in names}
    solver.close()

    out = {
        "metric": "Mpts·RK3-steps/s at %d² FP32" % n,
        "value": round(value, 2),
        "unit": "Mpts·steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: forced beta-plane turbulence from rest, state resident in HBM",
        "config": {"workload": "forced beta-plane turbulence %d^2 FP32" % n, "n": n,
                   "beta": 10.0, "drag": 0.01, "hyper_p": 4, "viscosity_rate": 1.0,
                   "force_k": 32, "force_width": 2, "injection_rate": 1e-3,
                   "force_amp": round(params["force_amp"], 6), "cfl": 0.5,
                   "final_dt": diag["dt"], "final_energy": diag["energy"]},
        "pct_hbm_roofline_208B": round(100.0 * B_CANONICAL * value * 1e6 / (HBM_PEAK_GBS * 1e9), 2),
        "roofline": {"bound": "hbm", "kernel": names[dom], "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(names[dom], n), "avg_launch_us": round(avg_s * 1e6, 2),
                     "bytes_per_launch": bytes_launch,
                     "limiter": "latency of its chain of loads, barriers and transforms at 2 "
                                "workgroups per CU (LDS- and register-limited): its traffic is "
                                "1.05x the algorithmic bytes and 32.7 M VALU instructions per "
                                "launch are about half its time (profiles/step_kernels_4096_"
                                "forced_turbulence.md)"},
        "kernels": kernels,
        "warmup_settle": {"steps": settle_steps, "seconds": SETTLE_S},
    }
    if not args.no_workloads:
        # the second half of the metric: batched-ensemble and served throughput, each against a
        # loop over the plain single-instance path on the same GPU
        ens = bench_ensemble(s2t, 100, 10, 20, tune=False)
        out["ensemble"] = {k: ens[k] for k in ("value", "unit", "ms_per_step", "plain_loop",
                                               "speedup_vs_plain_loop", "per_class_graph")}
        out["ensemble"]["workload"] = ens["config"]["workload"]
        srv = bench_serve(s2t, [10.0, 30.0, 60.0], 80)
        out["serve"] = {"saturation_mpts_steps_per_s": srv["saturation_mpts_steps_per_s"],
                        "plain_loop": srv["plain_loop"],
                        "served_vs_plain": srv["served_vs_plain"],
                        "sweep": [{k: l[k] for k in ("rate_req_s", "p50_latency_s",
                                                     "p99_latency_s", "mpts_steps_per_s")}
                                  for l in srv["sweep"]],
                        "workload": "80 requests, N in {128..2048}^2, steps log-uniform "
                                    "100..5000, four scenarios, 200 GB budget"}
        # the 1024^2 decaying workload with BFECC dye on the 2048^2 grid
        dec = bench_decaying(s2t, 300, 20)
        out["decaying"] = {k: dec[k] for k in ("value", "unit", "ms_per_step", "kernels_us")}
        out["decaying"]["workload"] = dec["config"]["workload"]
    if not args.no_cpu_baseline:
        # the host's cores for the threaded variant (at most 16, the box's CPU share)
        threads = min(16, len(os.sched_getaffinity(0)))
        out["cpu_baseline"] = cpu_baseline(n, params, steps=args.cpu_steps, threads=threads)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
