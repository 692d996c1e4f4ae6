"""Throughput of the spectral 2-D turbulence step on one MI355X.

Workload (forced beta-plane turbulence, 4096^2 FP32): beta = 10, drag mu = 0.01, hyperviscosity
order 4 at rate 1 at the cut-off, stochastic forcing in the annulus k_f = 32 +- 2 with the
amplitude that injects energy at 1e-3 per unit time, CFL 0.5 adaptive step, from rest. A step is
one full IF-RK3 step (three right-hand-side evaluations, forcing, CFL reduction) on the device.

Prints one JSON line. `value` is Mpts*RK3-steps/s with the state resident in HBM; `roofline`
describes the dominant kernel, timed with HIP events on the solver's stream during the timed
steps; `cpu_baseline` times the FP64 NumPy oracle step on the host cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 4096] [--no-cpu-baseline]
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


def measured_traffic(kernel, n):
    """HBM bytes per launch of `kernel` at grid n from the committed rocprofv3 counter summary.

    The summary (profiles/step_kernels_<n>_forced_turbulence.md) holds FETCH_SIZE and WRITE_SIZE
    per dispatch of this workload, scaled by the copy-kernel calibration, as `TRAFFIC <name> <bytes>`
    lines. The step variant without pointer forcing is the one the bench runs. None when absent.
    """
    path = os.path.join(ROOT, "profiles", "step_kernels_%d_forced_turbulence.md" % n)
    if not os.path.exists(path):
        return None
    prefix = "s2t::%s<%d," % (kernel, n)
    suffix = ",0,false>" if kernel == "k_cols" else ",false>"
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 3 and parts[0] == "TRAFFIC" and parts[1].startswith(prefix) \
                    and parts[1].endswith(suffix):
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


def cpu_baseline(n, params, steps=1):
    """Time the FP64 NumPy oracle step (same operation order as reference/spectral2d.py)."""
    sys.path.insert(0, ROOT)
    from oracle import gpusem
    from oracle.spectral import Scheme, nu_from_rate
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
    el = time.perf_counter() - t0
    return dict(value=n * n * steps / el / 1e6, unit="Mpts·steps/s", cores=1, kind="port",
                sample="%d IF-RK3 step(s) at %d^2 of the FP64 NumPy oracle (numpy.fft, one thread),"
                       " CFL dt and stochastic forcing included, %.1f s" % (steps, n, el))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--copy-probe", type=float, default=0.0,
                    help="only run a plain device copy of this many bytes (counter calibration)")
    args = ap.parse_args(argv)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: a single instance does not shard across GPUs "
              "in this build (see DESIGN.md); use --gpus 1", file=sys.stderr)
        return 2

    s2t = load_binding()
    if args.copy_probe > 0:
        ms = s2t.copy_probe(args.copy_probe, reps=5)
        print(json.dumps({"copy_bytes": args.copy_probe, "ms": ms,
                          "GBps": 2 * args.copy_probe / ms / 1e6}))
        return 0
    n = args.n
    params = workload_params(n)
    solver = s2t.Solver(n)
    solver.configure(**params)
    solver.seed_spectrum(k0=1.0, amplitude=0.0, seed=1)  # from rest
    solver.step(args.warmup)
    solver.sync()

    solver.profile(True)
    t0 = time.perf_counter()
    solver.step(args.steps)
    solver.sync()
    el = time.perf_counter() - t0
    solver.profile(False)
    prof = {k: solver.profile_read(k) for k in (s2t.PROF_ROWS, s2t.PROF_COLS, s2t.PROF_OTHER)}
    diag = solver.diagnostics()
    if not np.isfinite(diag["energy"]) or diag["nonfinite"]:
        print("non-finite state after the timed steps", file=sys.stderr)
        return 1

    value = n * n * args.steps / el / 1e6
    ms = el * 1e3 / args.steps
    # dominant kernel and its algorithmic bytes per launch (mean over the three stages)
    names = {s2t.PROF_ROWS: "k_rows", s2t.PROF_COLS: "k_cols"}
    dom = max((s2t.PROF_ROWS, s2t.PROF_COLS), key=lambda k: prof[k][1])
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
        "config": {"workload": "forced beta-plane turbulence %d^2 FP32 (config 3)" % n, "n": n,
                   "beta": 10.0, "drag": 0.01, "hyper_p": 4, "viscosity_rate": 1.0,
                   "force_k": 32, "force_width": 2, "injection_rate": 1e-3,
                   "force_amp": round(params["force_amp"], 6), "cfl": 0.5,
                   "final_dt": diag["dt"], "final_energy": diag["energy"]},
        "pct_hbm_roofline_208B": round(100.0 * B_CANONICAL * value * 1e6 / (HBM_PEAK_GBS * 1e9), 2),
        "roofline": {"bound": "hbm", "kernel": names[dom], "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(names[dom], n), "avg_launch_us": round(avg_s * 1e6, 2),
                     "bytes_per_launch": bytes_launch},
        "kernels": kernels,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(n, params, steps=args.cpu_steps)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
