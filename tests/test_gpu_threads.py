"""Solvers and batches stepped from several host threads at once. Every solver owns its buffers,
its stream and (with dye) a dye stream and two events, every batch its members, streams and
captured graphs; the library keeps no other state between calls, and the bindings release the
interpreter lock while a call runs. So two solvers and a batch stepped at the same time from three
threads end bit for bit where each ends when stepped alone."""

import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DECAY = dict(beta=0.0, drag=0.0, viscosity=6.0, hyper_p=4.0, force_amp=0.0, cfl=0.5,
             dt_min=1e-5, dt_max=0.02, dye_fade=1.0)
FORCED = dict(beta=10.0, drag=0.01, viscosity=1.0, hyper_p=4.0, force_k=32.0, force_width=2.0,
              force_amp=6.0, cfl=0.5)


def run_dye(s2t, steps):
    with s2t.Solver(512, dye_scale=2) as s:
        s.configure(**DECAY)
        s.seed_bandlimited(8.0, 16.0, 6.0, seed=4)
        s.paint_dye("bands")
        for _ in range(steps):
            s.step(1, dye=True)
        return s.read_state(), s.read_dye().view(np.uint16).copy()


def run_forced(s2t, steps):
    with s2t.Solver(1024) as s:
        s.configure(**FORCED)
        s.seed_spectrum(k0=14.0, amplitude=6.0, seed=9)
        s.step(steps)
        return s.read_state(), s.diagnostics()["energy"]


BATCH_SIZES = [256, 512, 512, 1024]


def run_batch(s2t, steps):
    with s2t.Batch(BATCH_SIZES) as b:
        for i in range(len(BATCH_SIZES)):
            b.configure(i, **FORCED)
            b.seed_spectrum(i, k0=14.0, amplitude=6.0, seed=20 + i)
        b.step(steps)
        return [b.read_state(i) for i in range(len(BATCH_SIZES))]


def test_three_threads_match_alone(s2t):
    steps = 100
    alone = {"dye": run_dye(s2t, steps), "forced": run_forced(s2t, steps),
             "batch": run_batch(s2t, steps)}
    out, errors = {}, []
    start = threading.Barrier(3)

    def worker(name, fn):
        try:
            start.wait()
            out[name] = fn(s2t, steps)
        except Exception as e:  # reported below, in the main thread
            errors.append((name, repr(e)))

    threads = [threading.Thread(target=worker, args=("dye", run_dye)),
               threading.Thread(target=worker, args=("forced", run_forced)),
               threading.Thread(target=worker, args=("batch", run_batch))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=60)
    assert not any(t.is_alive() for t in threads), "a worker did not finish"
    assert not errors, errors
    w, dye = out["dye"]
    assert np.array_equal(w, alone["dye"][0]) and np.array_equal(dye, alone["dye"][1])
    w, e = out["forced"]
    assert np.array_equal(w, alone["forced"][0]) and e == alone["forced"][1]
    assert all(np.array_equal(a, b) for a, b in zip(out["batch"], alone["batch"]))
