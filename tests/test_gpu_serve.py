"""Continuous-batching server (s2t_server_*): results alone and in the mix are bit-identical;
admission, rejection and latency statistics."""

import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PHYS = dict(beta=5.0, drag=0.01, viscosity=1.0, hyper_p=4.0, cfl=0.5, dt_min=1e-5, dt_max=0.02)


def test_alone_equals_in_mix(s2t):
    probe = dict(n=128, steps=40, scenario="merger", scenario_param=0.9, keep_state=True, **PHYS)
    with s2t.Server({128: 2, 256: 2}) as sv:
        rid = sv.submit(**probe)
        sv.drain(120)
        alone = sv.read_state(rid)
        st_alone, d_alone, _ = sv.poll(rid)
    with s2t.Server({64: 3, 128: 3, 256: 2}) as sv:
        others = []
        for k in range(6):
            others.append(sv.submit(n=[64, 128, 256][k % 3], steps=15 + 10 * k,
                                    scenario=["turbulence", "dipole", "taylor-green"][k % 3],
                                    seed=k + 3, **PHYS))
            if k == 2:
                rid = sv.submit(**probe)
            time.sleep(0.002)
        sv.drain(120)
        mixed = sv.read_state(rid)
        st_mix, d_mix, _ = sv.poll(rid)
        stats = sv.stats()
    assert st_alone == st_mix == s2t.REQ_DONE
    assert np.array_equal(alone, mixed)
    assert d_alone["stepIndex"] == 40 and d_mix["stepIndex"] == 40
    assert d_alone["energy"] == d_mix["energy"]
    assert stats["completed"] == 7 and stats["p99_latency_s"] >= stats["p50_latency_s"] > 0


def test_server_matches_plain_solver(s2t):
    n, steps = 64, 25
    with s2t.Server({n: 4}) as sv:
        rid = sv.submit(n=n, steps=steps, scenario="turbulence", k0=6.0, amplitude=5.0, seed=11,
                        keep_state=True, **PHYS)
        sv.drain(60)
        got = sv.read_state(rid)
    with s2t.Solver(n) as s:
        s.configure(**PHYS)
        s.seed_spectrum(k0=6.0, amplitude=5.0, seed=11)
        s.step(steps)
        assert np.array_equal(got, s.read_state())


@pytest.mark.parametrize("n, slots", [(2048, 2), (512, 16)])
def test_chunk_width_changes_keep_results(s2t, n, slots):
    """A chunk steps at the width measured fastest for its busy slots, chosen again at every poll
    (2048^2: two columns per workgroup while two slots are busy; 512^2: four columns with P_hat in
    LDS while all sixteen are, the field-parallel kernel once one has finished), so the kernels
    change under a running request: each result still equals its plain solver bit for bit."""
    reqs = [dict(n=n, steps=6 + 3 * (k % 4), scenario="turbulence", k0=n / 40, amplitude=5.0,
                 seed=30 + k, keep_state=True, **PHYS) for k in range(slots)]
    with s2t.Server({n: slots}) as sv:
        ids = [sv.submit(**r) for r in reqs]
        sv.drain(120)
        got = [sv.read_state(i) for i in ids]
    for r, g in zip(reqs, got):
        with s2t.Solver(n) as s:
            s.configure(**PHYS)
            s.seed_spectrum(k0=r["k0"], amplitude=5.0, seed=r["seed"])
            s.step(r["steps"])
            assert np.array_equal(g, s.read_state()), r["seed"]


def chunk_bytes(n, slots):
    """Device bytes the server charges for a chunk of `slots` instances of size n: 3 states and
    5 mixed planes per instance, 3 full fields and a state of scratch per chunk."""
    k = n // 3
    kc, nj = k + 1, 2 * k + 1
    state = kc * ((nj + 63) // 64 * 64) * 8
    plane = n * ((kc + 15) // 16 * 16) * 8
    return slots * (3 * state + 5 * plane) + n * n * 24 + state


def test_rejection_and_budget(s2t):
    """Unsupported sizes, and sizes of which not one instance fits the budget, are rejected; any
    other size is served whether or not a class was configured for it."""
    with s2t.Server({64: 2}, budget_bytes=1e6) as sv:
        big = sv.submit(n=2048, steps=5, **PHYS)
        odd = sv.submit(n=96, steps=5, **PHYS)
        other = sv.submit(n=128, steps=5, **PHYS)
        for rid in (big, odd):
            st, _, lat = sv.poll(rid)
            assert st == s2t.REQ_REJECTED and lat < 0
        sv.drain(60)
        assert sv.poll(other)[0] == s2t.REQ_DONE
        st = sv.stats()
        assert st["rejected"] == 2 and st["completed"] == 1 and st["peak_bytes"] <= 1e6
    with pytest.raises(s2t.S2TError, match="over the budget"):
        s2t.Server({2048: 8}, budget_bytes=1e6)


def test_admission_against_the_budget(s2t):
    """Chunks are allocated as requests need them and freed when idle to make room for another
    size: with a budget that holds one 4-slot chunk of 128^2, six 128^2 requests and three 64^2
    requests (64 has no configured class) all complete, the server never holds more than the
    budget, and results equal the plain solver's."""
    budget = 1.05 * chunk_bytes(128, 4)
    with s2t.Server({128: 4}, budget_bytes=budget) as sv:
        ids = [sv.submit(n=128, steps=20 + k, seed=k, keep_state=k == 0, **PHYS) for k in range(6)]
        ids += [sv.submit(n=64, steps=15 + k, seed=10 + k, keep_state=k == 0, **PHYS)
                for k in range(3)]
        sv.drain(120)
        for rid in ids:
            assert sv.poll(rid)[0] == s2t.REQ_DONE
        got = {128: sv.read_state(ids[0]), 64: sv.read_state(ids[6])}
        st = sv.stats()
    assert st["completed"] == 9 and 0 < st["peak_bytes"] <= budget
    for n, steps, seed in ((128, 20, 0), (64, 15, 10)):
        with s2t.Solver(n) as s:
            s.configure(**PHYS)
            s.seed_spectrum(k0=14.0, amplitude=6.0, seed=seed)
            s.step(steps)
            assert np.array_equal(got[n], s.read_state())


def test_queueing_beyond_slots(s2t):
    """With room for one 2-slot chunk, more requests than slots wait for slots to free and all
    complete."""
    with s2t.Server({64: 2}, budget_bytes=1.05 * chunk_bytes(64, 2)) as sv:
        ids = [sv.submit(n=64, steps=10 + k, seed=k, **PHYS) for k in range(7)]
        sv.drain(60)
        for rid in ids:
            st, d, lat = sv.poll(rid)
            assert st == s2t.REQ_DONE and d["stepIndex"] >= 10 and lat > 0


def test_cancel_and_completion_callbacks(s2t):
    """A queued request is dropped at once and a running one is stopped between steps, its slot
    going to the next request (whose result is still the plain solver's, bit for bit); every
    request reports its final status exactly once through the callback; a finished request
    cannot be cancelled."""
    import threading
    seen, lock = {}, threading.Lock()

    def on_final(rid, status):
        with lock:
            assert rid not in seen
            seen[rid] = status

    seed = dict(scenario="turbulence", k0=6.0, amplitude=5.0, seed=11)
    with s2t.Server({64: 1}, budget_bytes=1.05 * chunk_bytes(64, 1)) as sv:
        sv.set_callback(on_final)
        long_id = sv.submit(n=64, steps=10 ** 7, **seed, **PHYS)   # runs until cancelled
        queued = sv.submit(n=64, steps=10, **seed, **PHYS)        # waits for the one slot
        rejected = sv.submit(n=2048, steps=5, **PHYS)             # cannot fit in the budget
        t0 = time.time()
        while sv.poll(long_id)[0] != s2t.REQ_RUNNING:
            assert time.time() - t0 < 60
            time.sleep(0.005)
        sv.cancel(queued)
        assert sv.poll(queued)[0] == s2t.REQ_CANCELLED
        after = sv.submit(n=64, steps=7, keep_state=True, **seed, **PHYS)
        sv.cancel(long_id)
        sv.drain(120)
        assert sv.poll(long_id)[0] == s2t.REQ_CANCELLED
        st, d, lat = sv.poll(after)
        assert st == s2t.REQ_DONE and d["stepIndex"] == 7 and lat > 0
        got = sv.read_state(after)
        with pytest.raises(s2t.S2TError, match="already finished"):
            sv.cancel(after)
        stats = sv.stats()
    assert seen == {long_id: s2t.REQ_CANCELLED, queued: s2t.REQ_CANCELLED,
                    rejected: s2t.REQ_REJECTED, after: s2t.REQ_DONE}
    assert (stats["cancelled"], stats["completed"], stats["rejected"]) == (2, 1, 1)
    with s2t.Solver(64) as s:
        s.configure(**PHYS)
        s.seed_spectrum(k0=6.0, amplitude=5.0, seed=11)
        s.step(7)
        assert np.array_equal(got, s.read_state())


def test_worker_failure_fails_every_outstanding_request(s2t):
    """If the worker stops on an error (forced from the host: s2t_server_inject_failure(1) makes
    the second admission fail), every request still queued or running fails and is called back once,
    so no caller waits forever; the server then accepts nothing more and reports the error."""
    import threading
    seen, lock, done = {}, threading.Lock(), threading.Event()

    def on_final(rid, status):
        with lock:
            assert rid not in seen
            seen[rid] = status
            if len(seen) == 3:
                done.set()

    sv = s2t.Server({64: 1}, budget_bytes=1.05 * chunk_bytes(64, 1))
    try:
        sv.inject_failure(1)
        sv.set_callback(on_final)
        ids = [sv.submit(n=64, steps=20, seed=k, **PHYS) for k in range(3)]
        assert done.wait(60)
        # one slot: the first request runs to completion, the second admission fails
        assert [seen[i] for i in ids] == [s2t.REQ_DONE, s2t.REQ_FAILED, s2t.REQ_FAILED]
        with pytest.raises(s2t.S2TError, match="injected"):
            sv.submit(n=64, steps=5, **PHYS)
        with pytest.raises(s2t.S2TError, match="injected"):
            sv.drain(1)
    finally:
        sv.close()
