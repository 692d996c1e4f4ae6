"""Configuration checks, kernel-kind profiling, and the two stage-0 row-kernel forms.

Stage 0 of a step transforms (omega, s = u + v, d = u - v) row pairs when no dye velocity is
needed, and (v pair, omega + i u per row) when it is; both give the same step to FP32 round-off
and the same CFL maximum |u| + |v| = max(|s|, |d|).
"""

import numpy as np
import pytest

from oracle.spectral import random_field

pytestmark = pytest.mark.gpu

PHYS = dict(beta=8.0, drag=0.01, viscosity=1.0, hyper_p=4.0, cfl=0.5, dt_min=1e-5, dt_max=0.02)


@pytest.mark.parametrize("bad, msg", [
    (dict(hyper_p=2.5), "hyper_p must be an integer"),
    (dict(hyper_p=0.0), "hyper_p must be an integer"),
    (dict(hyper_p=17.0), "hyper_p must be an integer"),
    (dict(drag=-0.1), "drag must be non-negative"),
    (dict(viscosity=-1.0), "viscosity must be non-negative"),
    (dict(beta=1e6, dt_max=0.02), "beta"),
    (dict(beta=100.0, fixed_dt=200.0), "beta"),
])
def test_parameter_ranges(s2t, bad, msg):
    with s2t.Solver(64) as s:
        with pytest.raises(s2t.S2TError, match=msg):
            s.configure(**dict(PHYS, **bad))
        s.configure(**dict(PHYS, hyper_p=16.0))  # the largest accepted order
        s.configure(**PHYS)


def test_profile_kinds(s2t):
    with s2t.Solver(128) as s:
        s.configure(**PHYS)
        s.seed_spectrum(k0=8.0, amplitude=5.0, seed=3)
        s.step(1)
        s.profile(True, kinds=[s2t.PROF_COLS])
        s.step(2)
        s.profile(False)
        cols, rows = s.profile_read(s2t.PROF_COLS), s.profile_read(s2t.PROF_ROWS)
        assert cols[0] == 6 and cols[1] > 0
        assert rows[0] == 0 and rows[1] == 0
        s.profile(True)
        s.step(1)
        s.profile(False)
        assert s.profile_read(s2t.PROF_ROWS)[0] == 3 and s.profile_read(s2t.PROF_COLS)[0] == 3
        # one step in four: steps 0 and 4 of 7, all three stages of each
        s.profile(True, kinds=[s2t.PROF_COLS], every=4)
        s.step(3)
        s.step(4)
        s.profile(False)
        assert s.profile_read(s2t.PROF_COLS)[0] == 6 and s.profile_read(s2t.PROF_ROWS)[0] == 0
        with pytest.raises(s2t.S2TError, match="period"):
            s.profile(True, every=0)
        s.profile(True)
        s.step(2)
        s.profile(False)
        assert s.profile_read(s2t.PROF_COLS)[0] == 6


@pytest.mark.parametrize("n", [128, 1024])
def test_stage0_forms_agree(s2t, n):
    """One adaptive step with the sum / difference stage 0 (no dye) and with the per-row stage 0
    (dye on): the same state to 2e-6 relative, and the same CFL speed (hence dt) to the round-off
    of two different FP32 transform paths: max(|s|, |d|) against max(|u| + |v|), each from
    log2(n)-deep FFTs, so the bound is 5 u log2(n) with the FP32 unit round-off u = 2^-24
    (3.0e-6 at n = 1024)."""
    w = random_field(n, k0=n / 16, seed=11, amplitude=6.0)
    out = {}
    for dye in (False, True):
        with s2t.Solver(n) as s:
            s.configure(**PHYS)
            s.seed_physical(w)
            s.step(1, dye=dye)
            d = s.diagnostics()
            out[dye] = (s.read_state(), d["dt"], d["maxSpeed"])
    (a, dta, spa), (b, dtb, spb) = out[False], out[True]
    tol = 5 * 2.0**-24 * np.log2(n)
    print("n=%d speed rel diff %.2e, dt rel diff %.2e (bound %.2e)"
          % (n, abs(spa / spb - 1), abs(dta / dtb - 1), tol))
    assert abs(spa / spb - 1) < tol
    assert abs(dta / dtb - 1) < tol
    err = np.max(np.abs(a - b)) / np.max(np.abs(b))
    print("n=%d stage-0 forms: speed %.9g vs %.9g, state rel diff %.2e" % (n, spa, spb, err))
    assert err < 2e-6


@pytest.mark.parametrize("n", [1024, 4096])
def test_steps_are_deterministic(s2t, n):
    """The same state stepped twice (stochastic forcing on, adaptive dt) gives bit-identical
    states and clocks: every reduction in the step is order-independent (max, fixed-order sums),
    so any difference means a race or a hazard in a kernel."""
    params = dict(beta=10.0, drag=0.01, viscosity=1.0, hyper_p=4.0, force_amp=4.0, force_k=32.0,
                  force_width=2.0, cfl=0.5, dt_min=1e-5, dt_max=0.02)
    runs = []
    for _ in range(2):
        with s2t.Solver(n) as s:
            s.configure(**params)
            s.seed_spectrum(k0=n / 64, amplitude=6.0, seed=21)
            s.step(3)
            d = s.diagnostics()
            runs.append((s.read_state(), d["dt"], d["time"], d["maxSpeed"]))
    (a, *ca), (b, *cb) = runs
    assert ca == cb
    assert np.array_equal(a, b)


@pytest.mark.parametrize("n", [128, 1024])
def test_seed_bandlimited_stores_the_host_spectrum(s2t, n):
    """s2t_seed_bandlimited stores exactly s2t_bandlimited_spectrum (every mode is in band) and
    resets the clock like s2t_seed_spectrum."""
    want = s2t.bandlimited_spectrum(n, 8.0, 16.0, 6.0, seed=2)
    with s2t.Solver(n) as s:
        s.configure(**PHYS)
        s.seed_spectrum(k0=8.0, amplitude=5.0, seed=1)
        s.step(2)
        s.seed_bandlimited(8.0, 16.0, 6.0, seed=2)
        assert np.array_equal(s.read_state(), want)
        d = s.diagnostics()
        assert d["stepIndex"] == 0 and d["time"] == 0.0
        s.step(1)
        assert s.diagnostics()["stepIndex"] == 1


def test_checkpoint_resumes_a_forced_run(s2t):
    """A checkpoint is the state (s2t_read_state) and the clock (time and step index from the
    diagnostics); written into a new solver with s2t_write_state and s2t_set_clock, the run goes
    on with the same stochastic forcing phases and CFL steps as the run that never stopped."""
    n = 256
    phys = dict(PHYS, force_amp=4.0, force_k=20.0, force_width=2.5)
    with s2t.Solver(n) as a:
        a.configure(**phys)
        a.seed_spectrum(k0=12.0, amplitude=5.0, seed=21)
        a.step(4)
        ck_state, ck = a.read_state(), a.diagnostics()
        a.step(6)
        want, dwant = a.read_state(), a.diagnostics()
    with s2t.Solver(n) as b:
        b.configure(**phys)
        b.write_state(ck_state)
        b.set_clock(ck["time"], ck["stepIndex"])
        d = b.diagnostics()
        assert d["stepIndex"] == 4 and d["time"] == ck["time"]
        b.step(6)
        got, dgot = b.read_state(), b.diagnostics()
    assert dgot["stepIndex"] == dwant["stepIndex"] == 10
    assert dgot["time"] == pytest.approx(dwant["time"], rel=1e-12)
    # the resumed run re-primes its velocity planes from the written state with the prime
    # kernel, which the uninterrupted run took from its last column kernel: FP32 round-off only
    assert np.max(np.abs(got - want)) <= 1e-6 * np.max(np.abs(want))
    # without the clock the forcing phases restart at step 0 and the runs part
    with s2t.Solver(n) as c:
        c.configure(**phys)
        c.write_state(ck_state)
        c.step(6)
        assert np.max(np.abs(c.read_state() - want)) > 1e-3 * np.max(np.abs(want))


def test_zero_and_negative_step_counts(s2t):
    """step(0) changes nothing (state, clock, diagnostics); a negative count is an error."""
    with s2t.Solver(128) as s:
        s.configure(**PHYS)
        s.seed_spectrum(k0=8.0, amplitude=5.0, seed=3)
        s.step(2)
        before, d0 = s.read_state(), s.diagnostics()
        s.step(0)
        assert np.array_equal(s.read_state(), before)
        d1 = s.diagnostics()
        assert (d1["stepIndex"], d1["time"], d1["energy"]) == (d0["stepIndex"], d0["time"],
                                                               d0["energy"])
        with pytest.raises(s2t.S2TError, match="nsteps"):
            s.step(-1)
