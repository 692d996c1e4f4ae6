"""Column-group width tuning (s2t_column_widths / s2t_set_column_width / s2t_autotune and the
batch form): every compiled width gives the same state bit for bit, and tuning leaves the
solver's own state alone."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PHYS = dict(beta=8.0, drag=0.01, viscosity=1.0, hyper_p=4.0, force_amp=3.0, force_k=12.0,
            force_width=2.0, cfl=0.5, dt_min=1e-5, dt_max=0.02)


def run(s2t, n, width, steps=4):
    with s2t.Solver(n) as s:
        s.configure(**PHYS)
        s.seed_spectrum(k0=max(2.0, n / 32), amplitude=5.0, seed=4)
        s.set_column_width(width)
        s.step(steps)
        return s.read_state(), s.diagnostics()


@pytest.mark.parametrize("n", [128, 256, 512, 1024, 2048, 4096])
def test_every_width_is_bit_identical(s2t, n):
    widths = s2t.column_widths(n)
    ref, d0 = run(s2t, n, 0)
    for w in widths:
        got, d = run(s2t, n, w)
        assert np.array_equal(got, ref), (n, w)
        assert d["stepIndex"] == d0["stepIndex"] == 4 and d["dt"] == d0["dt"]


SPLATS = [dict(x=0.3, y=0.6, dx=0.02, dy=-0.01, radius=0.05, amp=400.0, color=(1, 0, 0), dyeAmp=0.5),
          dict(x=0.95, y=0.02, dx=-0.01, dy=0.03, radius=0.03, amp=250.0, color=(0, 1, 0),
               dyeAmp=0.2)]


@pytest.mark.parametrize("n", [512, 1024, 2048])
def test_pointer_forcing_every_width(s2t, n):
    """With pointer forcing (the stage-0 transform of the force field: a thread group of its own in
    the three-group field-parallel kernel, a transform ahead of P' and Q' in the two-group one)
    every compiled width gives the same state bit for bit."""
    res = {}
    for w in s2t.column_widths(n):
        with s2t.Solver(n) as s:
            s.configure(**PHYS)
            s.seed_spectrum(k0=max(2.0, n / 32), amplitude=5.0, seed=4)
            s.set_column_width(w)
            s.set_splats(SPLATS)
            s.step(3)
            res[w] = s.read_state()
    ref = res[s2t.column_widths(n)[0]]
    assert np.abs(ref).max() > 0
    for w, got in res.items():
        assert np.array_equal(got, ref), (n, w)


@pytest.mark.parametrize("n", [512, 1024])
@pytest.mark.parametrize("case", ["plain", "dye", "pointer"])
def test_row_kernels_bit_identical(s2t, n, case):
    """The field-parallel row kernel (three thread groups per row pair) and the one-group row
    kernel give the same state (and dye) bit for bit: paired stages, stage 0 with the dye velocity,
    and stage 0 with pointer forcing."""
    res = []
    for fp in (True, False):
        with s2t.Solver(n) as s:
            s.configure(**PHYS)
            s.seed_spectrum(k0=max(2.0, n / 32), amplitude=5.0, seed=4)
            s.set_row_kernel(fp)
            if case == "pointer":
                s.set_splats(SPLATS)
            if case == "dye":
                s.paint_dye("bands")
                for _ in range(3):
                    s.step(1, dye=True)
                res.append((s.read_state(), s.read_dye()))
            else:
                s.step(3)
                res.append((s.read_state(), None))
    assert np.abs(res[0][0]).max() > 0
    assert np.array_equal(res[0][0], res[1][0])
    if case == "dye":
        assert np.array_equal(res[0][1].view(np.uint16), res[1][1].view(np.uint16))


def test_bad_width_rejected(s2t):
    with s2t.Solver(256) as s:
        with pytest.raises(s2t.S2TError, match="not compiled"):
            s.set_column_width(3)
        s.set_column_width(0)


def test_autotune_keeps_state_and_results(s2t):
    n = 1024
    with s2t.Solver(n) as s:
        s.configure(**PHYS)
        s.seed_spectrum(k0=32.0, amplitude=5.0, seed=4)
        before = s.read_state()
        w, ms = s.autotune(steps=3)
        assert w in s2t.column_widths(n) and sorted(ms) == sorted(s2t.column_widths(n))
        assert all(v > 0 for v in ms.values()) and ms[w] == min(ms.values())
        assert np.array_equal(s.read_state(), before)
        assert s.diagnostics()["stepIndex"] == 0
        s.step(4)
        got = s.read_state()
    ref, _ = run(s2t, n, 0)
    assert np.array_equal(got, ref)


def test_batch_autotune_members_equal_plain_solvers(s2t):
    sizes = [256] * 6 + [1024] * 2
    with s2t.Batch(sizes) as b:
        for i, n in enumerate(sizes):
            b.configure(i, **PHYS)
            b.seed_spectrum(i, k0=max(2.0, n / 32), amplitude=5.0, seed=4)
        b.step(1)
        chosen = b.autotune(steps=3)
        assert set(chosen) == {256, 1024}
        assert all(chosen[n] in s2t.column_widths(n) for n in chosen)
        b.step(3)   # replayed graphs, captured again with the chosen widths
        got = [b.read_state(i) for i in (0, 6)]
    for g, n in zip(got, (256, 1024)):
        ref, _ = run(s2t, n, 0)
        assert np.array_equal(g, ref)


def test_batch_default_widths(s2t):
    """A size class of two or more members defaults to the width measured fastest for that many
    instances: 4 columns per workgroup once the class fills the chip (128 and 256 from two members,
    512 from sixteen, 1024 from eight, 2048 from four), 2 for two or three of 2048, else the size's
    single-instance default (the first listed), which a single member keeps too."""
    sizes = [512] * 2 + [1024] * 8 + [2048] * 2 + [128] * 3 + [256]
    with s2t.Batch(sizes) as b:
        assert b.column_widths() == {512: s2t.column_widths(512)[0], 1024: 4, 2048: 2, 128: 4,
                                     256: s2t.column_widths(256)[0]}
    with s2t.Batch([2048] * 4 + [512] * 16) as b:
        assert b.column_widths() == {2048: 4, 512: 4}
