"""Pins for the oracle functions of the GPU-only semantics that no reference fixture covers:
the stochastic forcing phase composition, the fillSpectrum seeding, the scenario geometry and the
'bands' dye pattern (oracle/gpusem.py).

Each pin is something the mathematics fixes or a known answer computed by independent code
(tests/golden/hash_kat.c, plain C uint32 arithmetic), chosen so that a plausible misreading of
the reference fails it. The misreadings are planted in the oracle itself (on the CPU, through
monkeypatch) and each planted copy is shown to fail.
"""

import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from oracle import gpusem as gs
from oracle.spectral import TWO_PI, Grid

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# Output of tests/golden/hash_kat.c: (row, col, sign of the sine, phase of the canonical mode).
# The forcing phase is TWO_PI * hashf(idx, stepIndex) in f32; fillSpectrum's is FP64.
FORCE_KAT = [  # (row, col, sign, f32 phase): N = 64, stepIndex = 3
    (2, 9, 1, 5.45448112),
    (2, 55, 1, 3.75060368),
    (62, 9, -1, 3.75060368),
    (0, 10, 1, 0.394328386),
    (0, 54, -1, 0.394328386),
    (5, 61, 1, 1.2775768),
]
FILL_KAT = [  # (row, col, sign, FP64 phase): n = 32, seed = 5
    (3, 4, 1, 2.3422795981698106),
    (3, 28, 1, 1.8553166380936141),
    (29, 4, -1, 1.8553166380936141),
    (1, 2, 1, 0.66788076620043701),
    (0, 7, 1, 4.4324711424888914),
    (0, 25, -1, 4.4324711424888914),
]


def _wrap(a):
    return (np.asarray(a) + np.pi) % TWO_PI - np.pi


def test_hash_kat_generator_agrees():
    """The constants above are what tests/golden/hash_kat.c prints (rebuilt with gcc here)."""
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    exe = os.path.join(os.environ.get("TMPDIR", "/tmp"), "s2t_hash_kat_%d" % os.getpid())
    subprocess.run([gcc, "-O0", "-o", exe, os.path.join(GOLDEN, "hash_kat.c"), "-lm"], check=True)
    try:
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    finally:
        os.remove(exe)
    rows = [tuple(float(x) for x in m) for m in
            re.findall(r"\((-?\d+), (-?\d+), (-?\d+), ([0-9.e+-]+)\)", out)]
    want = [tuple(float(x) for x in r) for r in FORCE_KAT + FILL_KAT]
    assert len(rows) == len(want)
    for got, w in zip(rows, want):
        assert got[:3] == w[:3]
        assert abs(got[3] - w[3]) <= 1e-15 * max(1.0, abs(w[3]))


# ---------------------------------------------------------------------------------------------
# stochastic forcing (solver.wgsl:77-92, common.wgsl:76-84)
# ---------------------------------------------------------------------------------------------
def _forcing_kat_error(n=64, step=3):
    """Largest phase error of the oracle's forcing at the FORCE_KAT modes.

    Derivation of one entry by hand (mode row 2, col 9 of N = 64 at stepIndex 3): kx = 9, ky = 2
    > 0, so the mode is canonical and idx = 2 * 64 + 9 = 137; hash1(3) = 2131687100 (the C
    known answer); hash1(137 ^ 2131687100) = 3728493940; hashf = 3728493940 * 2^-32 =
    0.868107617 (f32); phase = f32(2 pi) * 0.868107617 = 5.45448112. The value is
    mag (cos phase, sign sin phase). Mode (2, 55) has kx = -9 < 0 with ky = 2 > 0: canonical
    under ky-first, so it hashes its own index 183; mode (62, 9) is its mirror and takes
    idx 183 with sign -1. Row 0 modes are canonical for kx > 0 only.
    """
    f = gs.stochastic_forcing(n, 2.0, 9.0, 3.0, step, 0.01)
    err = 0.0
    for row, col, sgn, ph in FORCE_KAT:
        v = f[row, col]
        assert abs(v) > 0, (row, col)
        err = max(err, abs(_wrap(np.angle(v) - sgn * ph)))
    return err


def _annulus_phases(step, n=256, kf=32.0, width=2.0):
    """Phases of the canonical annulus modes of the oracle's forcing at one step."""
    f = gs.stochastic_forcing(n, 1.0, kf, width, step, 0.01)
    g = Grid(n)
    sel = (np.abs(f) > 0) & gs.canonical_half(g.kx, g.ky)
    return np.angle(f[sel])


def _forcing_statistics():
    """(|<e^{i phi}>|, |<e^{2 i phi}>|, max |<e^{i(phi_s - phi_{s+l})}>| for l = 1, 2), and the
    1/sqrt(count) scale of a mean of unit phasors."""
    ph = [_annulus_phases(s) for s in range(4)]
    cnt = ph[0].size
    m1 = max(abs(np.mean(np.exp(1j * p))) for p in ph)
    m2 = max(abs(np.mean(np.exp(2j * p))) for p in ph)
    lag = max(abs(np.mean(np.exp(1j * (ph[s] - ph[s + l])))) for l in (1, 2) for s in range(4 - l))
    return m1, m2, lag, 1.0 / np.sqrt(cnt)


def test_forcing_known_answers():
    assert _forcing_kat_error() < 2e-6


def test_forcing_phase_statistics():
    """A uniform phase on [0, 2 pi) has zero mean e^{i phi} and e^{2 i phi}, and fresh phases every
    step have zero step-to-step correlation: each mean is O(1/sqrt(count)) over ~600 modes."""
    m1, m2, lag, scale = _forcing_statistics()
    assert scale < 0.05
    assert m1 < 4 * scale and m2 < 4 * scale and lag < 4 * scale


def test_forcing_misreadings_fail(monkeypatch):
    """Each plausible misreading of the phase, planted in the oracle, fails a pin above."""
    hashf, hash1 = gs.hashf, gs.hash1
    # (a) phase pi u instead of 2 pi u: <e^{i phi}> = 2i/pi, and the known answers move
    monkeypatch.setattr(gs, "force_phase", lambda idx, s: np.float64(np.pi) * hashf(idx, s))
    assert _forcing_kat_error() > 0.1
    assert _forcing_statistics()[0] > 0.5
    monkeypatch.undo()
    # (b) hash1(idx ^ step) instead of hash1(idx ^ hash1(step)): statistically indistinguishable,
    # but the known answers move
    monkeypatch.setattr(gs, "force_phase", lambda idx, s: gs.F32_TWO_PI.astype(np.float64) * (
        hash1(np.asarray(idx, np.uint32) ^ np.uint32(s)).astype(np.float64) * 2.0 ** -32))
    assert _forcing_kat_error() > 0.1
    monkeypatch.undo()
    # (c) hash1(idx) ^ step (hash, then xor): phases change only in their low bits step to step
    monkeypatch.setattr(gs, "force_phase", lambda idx, s: TWO_PI * (
        (hash1(np.asarray(idx, np.uint32)) ^ np.uint32(s)).astype(np.float64) * 2.0 ** -32))
    assert _forcing_statistics()[2] > 0.9
    monkeypatch.undo()
    # (d) kx-first canonical half (kx > 0, or kx = 0 and ky > 0): Hermitian too, but the known
    # answers of the modes with ky > 0 > kx move
    monkeypatch.setattr(gs, "canonical_half", lambda kx, ky: (kx > 0) | ((kx == 0) & (ky > 0)))
    assert _forcing_kat_error() > 0.1
    monkeypatch.undo()
    assert _forcing_kat_error() < 2e-6


# ---------------------------------------------------------------------------------------------
# fillSpectrum (src/gpu/solver.js:384-418)
# ---------------------------------------------------------------------------------------------
def _fill_errors(n=32, k0=8.0, seed=5):
    """(phase error at FILL_KAT, magnitude-ratio error) of the oracle's fillSpectrum.

    Derivation of one entry by hand (mode row 3, col 4 of n = 32, seed 5): ky = 3 > 0, canonical,
    idx = 3 * 32 + 4 = 100; salt = hash32(5) = 2161170183; hash32(100 ^ 2161170183) =
    1601100999; phase = 2 pi * 1601100999 / 2^32 = 2.3422795981698106. Magnitudes before the
    common rms scale are sqrt(k^2 / (1 + (|k|/k0)^8)), so |w(k1)| / |w(k2)| is
    sqrt(env(k1) / env(k2)) whatever the scale.
    """
    w = gs.fill_spectrum(n, k0=k0, amplitude=6.0, seed=seed).astype(np.complex128)
    perr = max(abs(_wrap(np.angle(w[r, c]) - s * ph)) for r, c, s, ph in FILL_KAT)

    def env(kx, ky):
        k2 = kx * kx + ky * ky
        return k2 / (1 + (np.sqrt(k2) / k0) ** 8)

    # (3, 4): |k| = 5; (1, 2): |k| = sqrt 5; (0, 7): |k| = 7; (6, 9): kx = 9, ky = 6
    pairs = [((3, 4), (1, 2)), ((0, 7), (3, 4)), ((6, 9), (0, 7))]
    rerr = 0.0
    for (r1, c1), (r2, c2) in pairs:
        got = abs(w[r1, c1]) / abs(w[r2, c2])
        want = np.sqrt(env(c1, r1) / env(c2, r2))
        rerr = max(rerr, abs(got / want - 1))
    return perr, rerr


def test_fill_spectrum_known_answers():
    perr, rerr = _fill_errors()
    assert perr < 1e-6 and rerr < 1e-6


def test_fill_spectrum_phase_statistics():
    n = 256
    w = gs.fill_spectrum(n, k0=40.0, amplitude=6.0, seed=9)
    g = Grid(n)
    sel = (np.abs(w) > 0) & gs.canonical_half(g.kx, g.ky)
    ph = np.angle(w[sel])
    scale = 1.0 / np.sqrt(ph.size)
    assert abs(np.mean(np.exp(1j * ph))) < 4 * scale
    assert abs(np.mean(np.exp(2j * ph))) < 4 * scale


def test_fill_spectrum_misreadings_fail(monkeypatch):
    hash1 = gs.hash1
    # (a) the seed as salt without hashing it
    monkeypatch.setattr(gs, "fill_phase", lambda idx, seed: TWO_PI * hash1(
        np.asarray(idx, np.uint32) ^ np.uint32(seed)).astype(np.float64) / 4294967296.0)
    assert _fill_errors()[0] > 0.1
    monkeypatch.undo()
    # (b) hash32(idx) ^ hash32(seed) (hash, then salt)
    monkeypatch.setattr(gs, "fill_phase", lambda idx, seed: TWO_PI * (
        hash1(np.asarray(idx, np.uint32)) ^ hash1(np.uint32(seed))).astype(np.float64)
        / 4294967296.0)
    assert _fill_errors()[0] > 0.1
    monkeypatch.undo()
    # (c) the envelope as the magnitude (no square root): the magnitude ratios move
    monkeypatch.setattr(gs, "fill_amplitude", lambda k2, k0: k2 / (1.0 + (np.sqrt(k2) / k0) ** 8))
    assert _fill_errors()[1] > 0.1
    monkeypatch.undo()
    assert _fill_errors()[0] < 1e-6


# ---------------------------------------------------------------------------------------------
# scenarios (src/scenarios.js:30-162): moments of the physical fields
# ---------------------------------------------------------------------------------------------
def _moments(f, n, mask=None):
    """Circulation sum(f) h^2 and centroid / second central moments of f over the grid."""
    h = TWO_PI / n
    y, x = np.meshgrid(np.arange(n) * h, np.arange(n) * h, indexing="ij")
    if mask is not None:
        f = np.where(mask, f, 0.0)
    gam = f.sum() * h * h
    cx = (f * x).sum() * h * h / gam
    cy = (f * y).sum() * h * h / gam
    vx = (f * (x - cx) ** 2).sum() * h * h / gam
    vy = (f * (y - cy) ** 2).sum() * h * h / gam
    return gam, cx, cy, vx, vy


def test_taylor_green_spectrum():
    """6 cos 2x cos 2y has exactly the four modes (+-2, +-2), each 6/4 N^2 (unscaled FFT)."""
    n = 64
    w = np.fft.fft2(gs.scenario_field("taylor-green", n))
    g = Grid(n)
    four = (np.abs(g.kx) == 2) & (np.abs(g.ky) == 2)
    assert np.allclose(w[four], 1.5 * n * n, rtol=1e-12)
    assert np.max(np.abs(w[~four])) < 1e-9 * n * n


def _dipole_error(f, n):
    """Largest deviation of the dipole field's moments from their closed forms.

    Cores of circulation +Gamma and -Gamma (Gamma = 12, a = 0.3) at x = 0.6 pi and
    y = pi -+ d/2 (d = 0.9), the positive one below. A Gaussian exp(-r^2/a^2) integrates to
    pi a^2 with variance s^2 = a^2/2 per axis (the cut at r = 5a changes that by e^-25), so with
    Y = y - pi, X = x - 0.6 pi the moments are: sum omega = 0; sum omega Y = -Gamma d;
    sum omega Y X = 0; sum omega X^2 Y = -Gamma d s^2; sum omega Y^3 = -2 Gamma ((d/2)^3 +
    3 (d/2) s^2). They pin the circulation, the separation, the sign, the column and the width
    (the cores overlap, so the two signs cannot be integrated apart).
    """
    gam, d, s2 = 12.0, 0.9, 0.3 ** 2 / 2
    h = TWO_PI / n
    y, x = np.meshgrid(np.arange(n) * h, np.arange(n) * h, indexing="ij")
    Y, X = y - np.pi, x - 0.6 * np.pi

    def mom(w):
        return float((f * w).sum() * h * h)

    want = [(mom(1.0), 0.0, 1.0), (mom(Y), -gam * d, gam * d), (mom(Y * X), 0.0, 1.0),
            (mom(X * X * Y), -gam * d * s2, gam * d * s2),
            (mom(Y ** 3), -2 * gam * ((d / 2) ** 3 + 3 * (d / 2) * s2), 1.0)]
    return max(abs(g - w) / sc for g, w, sc in want)


def test_dipole_moments():
    n = 256
    assert _dipole_error(gs.scenario_field("dipole", n), n) < 1e-9


def test_dipole_misreadings_fail(monkeypatch):
    """Planted in the oracle: the signs swapped (positive core above), the separation d read as
    the offset of each core, the cores on the column x = 0.6 instead of 0.6 pi."""
    n = 128
    assert _dipole_error(-gs.scenario_field("dipole", n), n) > 0.5
    blobs = gs.blobs
    for change in (lambda b: dict(b, y=np.pi + 2 * (b["y"] - np.pi)),
                   lambda b: dict(b, x=0.6)):
        monkeypatch.setattr(gs, "blobs", lambda m, items: blobs(m, [change(b) for b in items]))
        assert _dipole_error(gs.scenario_field("dipole", n), n) > 0.05
        monkeypatch.undo()


@pytest.mark.parametrize("sep", [0.5, 1.0, 1.6])
def test_merger_cores(sep):
    """Two equal cores (amplitude 30, a = 0.22) at (pi -+ d/2, pi): total circulation
    2 * 30 pi a^2, centroid (pi, pi), <(x - pi)^2> = d^2/4 + a^2/2, <(y - pi)^2> = a^2/2."""
    n = 256
    a = 0.22
    gam, cx, cy, vx, vy = _moments(gs.scenario_field("merger", n, {"mergerSep": sep}), n)
    assert gam == pytest.approx(2 * 30 * np.pi * a * a, rel=1e-6)
    assert cx == pytest.approx(np.pi, abs=1e-9) and cy == pytest.approx(np.pi, abs=1e-9)
    assert vx == pytest.approx(sep ** 2 / 4 + a * a / 2, rel=1e-5)
    assert vy == pytest.approx(a * a / 2, rel=1e-5)


@pytest.mark.parametrize("count", [3, 5, 6, 8])
def test_polygon_cores(count):
    """count cores (amplitude 34, a = 0.19) at angles 2 pi j / count around (pi, pi) on radii
    1.25 (1 + 0.02 cos(2 pi m j / count)), m = max(2, floor(count / 2)): each sector's core has
    circulation 34 pi a^2 and its centroid at that radius and angle."""
    n = 512
    a = 0.19
    f = gs.scenario_field("polygon", n, {"polygonN": count})
    h = TWO_PI / n
    y, x = np.meshgrid(np.arange(n) * h, np.arange(n) * h, indexing="ij")
    ang = np.mod(np.arctan2(y - np.pi, x - np.pi), TWO_PI)
    m = max(2, count // 2)
    for j in range(count):
        th = TWO_PI * j / count
        sector = np.abs(_wrap(ang - th)) < np.pi / count
        gam, cx, cy, _, _ = _moments(f, n, sector)
        r = 1.25 * (1 + 0.02 * np.cos(TWO_PI * m * j / count))
        tol = 2e-3 if count < 8 else 5e-3  # tails of the neighbouring cores
        assert gam == pytest.approx(34 * np.pi * a * a, rel=tol)
        assert np.hypot(cx - np.pi, cy - np.pi) == pytest.approx(r, abs=tol * 0.5)
        assert _wrap(np.arctan2(cy - np.pi, cx - np.pi) - th) == pytest.approx(0, abs=tol)


def test_shear_layer_profile():
    """omega = 0.05 cos x - du/dy with u = tanh((y - pi/2)/delta) below y = pi and its mirror
    above, delta = max(2 pi/30, 12 pi/N): the x-mean of omega integrates to -2 tanh(pi/(2 delta))
    over the lower half (the velocity jump of the layer) and +2 tanh(...) over the upper half;
    the sech^2 profile has second moment delta^2 pi^2 / 12; the kx = 1 perturbation is 0.05."""
    for n in (128, 512):
        f = gs.scenario_field("shear layer", n)
        h = TWO_PI / n
        delta = max(TWO_PI / 30, 6 * TWO_PI / n)
        y = np.arange(n) * h
        prof = f.mean(axis=1)
        low = y <= np.pi
        jump = np.tanh(np.pi / (2 * delta))
        assert prof[low].sum() * h == pytest.approx(-2 * jump, rel=2e-3)
        assert prof[~low].sum() * h == pytest.approx(2 * jump, rel=2e-3)
        m2 = (prof[low] * (y[low] - np.pi / 2) ** 2).sum() / prof[low].sum()
        assert m2 == pytest.approx(delta ** 2 * np.pi ** 2 / 12, rel=2e-3)
        pert = np.fft.fft(f, axis=1)[:, 1] / n
        assert np.allclose(pert, 0.025, atol=1e-12)


# ---------------------------------------------------------------------------------------------
# 'bands' dye pattern (src/gpu/solver.js:446-473)
# ---------------------------------------------------------------------------------------------
def _trunc_ok(got, exact):
    """binary16 by mantissa truncation: at most the exact value and within one half ulp below."""
    got = np.asarray(got, np.float64)
    exact = np.asarray(exact, np.float64)
    ulp = 2.0 ** (np.floor(np.log2(np.maximum(np.abs(exact), 2 ** -14))) - 10)
    return np.all(got <= exact + 1e-12) and np.all(got > exact - ulp - 1e-12)


def _bands_ok(d):
    """band(fy) = 1/2 + tanh(6 sin(12 pi fy))/2 is 1/2 where sin(12 pi fy) = 0 (rows 0, m/4,
    m/2), about 1 at fy = 3/8 and about 0 at fy = 1/8 (six bright bands, repeating every m/6
    rows); mix(fx) = 1/2 + sin(2 pi fx)/2 is 1 at fx = 1/4 (warm) and 0 at fx = 3/4 (cool).
    Each channel is 0.9 band (warm mix + cool (1 - mix)), alpha 1, truncated to binary16."""
    m = d.shape[0]
    d = d.astype(np.float64)
    warm = np.array([1.0, 0.45, 0.22])
    cool = np.array([0.28, 0.62, 1.0])
    hi = 0.5 + 0.5 * np.tanh(6.0)
    cases = [  # (row, col, expected rgb)
        (m // 4, m // 4, 0.9 * 0.5 * warm),
        (m // 2, 3 * m // 4, 0.9 * 0.5 * cool),
        (0, m // 2, 0.9 * 0.5 * 0.5 * (warm + cool)),
        (3 * m // 8, m // 4, 0.9 * hi * warm),
        (3 * m // 8, 3 * m // 4, 0.9 * hi * cool),
        (m // 8, m // 4, 0.9 * (1 - hi) * warm),
    ]
    ok = all(_trunc_ok(d[r, c, :3], rgb) for r, c, rgb in cases)
    ok = ok and bool(np.all(d[..., 3] == 1.0))
    # six bands: the row profile at the warm column crosses the half level 12 times
    bright = d[:, m // 4, 0] > 0.45
    return ok and np.count_nonzero(bright != np.roll(bright, 1)) == 12


def test_paint_bands_landmarks():
    assert _bands_ok(gs.paint_bands(256))


def test_paint_bands_misreadings_fail():
    """Planted in the oracle's output: x and y swapped (vertical bands), warm and cool swapped."""
    d = gs.paint_bands(256)
    assert not _bands_ok(d.transpose(1, 0, 2))
    assert not _bands_ok(np.roll(d, 128, axis=1))  # mix(fx + 1/2) = 1 - mix(fx)
