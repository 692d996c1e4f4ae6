"""Pins for the oracle's restatement of GPU-only semantics (oracle/gpusem.py).

None of these functions has an FP64 counterpart in the reference; they are pinned by known
answers from independent integer arithmetic and by what the mathematics fixes.
"""

import numpy as np
import pytest

from oracle import gpusem as gs
from oracle.spectral import TWO_PI, Grid, Scheme

# hash1 known answers computed with C uint32 arithmetic (gcc), independent of numpy's wrap-around
HASH_KAT = [(0, 129708002), (1, 2831084092), (2, 2055130248), (7, 2120684060),
            (1000, 2163870295), (123456789, 4272394698), (4294967295, 3861530882),
            (2891336453, 582399676)]


def test_hash_known_answers():
    for x, h in HASH_KAT:
        assert int(gs.hash1(np.uint32(x))) == h
    assert int(gs.hash1(np.uint32(5) ^ gs.hash1(np.uint32(3)))) == 1861624210


def test_hashf_uniform():
    v = gs.hashf(np.arange(200000, dtype=np.uint32), np.uint32(17)).astype(np.float64)
    assert v.min() >= 0 and v.max() <= 1.0
    hist, _ = np.histogram(v, bins=20, range=(0, 1))
    assert np.all(np.abs(hist / 10000.0 - 1) < 0.05)


def test_gpu_dt():
    assert gs.gpu_dt(0.0, 0.5, 0.1, 1e-5, 0.02) == 0.02
    assert gs.gpu_dt(float("nan"), 0.5, 0.1, 1e-5, 0.02) == 0.02
    assert gs.gpu_dt(100.0, 0.5, 0.1, 1e-5, 0.02) == pytest.approx(5e-4)
    assert gs.gpu_dt(1e9, 0.5, 0.1, 1e-5, 0.02) == 1e-5


def test_stochastic_forcing_is_hermitian_and_in_annulus():
    n, kf, width = 64, 10.0, 2.0
    f = gs.stochastic_forcing(n, 3.0, kf, width, 12, 0.01)
    g = Grid(n)
    m = g.mirror
    assert np.max(np.abs(f - np.conj(f[np.ix_(m, m)]))) < 1e-9 * np.max(np.abs(f))
    kr = np.sqrt(g.k2)
    nz = np.abs(f) > 0
    assert nz.sum() > 0
    assert np.all(np.abs(kr[nz] - kf) < width * np.sqrt(np.log(1e4)) + 1e-9)
    assert np.all(g.band[nz])
    # real physical field
    assert np.max(np.abs(np.fft.ifft2(f).imag)) < 1e-9 * np.max(np.abs(np.fft.ifft2(f).real))


def test_stochastic_forcing_scaling():
    n = 64
    f1 = gs.stochastic_forcing(n, 2.0, 10.0, 2.0, 5, 0.01)
    f2 = gs.stochastic_forcing(n, 2.0, 10.0, 2.0, 5, 0.04)
    f3 = gs.stochastic_forcing(n, 4.0, 10.0, 2.0, 5, 0.01)
    assert np.allclose(f2, f1 / 2.0)
    assert np.allclose(f3, 2.0 * f1)
    f4 = gs.stochastic_forcing(n, 2.0, 10.0, 2.0, 6, 0.01)
    assert np.allclose(np.abs(f4), np.abs(f1)) and not np.allclose(f4, f1)


def test_injection_rate_closed_form():
    """Energy injected from rest by forcing alone matches 1/2 (0.01 A)^2 sum env^2/k^2."""
    n, amp, kf, width, dt = 64, 3.0, 10.0, 2.0, 0.01
    s = Scheme(n)
    g = s.g
    e = []
    for step in range(400):
        w = np.zeros((n, n), np.complex128)
        w += dt * gs.stochastic_forcing(n, amp, kf, width, step, dt)
        e.append(0.5 * np.sum(np.abs(w) ** 2 * g.inv_k2) / n ** 4)
    rate = np.mean(e) / dt
    assert rate == pytest.approx(gs.injection_rate(n, amp, kf, width), rel=0.1)
    assert gs.amp_for_rate(gs.injection_rate(n, amp, kf, width), n, kf, width) == pytest.approx(amp)


def test_pointer_force_is_curl_of_gaussian_force():
    """Closed-form curl vs a centred finite difference of the body force on a fine grid."""
    n = 512
    sp = dict(x=0.4, y=0.55, dx=0.03, dy=-0.02, radius=0.05, amp=7.0)
    got = gs.pointer_force_physical(n, [sp])
    c = (np.arange(n) + 0.5) / n
    y, x = np.meshgrid(c, c, indexing="ij")
    rx = gs._sep(x, sp["x"])
    ry = gs._sep(y, sp["y"])
    g = np.exp(-(rx ** 2 + ry ** 2) / sp["radius"] ** 2)
    fx, fy = sp["amp"] * g * sp["dx"], sp["amp"] * g * sp["dy"]
    h = 1.0 / n
    curl = (np.roll(fy, -1, 1) - np.roll(fy, 1, 1)) / (2 * h) - (np.roll(fx, -1, 0) - np.roll(fx, 1, 0)) / (2 * h)
    inside = rx ** 2 + ry ** 2 < 4 * sp["radius"] ** 2
    assert np.max(np.abs(got[inside] - curl[inside])) / np.max(np.abs(curl)) < 5e-3
    flipped = gs.pointer_force_physical(n, [dict(sp, dx=-sp["dx"], dy=-sp["dy"])])
    assert np.allclose(flipped, -got)
    assert abs(got.mean()) < 1e-6 * np.max(np.abs(got))


def test_fill_spectrum():
    n = 64
    w = gs.fill_spectrum(n, k0=8.0, amplitude=6.0, seed=3).astype(np.complex128)
    g = Grid(n)
    m = g.mirror
    assert np.max(np.abs(w - np.conj(w[np.ix_(m, m)]))) < 1e-5 * np.max(np.abs(w))
    assert np.all(w[~g.band] == 0) and w[0, 0] == 0
    omega = np.fft.ifft2(w).real
    assert np.sqrt(np.mean(omega ** 2)) == pytest.approx(6.0, rel=1e-5)
    w2 = gs.fill_spectrum(n, k0=8.0, amplitude=6.0, seed=4)
    assert np.allclose(np.abs(w2), np.abs(w), rtol=1e-5) and not np.allclose(w2, w)


def test_scenarios_match_formulas():
    n = 64
    x = np.arange(n) * TWO_PI / n
    tg = gs.scenario_field("taylor-green", n)
    assert tg[5, 7] == pytest.approx(6 * np.cos(2 * x[7]) * np.cos(2 * x[5]))
    dip = gs.scenario_field("dipole", n)
    assert abs(dip.sum()) < 1e-6 * np.abs(dip).sum()  # equal and opposite cores
    mer = gs.scenario_field("merger", n, {"mergerSep": 1.0})
    assert mer.max() == pytest.approx(30.0, rel=0.1) and mer.min() >= 0
    with pytest.raises(KeyError):
        gs.scenario_field("nope", n)


def test_to_half_truncates():
    assert int(gs.to_half_truncating(np.float32(1.0))) == 0x3C00
    # 1 + 2^-11 + 2^-12 rounds up under RNE but truncates here
    v = np.float32(1.0 + 2.0 ** -11 + 2.0 ** -12)
    assert int(gs.to_half_truncating(v)) == 0x3C00
    assert int(np.float16(v).view(np.uint16)) == 0x3C01
    assert int(gs.to_half_truncating(np.float32(-2.5))) == int(np.float16(-2.5).view(np.uint16))


def _dye(m, seed=0):
    rng = np.random.default_rng(seed)
    d = np.ones((m, m, 4), np.float16)
    d[..., :3] = rng.uniform(0, 1, (m, m, 3)).astype(np.float16)
    return d


def test_dye_invariants():
    m, n = 32, 16
    d = _dye(m)
    zero = np.zeros((n, n, 2))
    assert np.array_equal(gs.dye_step(d, zero, 0.01), d)
    const = np.full((m, m, 4), 0.5, np.float16)
    const[..., 3] = 1
    vel = np.random.default_rng(1).normal(size=(n, n, 2))
    assert np.allclose(gs.dye_step(const, vel, 0.01).astype(float), const.astype(float))
    # uniform velocity moving exactly 3 texels in x: dt/2pi * u * m = 3
    dt = 0.05
    u = 3.0 * TWO_PI / (dt * m)
    vel = np.zeros((n, n, 2))
    vel[..., 0] = u
    out = gs.dye_step(d, vel, dt)
    assert np.allclose(out.astype(float), np.roll(d, 3, axis=1).astype(float), atol=1e-3)
    # limiter: output within the range of the input
    vel = np.random.default_rng(2).normal(size=(n, n, 2)) * 20
    out = gs.dye_step(d, vel, 0.01).astype(float)
    assert out[..., :3].min() >= d[..., :3].astype(float).min() - 1e-3
    assert out[..., :3].max() <= d[..., :3].astype(float).max() + 1e-3


@pytest.mark.parametrize("kw", [8, 12])
def test_dye_bfecc_half_texel_closed_form(kw):
    """A uniform shift by half a texel of a cosine dye field: bilinear filtering multiplies the
    mode exp(i theta j) by A = 1 - a + a exp(-i theta) (forward trace, a = 1/2) and the back trace
    by conj(A), so BFECC gives A (3/2 - |A|^2 / 2) times the mode, then the fade factor.
    Compared where the limiter cannot bind (the closed-form value inside the range of the two
    texels around the trace foot); pins the 3/2, -1/2 blend, both trace directions, h = dt / 2 pi
    and the fade, which the integer shift (A = a pure shift) cannot tell apart."""
    m, n, dt, fade, amp = 64, 16, 0.05, 0.9, 0.25
    theta = 2 * np.pi * kw / m
    j = np.arange(m)
    phi0 = np.empty((m, m, 4), np.float16)
    for ch, psi in enumerate((0.0, 0.7, 1.9)):
        phi0[..., ch] = (0.5 + amp * np.cos(theta * j + psi))[None, :]
    phi0[..., 3] = 1
    vel = np.zeros((n, n, 2))
    vel[..., 0] = 0.5 * TWO_PI / (dt * m)      # h u m = 1/2 texel
    out = gs.dye_step(phi0, vel, dt, fade=fade).astype(np.float64)
    x = phi0.astype(np.float64)
    th = 2 * np.pi * np.fft.fftfreq(m, 1.0 / m) / m
    Ak = 0.5 + 0.5 * np.exp(-1j * th)
    for ch in range(3):
        # the stored field is the fp16 rounding of the cosine: apply the symbol to all its modes
        spec = np.fft.fft(x[0, :, ch])
        want = np.real(np.fft.ifft(spec * Ak * (1.5 - 0.5 * abs(Ak) ** 2))) * fade
        lo = np.minimum(x[0, :, ch], np.roll(x[0, :, ch], 1)) * fade
        hi = np.maximum(x[0, :, ch], np.roll(x[0, :, ch], 1)) * fade
        free = (want > lo + 2e-3) & (want < hi - 2e-3)
        assert free.sum() > m // 2
        got = out[:, :, ch]
        # fp16 intermediates: phi1 and phihat are stored at half precision, the output too
        assert np.max(np.abs(got[:, free] - want[None, free])) < 1.5e-3
        # the blend is visible: plain semi-Lagrangian (gain A alone) misses by far more
        plain = np.real(np.fft.ifft(spec * Ak)) * fade
        assert np.max(np.abs(plain[free] - want[free])) > 5e-3
    assert np.all(out[..., 3] == 1)


@pytest.mark.parametrize("cx, cy", [(0.3, 0.6), (0.98, 0.01)])
def test_dye_splat_gaussian_integral(cx, cy):
    """A dye splat on a black field at rest: its mass is the Gaussian integral dyeAmp pi sigma^2
    times the colour (cut off at r = 3 sigma, a factor 1 - e^-9 that this bound does not resolve),
    it peaks at the texel nearest the centre with dyeAmp exp(-r^2 / sigma^2) there, and it wraps
    periodically (a splat at a corner keeps its mass). Pins the width (sigma^2, not 2 sigma^2),
    the amplitude, the colour and the periodic separation."""
    m, n, sig, amp = 256, 16, 0.04, 0.8
    colour = (1.0, 0.5, 0.25)
    d = np.zeros((m, m, 4), np.float16)
    d[..., 3] = 1
    sp = dict(x=cx, y=cy, radius=sig, dyeAmp=amp, color=colour)
    out = gs.dye_step(d, np.zeros((n, n, 2)), 0.01, fade=1.0, splats=[sp]).astype(np.float64)
    mass = amp * np.pi * sig ** 2 * (1 - np.exp(-9.0))
    for ch, c in enumerate(colour):
        assert abs(out[..., ch].sum() / m ** 2 / (c * mass) - 1) < 2e-3
    iy, ix = np.unravel_index(np.argmax(out[..., 0]), (m, m))
    r2 = (((ix + 0.5) / m - cx + 0.5) % 1 - 0.5) ** 2 + (((iy + 0.5) / m - cy + 0.5) % 1 - 0.5) ** 2
    assert r2 <= 0.5 / m ** 2
    assert abs(out[iy, ix, 0] - amp * np.exp(-r2 / sig ** 2)) < 1e-3
    assert np.all(out[..., 3] == 1)


# ---------------------------------------------------------------------------------------------
# band-limited random seed (s2t_bandlimited_spectrum)
# ---------------------------------------------------------------------------------------------
def test_bandlimited_spectrum_invariants():
    n, kmin, kmax, amp = 256, 8.0, 16.0, 6.0
    w = gs.bandlimited_spectrum(n, kmin, kmax, amp, seed=2).astype(np.complex128)
    g = Grid(n)
    kr = np.sqrt(g.k2)
    support = g.band & (g.k2 > 0) & (kr >= kmin) & (kr <= kmax)
    assert np.all(w[~support] == 0) and np.all(w[support] != 0)
    assert w[0, 0] == 0
    # Hermitian: w(-k) = conj w(k), so the physical field is real
    assert np.array_equal(w[g.mirror][:, g.mirror], np.conj(w))
    phys = np.fft.ifft2(w)
    assert np.max(np.abs(phys.imag)) < 1e-12 * np.max(np.abs(phys.real))
    assert abs(np.sqrt(np.mean(phys.real ** 2)) / amp - 1) < 1e-6
    # determinism and seed dependence
    assert np.array_equal(gs.bandlimited_spectrum(n, kmin, kmax, amp, seed=2), w.astype(np.complex64))
    assert not np.array_equal(gs.bandlimited_spectrum(n, kmin, kmax, amp, seed=3), w.astype(np.complex64))


def test_bandlimited_amplitudes_are_gaussian():
    """|value| = |g| for a standard normal g: E|g| / sqrt(E g^2) = sqrt(2/pi) = 0.798 (a uniform
    amplitude would give 0.866, a constant one 1). About 2,500 canonical modes here."""
    n = 256
    w = gs.bandlimited_spectrum(n, 0.0, 40.0, 1.0, seed=7)
    g = Grid(n)
    canon = ((g.ky > 0) | ((g.ky == 0) & (g.kx > 0))) & (w != 0)
    a = np.abs(w[canon]).astype(np.float64)
    assert a.size > 2000
    ratio = np.mean(a) / np.sqrt(np.mean(a * a))
    assert abs(ratio - np.sqrt(2 / np.pi)) < 0.02
