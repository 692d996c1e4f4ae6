"""Closed forms and invariants that pin the FP64 scheme (reference/spectral2d.py:132-229).

A dropped term, a sign error or a transposed wavenumber fails at least one of these.
"""

import os

import numpy as np
import pytest

from oracle.spectral import TWO_PI, Scheme, grid_coords, nu_from_rate, random_field


def test_packed_equals_direct():
    s = Scheme(64)
    s.set_physical(random_field(64, k0=8, seed=1))
    direct = s.nonlinear_direct(s.w)
    assert np.max(np.abs(s.nonlinear(s.w) - direct)) / np.max(np.abs(direct)) < 1e-12


def test_incompressible_and_curl():
    s = Scheme(64)
    s.set_physical(random_field(64, k0=8, seed=1))
    u, v = s.velocity()
    g = s.g
    div = np.fft.ifft2(1j * g.kx * np.fft.fft2(u) + 1j * g.ky * np.fft.fft2(v)).real
    curl = np.fft.ifft2(1j * g.kx * np.fft.fft2(v) - 1j * g.ky * np.fft.fft2(u)).real
    assert np.max(np.abs(div)) < 1e-10
    assert np.max(np.abs(curl - s.physical())) < 1e-10


@pytest.mark.parametrize("beta", [0.0, 12.0])
def test_inviscid_invariants(beta):
    s = Scheme(128, beta=beta)
    s.set_physical(random_field(128, k0=10, seed=2))
    e0, z0 = s.energy(), s.enstrophy()
    for _ in range(100):
        s.step(dt=2e-3)
    assert abs(s.energy() / e0 - 1) < 1e-7
    assert abs(s.enstrophy() / z0 - 1) < 1e-7


def test_rossby_phase():
    n, kx, ky, beta = 64, 3, 2, 8.0
    s = Scheme(n, beta=beta)
    w = np.zeros((n, n), np.complex128)
    w[ky, kx] = n * n * 1e-7
    s.w = w
    w0 = s.w[ky, kx]
    for _ in range(250):
        s.step(dt=0.5 / 250)
    expect = beta * kx / (kx * kx + ky * ky) * 0.5
    got = np.angle(s.w[ky, kx] / w0)
    assert abs(np.angle(np.exp(1j * (got - expect)))) < 1e-9
    assert abs(abs(s.w[ky, kx] / w0) - 1) < 1e-9


def test_taylor_green():
    n, k = 64, 2
    s = Scheme(n, nu=nu_from_rate(1 / 40.0, n, 1), p=1)
    gx, gy = grid_coords(n)
    s.set_physical(4.0 * np.cos(k * gx) * np.cos(k * gy))
    w0 = s.w.copy()
    assert np.max(np.abs(s.nonlinear(w0))) / np.max(np.abs(w0)) < 1e-14
    for _ in range(300):
        s.step(dt=3.0 / 300)
    expect = np.exp(-s.nu * (2.0 * k * k) * 3.0)
    assert abs(abs(s.w[k, k] / w0[k, k]) / expect - 1) < 1e-9


def test_cfl_dt():
    s = Scheme(64, cfl=0.4, dt_max=0.05)
    assert s.cfl_dt() == 0.05  # at rest
    s.set_physical(random_field(64, k0=8, seed=1, amplitude=6.0))
    u, v = s.velocity()
    assert s.cfl_dt() == pytest.approx(min(0.05, 0.4 * TWO_PI / 64 / np.max(np.abs(u) + np.abs(v))))


def test_diagnostics_parseval():
    s = Scheme(64)
    s.set_physical(random_field(64, k0=8, seed=4, amplitude=3.0))
    omega = s.physical()
    u, v = s.velocity()
    assert s.enstrophy() == pytest.approx(0.5 * np.mean(omega ** 2), rel=1e-12)
    assert s.energy() == pytest.approx(0.5 * np.mean(u ** 2 + v ** 2), rel=1e-12)
    assert np.sum(s.spectrum()) == pytest.approx(s.energy(), rel=1e-12)
    assert s.zonal_mean_u() == pytest.approx(u.mean(axis=1), abs=1e-13)


@pytest.mark.skipif(not os.environ.get("S2T_SLOW"), reason="long oracle run (about a minute); set S2T_SLOW=1")
def test_decaying_slope():
    s = Scheme(256, nu=nu_from_rate(0.5, 256, 4), p=4, cfl=0.4, dt_max=0.05)
    s.set_physical(random_field(256, k0=24, seed=5, amplitude=6.0))
    z0 = s.enstrophy()
    while s.t < 12.0:
        s.step()
    e = s.spectrum()
    b = np.arange(e.size)
    band = (b >= 8) & (b <= 45) & (e > 0)
    slope = np.polyfit(np.log(b[band]), np.log(e[band]), 1)[0]
    assert -4.2 < slope < -2.2
    assert 1 - s.enstrophy() / z0 > 0.05
