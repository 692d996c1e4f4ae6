"""The FP64 oracle against the reference's own golden vectors (tests/golden/fixtures.json).

Each check follows one key of reference/make_fixtures.py:19-76; the bound 1e-12 is far below the
reference harness tolerances (tests/index.html:136-201), so the oracle can serve as the checker.
"""

import numpy as np

from conftest import cplx, rel_err
from oracle.spectral import Scheme, nu_from_rate, random_field

N = 32


def fixture_scheme(fx):
    st = fx["step"]
    return Scheme(N, beta=st["beta"], mu=st["mu"], nu=nu_from_rate(1 / 1.5, N, st["hyperP"]),
                  p=st["hyperP"])


def test_fft(fixtures):
    f = fixtures["fft"]
    x = cplx(f["input"], N)
    assert rel_err(np.fft.fft2(x), cplx(f["forward"], N)) < 1e-12
    assert rel_err(np.fft.ifft2(x), cplx(f["inverse"], N)) < 1e-12


def test_nu_mapping(fixtures):
    """viscosity = 1/tau as a rate at kcut gives the fixture's nuH (tests/index.html:125,136)."""
    st = fixtures["step"]
    assert abs(nu_from_rate(1 / 1.5, N, 4) / st["nuH"] - 1) < 1e-14


def test_velocity_pack(fixtures):
    s = fixture_scheme(fixtures)
    s.set_physical(random_field(N, k0=6, seed=3))
    v = fixtures["velocity"]
    assert rel_err(s.w, cplx(v["what"], N)) < 1e-12
    a, b = s.packed_spectra(s.w)
    assert rel_err(a, cplx(v["packA"], N)) < 1e-12
    assert rel_err(b, cplx(v["packB"], N)) < 1e-12


def test_assemble(fixtures):
    s = fixture_scheme(fixtures)
    w = cplx(fixtures["velocity"]["what"], N)
    a = fixtures["assemble"]
    assert rel_err(s.product_transform(w), cplx(a["z"], N)) < 1e-12
    assert rel_err(-s.nhat_from_z(cplx(a["z"], N)), cplx(a["rhs"], N)) < 1e-12


def test_step_and_run20(fixtures):
    s = fixture_scheme(fixtures)
    st = fixtures["step"]
    s.w = cplx(st["before"], N)
    s.step(dt=st["dt"])
    assert rel_err(s.w, cplx(st["after"], N)) < 1e-12
    s.w = cplx(st["before"], N)
    for _ in range(20):
        s.step(dt=st["dt"])
    r = fixtures["run20"]
    assert rel_err(s.w, cplx(r["after"], N)) < 1e-12
    assert abs(s.energy() / r["energy"] - 1) < 1e-12
    assert abs(s.enstrophy() / r["enstrophy"] - 1) < 1e-12
