"""The line-FFT contract model (reference/fft_ref.py) against numpy.fft and brute force."""

import numpy as np
import pytest

from oracle.fftmodel import brute_dft, digit_reverse, dit_2d, dit_line, radix_schedule

SIZES = [16, 32, 64, 128, 256, 512, 1024]


@pytest.mark.parametrize("n", SIZES)
def test_schedule_and_bijection(n):
    r = radix_schedule(n)
    assert int(np.prod(r)) == n
    assert sorted(digit_reverse(i, r) for i in range(n)) == list(range(n))


def test_reversal_order_matters_at_radix2_sizes():
    """Reversal over the radix list in reverse stage order (README 'The FFT kernel')."""
    r = radix_schedule(128)
    wrong = [sum(((i // int(np.prod(r[:k]))) % r[k]) * int(np.prod(r[k + 1:])) for k in range(len(r)))
             for i in range(128)]
    right = [digit_reverse(i, r) for i in range(128)]
    assert wrong != right


def test_bad_sizes():
    for n in (3, 12, 24, 96, 1000):
        with pytest.raises(ValueError, match="must be 4\\^a or 2 \\* 4\\^a"):
            radix_schedule(n)


@pytest.mark.parametrize("n", SIZES)
def test_line_vs_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = np.fft.fft(x)
    assert np.max(np.abs(dit_line(x) - ref)) / np.max(np.abs(ref)) < 1e-13
    assert np.max(np.abs(dit_line(x, True) - np.fft.ifft(x))) * n / np.max(np.abs(ref)) < 1e-13


def test_brute_force_small():
    rng = np.random.default_rng(5)
    for n in (16, 32):
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        assert np.allclose(dit_line(x), brute_dft(x), atol=1e-12)
        assert np.allclose(np.fft.fft(x), brute_dft(x), atol=1e-12)


def test_2d_fixture(fixtures):
    from conftest import cplx, rel_err
    f = fixtures["fft"]
    x = cplx(f["input"], 32)
    assert rel_err(dit_2d(x), cplx(f["forward"], 32)) < 1e-13
    assert rel_err(dit_2d(x, True), cplx(f["inverse"], 32)) < 1e-13
