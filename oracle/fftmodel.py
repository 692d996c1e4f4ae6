"""Model of the reference's line-FFT contract (test infrastructure only).

* ``radix_schedule``: N must be 4^a or 2*4^a; radix-4 stages then an optional trailing radix-2
  (src/gpu/fft.js:20-31, reference/fft_ref.py:6-16). The error text follows fft.js:28.
* ``digit_reverse``: reversal taken over the radix list in reverse stage order
  (reference/fft_ref.py:19-24, README.md "The FFT kernel").
* ``dit_line``: the in-place radix-4 decimation-in-time transform of src/shaders/fft.wgsl:42-79
  with the radix-2 block of src/gpu/fft.js:4-14, forward unscaled, inverse scaled by 1/N.

Pinned by: ``numpy.fft`` (library routine) to 1e-13 at N = 32..4096, fixtures.json ``fft``,
bijectivity of the digit reversal, and a brute-force O(N^2) DFT at small N.
"""

import numpy as np


def radix_schedule(n):
    out = []
    m = int(n)
    while m % 4 == 0 and m > 1:
        out.append(4)
        m //= 4
    if m == 2:
        out.append(2)
    elif m != 1:
        raise ValueError("grid size %d must be 4^a or 2 * 4^a" % n)
    return out


def digit_reverse(i, radices):
    rev = 0
    for r in radices[::-1]:
        rev = rev * r + i % r
        i //= r
    return rev


def dit_line(x, inverse=False):
    x = np.asarray(x, np.complex128)
    n = x.size
    radices = radix_schedule(n)
    sgn = 1.0 if inverse else -1.0
    buf = np.empty(n, np.complex128)
    buf[[digit_reverse(i, radices) for i in range(n)]] = x
    m = 1
    for r in radices:
        span = r * m
        j = np.arange(n // r) % m
        base = (np.arange(n // r) // m) * span + j
        w = np.exp(1j * sgn * 2.0 * np.pi * j / span)
        legs = [buf[base + q * m] * w ** q for q in range(r)]
        if r == 4:
            s02, d02 = legs[0] + legs[2], legs[0] - legs[2]
            s13, d13 = legs[1] + legs[3], 1j * sgn * (legs[1] - legs[3])
            buf[base] = s02 + s13
            buf[base + m] = d02 + d13
            buf[base + 2 * m] = s02 - s13
            buf[base + 3 * m] = d02 - d13
        else:
            buf[base] = legs[0] + legs[1]
            buf[base + m] = legs[0] - legs[1]
        m = span
    return buf / n if inverse else buf


def dit_2d(a, inverse=False):
    """Row pass then column pass, each scaled by 1/N when inverse (src/gpu/fft.js:81-92)."""
    out = np.array(a, np.complex128)
    out = np.apply_along_axis(dit_line, 1, out, inverse)
    return np.apply_along_axis(dit_line, 0, out, inverse)


def brute_dft(x, inverse=False):
    n = len(x)
    k = np.arange(n)
    sgn = 1.0 if inverse else -1.0
    mat = np.exp(sgn * 2j * np.pi * np.outer(k, k) / n)
    y = mat @ np.asarray(x, np.complex128)
    return y / n if inverse else y
