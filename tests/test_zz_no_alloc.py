"""s2t_fill_spectrum_noalloc: fillSpectrum into caller device memory with no allocation.

Checked through the C ABI against the existing entry point it varies, s2t_seed_spectrum (same
spectrum, read back in the fixture layout): bit for bit, on two calls in a row that share the
library's fixed scratch, the second with a smaller grid; and against the FP64 oracle's
fillSpectrum to FP32 rounding. The caller's device buffers come from the HIP runtime the library
itself uses (hipMalloc through ctypes, by the test, not by the entry point).
"""

import ctypes

import numpy as np
import pytest

from oracle import gpusem

pytestmark = pytest.mark.gpu

HIP_D2H = 2


def _hip():
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    return hip


def _ok(rc, what):
    assert rc == 0, "%s returned %d" % (what, rc)


def _seeded(s2t, n, k0, amp, seed):
    with s2t.Solver(n) as s:
        s.seed_spectrum(k0=k0, amplitude=amp, seed=seed)
        return s2t.to_interleaved(s.read_state())


def test_noalloc_fill_matches_seed_spectrum(s2t):
    lib = s2t.load()
    hip = _hip()
    nmax = 512
    buf = ctypes.c_void_p()
    stream = ctypes.c_void_p()
    _ok(hip.hipMalloc(ctypes.byref(buf), 2 * nmax * nmax * 4), "hipMalloc")
    _ok(hip.hipStreamCreate(ctypes.byref(stream)), "hipStreamCreate")
    try:
        # two calls in a row on one stream, the second smaller; then one on the null stream
        for n, k0, amp, seed, strm in ((512, 20.0, 6.0, 7, stream), (256, 9.0, 2.5, 11, stream),
                                       (128, 14.0, 6.0, 1, None)):
            print("noalloc fill n=%d k0=%g amplitude=%g seed=%d" % (n, k0, amp, seed), flush=True)
            rc = lib.s2t_fill_spectrum_noalloc(n, k0, amp, seed, buf, strm)
            assert rc == 0, lib.s2t_last_error().decode()
            _ok(hip.hipStreamSynchronize(strm), "hipStreamSynchronize")
            got = np.empty(2 * n * n, np.float32)
            _ok(hip.hipMemcpy(got.ctypes.data, buf, got.nbytes, HIP_D2H), "hipMemcpy")
            want = _seeded(s2t, n, k0, amp, seed)
            assert np.array_equal(got, want), "n=%d: differs from s2t_seed_spectrum" % n
            ref = gpusem.fill_spectrum(n, k0=k0, amplitude=amp, seed=seed)
            w = s2t.from_interleaved(got, n)
            assert np.max(np.abs(w - ref)) <= 2e-6 * np.max(np.abs(ref))
        # amplitude 0: a fluid at rest
        rc = lib.s2t_fill_spectrum_noalloc(256, 9.0, 0.0, 3, buf, stream)
        assert rc == 0, lib.s2t_last_error().decode()
        _ok(hip.hipStreamSynchronize(stream), "hipStreamSynchronize")
        got = np.empty(2 * 256 * 256, np.float32)
        _ok(hip.hipMemcpy(got.ctypes.data, buf, got.nbytes, HIP_D2H), "hipMemcpy")
        assert not np.any(got)
        # beyond the fixed scratch: rejected before any device work
        assert lib.s2t_fill_spectrum_noalloc(8192, 9.0, 1.0, 3, buf, stream) == s2t.S2T_EINVAL
        assert "at most 4096" in lib.s2t_last_error().decode()
    finally:
        hip.hipStreamDestroy(stream)
        hip.hipFree(buf)
