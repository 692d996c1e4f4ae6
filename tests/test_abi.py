"""The C ABI library loads and exports every entry point of include/s2t.h (no GPU calls; the
host-only spectrum generator is called)."""

import os
import re

import numpy as np
import pytest

from conftest import ROOT


def header_symbols():
    text = open(os.path.join(ROOT, "include", "s2t.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(s2t_[a-z0-9_]+)\s*\(", text)))


def test_header_matches_binding(s2t):
    assert header_symbols() == sorted(s2t.EXPORTS)


def test_library_exports(s2t):
    lib = s2t.load()
    for name in header_symbols():
        assert hasattr(lib, name), name


def test_radix_schedule_and_bad_sizes(s2t):
    import ctypes
    lib = s2t.load()
    st, r2 = ctypes.c_int(), ctypes.c_int()
    assert lib.s2t_radix_schedule(128, ctypes.byref(st), ctypes.byref(r2)) == 0
    assert (st.value, r2.value) == (3, 1)
    assert lib.s2t_radix_schedule(4096, ctypes.byref(st), ctypes.byref(r2)) == 0
    assert (st.value, r2.value) == (6, 0)
    assert lib.s2t_radix_schedule(96, None, None) == s2t.S2T_EINVAL
    # message of src/gpu/fft.js:28
    assert lib.s2t_last_error().decode() == "grid size 96 must be 4^a or 2 * 4^a"
    h = ctypes.c_void_p()
    assert lib.s2t_create(100, 1, ctypes.byref(h)) == s2t.S2T_EINVAL
    assert "must be 4^a or 2 * 4^a" in lib.s2t_last_error().decode()
    assert lib.s2t_create(8192, 1, ctypes.byref(h)) == s2t.S2T_EINVAL
    assert lib.s2t_create(256, 3, ctypes.byref(h)) == s2t.S2T_EINVAL
    assert not h.value


def test_null_handles_rejected(s2t):
    lib = s2t.load()
    assert lib.s2t_step(None, 1, 0) == s2t.S2T_EINVAL
    assert lib.s2t_destroy(None) == 0


@pytest.mark.parametrize("n, kmin, kmax", [(128, 4.0, 12.0), (1024, 8.0, 16.0)])
def test_bandlimited_spectrum_host_matches_oracle(s2t, n, kmin, kmax):
    """s2t_bandlimited_spectrum is host code (no GPU): the library's spectrum equals the oracle's
    to one FP32 ulp (the FP64 normalisation sums run in different orders)."""
    from oracle import gpusem as gs
    got = s2t.bandlimited_spectrum(n, kmin, kmax, 6.0, seed=2)
    want = gs.bandlimited_spectrum(n, kmin, kmax, 6.0, seed=2)
    assert np.array_equal(got != 0, want != 0)
    scale = np.max(np.abs(want))
    assert np.max(np.abs(got - want)) <= np.finfo(np.float32).eps * scale
    with pytest.raises(s2t.S2TError, match="4\\^a"):
        s2t.bandlimited_spectrum(100, 8.0, 16.0)
    with pytest.raises(s2t.S2TError, match="kmin"):
        s2t.bandlimited_spectrum(128, 16.0, 8.0)


def test_column_widths_host(s2t):
    """s2t_column_widths (host only): every supported size lists its compiled column-group widths,
    the default first, each a workgroup of at most 1024 threads (the line FFT's first radix values
    per thread: 16 at 1024 to 4096, 4 at 16, else 8); a negative
    width -g is the field-parallel column kernel (g thread groups on one column), the default with
    three groups at 512 and 1024 and with two at 2048."""
    fp = {512: -3, 1024: -3, 2048: -2}
    for n in (16, 32, 64, 128, 256, 512, 1024, 2048, 4096):
        w = s2t.column_widths(n)
        assert len(w) == len(set(w)) >= 1
        per = n // (16 if n >= 1024 else 4 if n == 16 else 8)
        assert all((c >= 1 and c * per <= 1024) or (c in (-2, -3) and -c * per <= 1024) for c in w)
        assert [c for c in w if c < 0] == ([fp[n]] if n in fp else [])
        assert (w[0] < 0) == (n in fp)
    assert s2t.column_widths(4096) == [2] and s2t.column_widths(1024)[:2] == [-3, 1]
    assert s2t.column_widths(2048) == [-2, 1, 2, 4]
    with pytest.raises(s2t.S2TError, match="4\\^a"):
        s2t.column_widths(96)


def test_argument_checks_before_any_device_work(s2t):
    """Argument errors of the batch, server and solver entry points are found on the host (no GPU
    call is made), with the messages a caller sees."""
    import ctypes
    lib = s2t.load()
    h = ctypes.c_void_p()
    empty = (ctypes.c_int * 1)(0)
    assert lib.s2t_batch_create(empty, 0, ctypes.byref(h)) == s2t.S2T_EINVAL
    assert "at least one member" in lib.s2t_last_error().decode()
    ragged = (ctypes.c_int * 3)(256, 100, 512)
    assert lib.s2t_batch_create(ragged, 3, ctypes.byref(h)) == s2t.S2T_EINVAL
    assert "4^a" in lib.s2t_last_error().decode() and not h.value
    sizes, slots = (ctypes.c_int * 2)(256, 512), (ctypes.c_int * 2)(4, 0)
    assert lib.s2t_server_create(sizes, slots, 2, 0.0, ctypes.byref(h)) == s2t.S2T_EINVAL
    assert "slot counts" in lib.s2t_last_error().decode()
    slots = (ctypes.c_int * 2)(4, 4)
    assert lib.s2t_server_create(sizes, slots, -1, 0.0, ctypes.byref(h)) == s2t.S2T_EINVAL
    assert lib.s2t_server_create(sizes, slots, 2, -1.0, ctypes.byref(h)) == s2t.S2T_EINVAL
    assert lib.s2t_server_create(sizes, slots, 2, 1e6, ctypes.byref(h)) == s2t.S2T_EINVAL
    assert "over the budget" in lib.s2t_last_error().decode()
    # null handles everywhere
    assert lib.s2t_batch_step(None, 1) == s2t.S2T_EINVAL
    assert lib.s2t_server_submit(None, None, None) == s2t.S2T_EINVAL
    assert lib.s2t_autotune(None, 3, None, None, 0) == s2t.S2T_EINVAL
    assert lib.s2t_set_clock(None, 0.0, 0) == s2t.S2T_EINVAL
    assert lib.s2t_profile_sampled(None, 1, 4) == s2t.S2T_EINVAL
    # the no-allocation fill: sizes beyond its fixed scratch, bad sizes, bad k0, NULL output
    fake = ctypes.c_void_p(4096)  # never dereferenced: every call below fails its checks first
    assert lib.s2t_fill_spectrum_noalloc(8192, 8.0, 1.0, 1, fake, None) == s2t.S2T_EINVAL
    assert "exceeds its fixed scratch (at most 4096)" in lib.s2t_last_error().decode()
    assert lib.s2t_fill_spectrum_noalloc(96, 8.0, 1.0, 1, fake, None) == s2t.S2T_EINVAL
    assert "4^a" in lib.s2t_last_error().decode()
    assert lib.s2t_fill_spectrum_noalloc(256, 0.0, 1.0, 1, fake, None) == s2t.S2T_EINVAL
    assert lib.s2t_fill_spectrum_noalloc(256, 8.0, 1.0, 1, None, None) == s2t.S2T_EINVAL
    assert "NULL" in lib.s2t_last_error().decode()
