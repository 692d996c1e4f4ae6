"""ctypes binding of libs2t.so (include/s2t.h) for the tests, the bench and smoke checks.

The product host interface is the Node.js N-API addon in ``js/`` (the reference's own host
language); this module is the same C ABI seen from Python. It loads the in-tree library built by
``build.py`` and raises if that library is missing: there is no CPU fallback.

Load it by path (the directory name is not a Python identifier)::

    import importlib.util, os
    spec = importlib.util.spec_from_file_location(
        "s2t_amd", os.path.join(ROOT, "spectral-2d-turbulence_amd", "__init__.py"))
    s2t = importlib.util.module_from_spec(spec); spec.loader.exec_module(s2t)
"""

import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libs2t.so")

S2T_OK = 0
S2T_EINVAL = -1
S2T_EHIP = -2
S2T_ENOMEM = -3
S2T_ESTATE = -4
S2T_STEP_DYE = 1
PROF_ROWS, PROF_COLS, PROF_DYE, PROF_OTHER = 0, 1, 2, 3
MAX_SPLATS = 8

# every symbol include/s2t.h declares
EXPORTS = [
    "s2t_create", "s2t_destroy", "s2t_last_error", "s2t_radix_schedule", "s2t_configure",
    "s2t_nu_h", "s2t_set_splats", "s2t_seed_spectrum", "s2t_seed_physical",
    "s2t_bandlimited_spectrum", "s2t_seed_bandlimited", "s2t_write_state",
    "s2t_read_state", "s2t_reset_clock", "s2t_set_clock", "s2t_step", "s2t_sync", "s2t_diagnostics",
    "s2t_paint_dye", "s2t_dye_size", "s2t_read_dye", "s2t_write_dye", "s2t_write_velocity",
    "s2t_dye_step", "s2t_fft2", "s2t_velocity_pack", "s2t_assemble", "s2t_rhs", "s2t_profile",
    "s2t_profile_kinds", "s2t_profile_sampled", "s2t_profile_read", "s2t_kernel_bytes", "s2t_copy_probe",
    "s2t_column_widths", "s2t_set_column_width", "s2t_set_row_kernel", "s2t_autotune",
    "s2t_batch_create", "s2t_batch_destroy", "s2t_batch_size", "s2t_batch_configure",
    "s2t_batch_seed_spectrum", "s2t_batch_seed_physical", "s2t_batch_write_state",
    "s2t_batch_read_state", "s2t_batch_diagnostics", "s2t_batch_use_graphs", "s2t_batch_use_grouped", "s2t_batch_step",
    "s2t_batch_sync", "s2t_batch_autotune", "s2t_batch_column_width",
    "s2t_server_create", "s2t_server_destroy", "s2t_server_submit", "s2t_server_poll",
    "s2t_server_read_state", "s2t_server_cancel", "s2t_server_set_callback", "s2t_server_drain",
    "s2t_server_get_stats", "s2t_fill_spectrum_noalloc", "s2t_reset_dye", "s2t_prime",
    "s2t_primed", "s2t_encode_diagnostics", "s2t_copy_diagnostics",
    "s2t_read_copied_diagnostics", "s2t_buffer_create", "s2t_buffer_destroy", "s2t_buffer_write",
    "s2t_buffer_read", "s2t_buffer_copy", "s2t_fft2_device", "s2t_server_inject_failure",
    "s2t_dye_step_passes",
]
SCENARIO = {"turbulence": 0, "taylor-green": 1, "dipole": 2, "merger": 3}
REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_REJECTED, REQ_CANCELLED, REQ_FAILED = 0, 1, 2, 3, 4, 5


class Params(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in (
        "beta", "drag", "viscosity", "hyper_p", "force_amp", "force_k", "force_width", "cfl",
        "dt_min", "dt_max", "dye_fade", "fixed_dt")]


class SplatC(ctypes.Structure):
    _fields_ = [("x", ctypes.c_float), ("y", ctypes.c_float), ("dx", ctypes.c_float),
                ("dy", ctypes.c_float), ("color", ctypes.c_float * 3), ("radius", ctypes.c_float),
                ("amp", ctypes.c_float), ("dye_amp", ctypes.c_float)]


class Diag(ctypes.Structure):
    _fields_ = [("energy", ctypes.c_double), ("enstrophy", ctypes.c_double),
                ("dt", ctypes.c_double), ("max_speed", ctypes.c_double),
                ("time", ctypes.c_double), ("step_index", ctypes.c_uint32),
                ("nonfinite", ctypes.c_uint32)]


class Request(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("steps", ctypes.c_int), ("scenario", ctypes.c_int),
                ("keep_state", ctypes.c_int), ("scenario_param", ctypes.c_double),
                ("k0", ctypes.c_double), ("amplitude", ctypes.c_double), ("seed", ctypes.c_uint32),
                ("params", Params)]


# void (*)(uint64_t id, int status, void* user)
SERVER_CALLBACK = ctypes.CFUNCTYPE(None, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p)


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64),
                ("rejected", ctypes.c_uint64), ("cancelled", ctypes.c_uint64),
                ("p50_latency_s", ctypes.c_double),
                ("p99_latency_s", ctypes.c_double), ("mean_latency_s", ctypes.c_double),
                ("mpts_steps_per_s", ctypes.c_double), ("budget_bytes", ctypes.c_double),
                ("used_bytes", ctypes.c_double), ("peak_bytes", ctypes.c_double)]


class S2TError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (status %d)" % (msg, code))
        self.code = code


_LIB = None


def load(path=None):
    """Load libs2t.so; raises FileNotFoundError when it has not been built."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FileNotFoundError("%s is not built: run spectral-2d-turbulence_amd/build.py" % p)
    lib = ctypes.CDLL(p)
    vp, ip, fp, dp = ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_double
    u32 = ctypes.c_uint32
    sig = {
        "s2t_create": (ip, [ip, ip, ctypes.POINTER(vp)]),
        "s2t_destroy": (ip, [vp]),
        "s2t_last_error": (ctypes.c_char_p, []),
        "s2t_radix_schedule": (ip, [ip, ctypes.POINTER(ip), ctypes.POINTER(ip)]),
        "s2t_configure": (ip, [vp, ctypes.POINTER(Params)]),
        "s2t_nu_h": (ip, [vp, ctypes.POINTER(dp)]),
        "s2t_set_splats": (ip, [vp, ctypes.POINTER(SplatC), ip]),
        "s2t_seed_spectrum": (ip, [vp, dp, dp, u32]),
        "s2t_seed_physical": (ip, [vp, fp]),
        "s2t_bandlimited_spectrum": (ip, [ip, dp, dp, dp, u32, fp]),
        "s2t_seed_bandlimited": (ip, [vp, dp, dp, dp, u32]),
        "s2t_write_state": (ip, [vp, fp]),
        "s2t_read_state": (ip, [vp, fp]),
        "s2t_reset_clock": (ip, [vp]),
        "s2t_set_clock": (ip, [vp, dp, u32]),
        "s2t_step": (ip, [vp, ip, u32]),
        "s2t_sync": (ip, [vp]),
        "s2t_diagnostics": (ip, [vp, ctypes.POINTER(Diag), fp, fp]),
        "s2t_paint_dye": (ip, [vp, ip]),
        "s2t_dye_size": (ip, [vp, ctypes.POINTER(ip)]),
        "s2t_read_dye": (ip, [vp, fp]),
        "s2t_write_dye": (ip, [vp, fp]),
        "s2t_write_velocity": (ip, [vp, fp]),
        "s2t_dye_step": (ip, [vp, dp]),
        "s2t_fft2": (ip, [vp, fp, fp, ip]),
        "s2t_velocity_pack": (ip, [vp, fp, fp, fp]),
        "s2t_assemble": (ip, [vp, fp, fp, u32, dp, fp]),
        "s2t_rhs": (ip, [vp, u32, dp, fp]),
        "s2t_profile": (ip, [vp, ip]),
        "s2t_profile_kinds": (ip, [vp, ctypes.c_uint]),
        "s2t_profile_sampled": (ip, [vp, ctypes.c_uint, ctypes.c_int]),
        "s2t_profile_read": (ip, [vp, ip, ctypes.POINTER(ip), ctypes.POINTER(dp)]),
        "s2t_kernel_bytes": (ip, [vp, ip, ip, ctypes.POINTER(dp)]),
        "s2t_copy_probe": (ip, [dp, ip, ctypes.POINTER(dp)]),
        "s2t_column_widths": (ip, [ip, ctypes.POINTER(ip), ip, ctypes.POINTER(ip)]),
        "s2t_set_column_width": (ip, [vp, ip]),
        "s2t_set_row_kernel": (ip, [vp, ip]),
        "s2t_autotune": (ip, [vp, ip, ctypes.POINTER(ip), ctypes.POINTER(dp), ip]),
        "s2t_batch_create": (ip, [ctypes.POINTER(ctypes.c_int), ip, ctypes.POINTER(vp)]),
        "s2t_batch_destroy": (ip, [vp]),
        "s2t_batch_size": (ip, [vp, ip, ctypes.POINTER(ip)]),
        "s2t_batch_configure": (ip, [vp, ip, ctypes.POINTER(Params)]),
        "s2t_batch_seed_spectrum": (ip, [vp, ip, dp, dp, u32]),
        "s2t_batch_seed_physical": (ip, [vp, ip, fp]),
        "s2t_batch_write_state": (ip, [vp, ip, fp]),
        "s2t_batch_read_state": (ip, [vp, ip, fp]),
        "s2t_batch_diagnostics": (ip, [vp, ip, ctypes.POINTER(Diag), fp, fp]),
        "s2t_batch_use_graphs": (ip, [vp, ip]),
        "s2t_batch_use_grouped": (ip, [vp, ip]),
        "s2t_batch_step": (ip, [vp, ip]),
        "s2t_batch_sync": (ip, [vp]),
        "s2t_batch_autotune": (ip, [vp, ip]),
        "s2t_batch_column_width": (ip, [vp, ip, ctypes.POINTER(ip)]),
        "s2t_server_create": (ip, [ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ip,
                                   dp, ctypes.POINTER(vp)]),
        "s2t_server_destroy": (ip, [vp]),
        "s2t_server_submit": (ip, [vp, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64)]),
        "s2t_server_poll": (ip, [vp, ctypes.c_uint64, ctypes.POINTER(ip), ctypes.POINTER(Diag),
                                 ctypes.POINTER(dp)]),
        "s2t_server_read_state": (ip, [vp, ctypes.c_uint64, fp]),
        "s2t_server_cancel": (ip, [vp, ctypes.c_uint64]),
        "s2t_server_set_callback": (ip, [vp, SERVER_CALLBACK, vp]),
        "s2t_server_drain": (ip, [vp, dp]),
        "s2t_server_get_stats": (ip, [vp, ctypes.POINTER(ServerStats)]),
        "s2t_fill_spectrum_noalloc": (ip, [ip, dp, dp, ctypes.c_uint32, vp, vp]),
        "s2t_reset_dye": (ip, [vp]),
        "s2t_prime": (ip, [vp]),
        "s2t_primed": (ip, [vp, ctypes.POINTER(ip)]),
        "s2t_encode_diagnostics": (ip, [vp]),
        "s2t_copy_diagnostics": (ip, [vp]),
        "s2t_read_copied_diagnostics": (ip, [vp, ctypes.POINTER(Diag), vp, vp]),
        "s2t_buffer_create": (ip, [ctypes.c_size_t, ctypes.POINTER(vp)]),
        "s2t_buffer_destroy": (ip, [vp]),
        "s2t_buffer_write": (ip, [vp, ctypes.c_size_t, vp, ctypes.c_size_t]),
        "s2t_buffer_read": (ip, [vp, ctypes.c_size_t, vp, ctypes.c_size_t]),
        "s2t_buffer_copy": (ip, [vp, ctypes.c_size_t, vp, ctypes.c_size_t, ctypes.c_size_t]),
        "s2t_fft2_device": (ip, [vp, vp, ip]),
        "s2t_server_inject_failure": (ip, [vp, ip]),
        "s2t_dye_step_passes": (ip, [vp, dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _LIB = lib
    return lib


def _check(lib, rc):
    if rc != S2T_OK:
        raise S2TError(rc, lib.s2t_last_error().decode())
    return rc


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def to_interleaved(z):
    """complex N x N -> float32 interleaved (re, im) in the fixture layout."""
    z = np.asarray(z)
    out = np.empty(z.shape + (2,), np.float32)
    out[..., 0] = z.real
    out[..., 1] = z.imag
    return np.ascontiguousarray(out.reshape(-1))


def from_interleaved(a, n):
    a = np.asarray(a, np.float64).reshape(n, n, 2)
    return a[..., 0] + 1j * a[..., 1]


DEFAULTS = dict(beta=0.0, drag=0.01, viscosity=0.67, hyper_p=4.0, force_amp=0.0, force_k=22.0,
                force_width=2.5, cfl=0.5, dt_min=1e-5, dt_max=0.02, dye_fade=0.999, fixed_dt=0.0)


class Solver:
    """One solver instance on the current HIP device (mirrors the reference Solver object)."""

    def __init__(self, n, dye_scale=1):
        self.lib = load()
        h = ctypes.c_void_p()
        _check(self.lib, self.lib.s2t_create(int(n), int(dye_scale), ctypes.byref(h)))
        self.h = h
        self.n = int(n)
        self.dye_scale = int(dye_scale)
        self.params = dict(DEFAULTS)

    def close(self):
        if self.h:
            self.lib.s2t_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def configure(self, **kw):
        unknown = set(kw) - set(DEFAULTS)
        if unknown:
            raise KeyError("unknown parameters: %s" % sorted(unknown))
        self.params.update(kw)
        p = Params(**{k: float(v) for k, v in self.params.items()})
        _check(self.lib, self.lib.s2t_configure(self.h, ctypes.byref(p)))

    def nu_h(self):
        v = ctypes.c_double()
        _check(self.lib, self.lib.s2t_nu_h(self.h, ctypes.byref(v)))
        return v.value

    def set_splats(self, splats):
        arr = (SplatC * max(1, len(splats)))()
        for i, s in enumerate(splats[:MAX_SPLATS]):
            arr[i].x, arr[i].y, arr[i].dx, arr[i].dy = s["x"], s["y"], s["dx"], s["dy"]
            c = s.get("color", (0.0, 0.0, 0.0))
            arr[i].color[0], arr[i].color[1], arr[i].color[2] = c
            arr[i].radius, arr[i].amp = s["radius"], s["amp"]
            arr[i].dye_amp = s.get("dyeAmp", 0.0)
        _check(self.lib, self.lib.s2t_set_splats(self.h, arr, len(splats)))

    def seed_spectrum(self, k0=14.0, amplitude=6.0, seed=1):
        _check(self.lib, self.lib.s2t_seed_spectrum(self.h, float(k0), float(amplitude),
                                                    int(seed) & 0xFFFFFFFF))

    def seed_physical(self, field):
        f = np.ascontiguousarray(np.asarray(field, np.float32).reshape(-1))
        assert f.size == self.n * self.n
        _check(self.lib, self.lib.s2t_seed_physical(self.h, _ptr(f)))

    def seed_bandlimited(self, kmin=8.0, kmax=16.0, amplitude=6.0, seed=2):
        """s2t_seed_bandlimited: random in-band modes with kmin <= |k| <= kmax, rms amplitude."""
        _check(self.lib, self.lib.s2t_seed_bandlimited(self.h, float(kmin), float(kmax),
                                                       float(amplitude), int(seed) & 0xFFFFFFFF))

    def write_state(self, w):
        a = to_interleaved(np.asarray(w).reshape(self.n, self.n))
        _check(self.lib, self.lib.s2t_write_state(self.h, _ptr(a)))

    def read_state(self):
        a = np.empty(2 * self.n * self.n, np.float32)
        _check(self.lib, self.lib.s2t_read_state(self.h, _ptr(a)))
        return from_interleaved(a, self.n)

    def reset_clock(self):
        _check(self.lib, self.lib.s2t_reset_clock(self.h))

    def set_clock(self, time, step_index):
        """s2t_set_clock: simulated time and step index (resume from a checkpoint)."""
        _check(self.lib, self.lib.s2t_set_clock(self.h, float(time), int(step_index) & 0xFFFFFFFF))

    def step(self, nsteps=1, dye=False):
        _check(self.lib, self.lib.s2t_step(self.h, int(nsteps), S2T_STEP_DYE if dye else 0))

    def sync(self):
        _check(self.lib, self.lib.s2t_sync(self.h))

    def diagnostics(self):
        d = Diag()
        spec = np.empty(self.n // 2 + 1, np.float64)
        zonal = np.empty(self.n, np.float64)
        _check(self.lib, self.lib.s2t_diagnostics(self.h, ctypes.byref(d), _ptr(spec),
                                                  _ptr(zonal)))
        return dict(energy=d.energy, enstrophy=d.enstrophy, dt=d.dt, maxSpeed=d.max_speed,
                    time=d.time, stepIndex=d.step_index, nonfinite=d.nonfinite,
                    spectrum=spec, zonal=zonal)

    def paint_dye(self, pattern="bands"):
        _check(self.lib, self.lib.s2t_paint_dye(self.h, {"none": 0, "bands": 1}[pattern]))

    def dye_size(self):
        m = ctypes.c_int()
        _check(self.lib, self.lib.s2t_dye_size(self.h, ctypes.byref(m)))
        return m.value

    def read_dye(self):
        m = self.dye_size()
        a = np.empty(m * m * 4, np.uint16)
        _check(self.lib, self.lib.s2t_read_dye(self.h, _ptr(a)))
        return a.view(np.float16).reshape(m, m, 4)

    def write_dye(self, rgba):
        a = np.ascontiguousarray(np.asarray(rgba, np.float16)).view(np.uint16).reshape(-1)
        _check(self.lib, self.lib.s2t_write_dye(self.h, _ptr(a)))

    def write_velocity(self, uv):
        a = np.ascontiguousarray(np.asarray(uv, np.float32).reshape(-1))
        _check(self.lib, self.lib.s2t_write_velocity(self.h, _ptr(a)))

    def dye_step(self, dt=0.0):
        """One dye update (dt <= 0: the last step's dt on the device); enqueues only."""
        _check(self.lib, self.lib.s2t_dye_step(self.h, float(dt)))

    def dye_step_passes(self, dt=0.0):
        """The same update as three launches (the unfused passes; test hook); enqueues only."""
        _check(self.lib, self.lib.s2t_dye_step_passes(self.h, float(dt)))

    def reset_dye(self):
        _check(self.lib, self.lib.s2t_reset_dye(self.h))

    def prime(self):
        _check(self.lib, self.lib.s2t_prime(self.h))

    @property
    def primed(self):
        v = ctypes.c_int()
        _check(self.lib, self.lib.s2t_primed(self.h, ctypes.byref(v)))
        return bool(v.value)

    def encode_diagnostics(self):
        _check(self.lib, self.lib.s2t_encode_diagnostics(self.h))

    def copy_diagnostics(self):
        _check(self.lib, self.lib.s2t_copy_diagnostics(self.h))

    def read_copied_diagnostics(self):
        d = Diag()
        spec = np.empty(self.n // 2 + 1, np.float64)
        zonal = np.empty(self.n, np.float64)
        _check(self.lib, self.lib.s2t_read_copied_diagnostics(self.h, ctypes.byref(d),
                                                              _ptr(spec), _ptr(zonal)))
        return dict(energy=d.energy, enstrophy=d.enstrophy, dt=d.dt, maxSpeed=d.max_speed,
                    time=d.time, stepIndex=d.step_index, nonfinite=d.nonfinite,
                    spectrum=spec, zonal=zonal)

    def fft2_device(self, dev_ptr, inverse=False):
        """In-place 2-D FFT of an n x n complex device buffer (DeviceBuffer.ptr); enqueues."""
        _check(self.lib, self.lib.s2t_fft2_device(self.h, dev_ptr, 1 if inverse else 0))

    # test hooks
    def fft2(self, z, inverse=False):
        a = to_interleaved(np.asarray(z).reshape(self.n, self.n))
        out = np.empty_like(a)
        _check(self.lib, self.lib.s2t_fft2(self.h, _ptr(a), _ptr(out), 1 if inverse else 0))
        return from_interleaved(out, self.n)

    def velocity_pack(self, w):
        a = to_interleaved(np.asarray(w).reshape(self.n, self.n))
        pa = np.empty_like(a)
        pb = np.empty_like(a)
        _check(self.lib, self.lib.s2t_velocity_pack(self.h, _ptr(a), _ptr(pa), _ptr(pb)))
        return from_interleaved(pa, self.n), from_interleaved(pb, self.n)

    def assemble(self, z, fhat=None, step_index=0, dt=0.01):
        a = to_interleaved(np.asarray(z).reshape(self.n, self.n))
        f = None if fhat is None else to_interleaved(np.asarray(fhat).reshape(self.n, self.n))
        out = np.empty_like(a)
        _check(self.lib, self.lib.s2t_assemble(self.h, _ptr(a), None if f is None else _ptr(f),
                                               int(step_index), float(dt), _ptr(out)))
        return from_interleaved(out, self.n)

    def rhs(self, step_index=0, dt=0.01):
        out = np.empty(2 * self.n * self.n, np.float32)
        _check(self.lib, self.lib.s2t_rhs(self.h, int(step_index), float(dt), _ptr(out)))
        return from_interleaved(out, self.n)

    # profiling
    def profile(self, enable, kinds=None, every=1):
        """Event timing of s2t_step's launches: all kernel kinds, or only `kinds` (PROF_*), in
        one step of `every`."""
        if enable and (kinds is not None or every != 1):
            mask = (1 << 4) - 1 if kinds is None else 0
            for k in kinds or ():
                mask |= 1 << int(k)
            _check(self.lib, self.lib.s2t_profile_sampled(self.h, mask, int(every)))
        else:
            _check(self.lib, self.lib.s2t_profile(self.h, 1 if enable else 0))

    def profile_read(self, kind):
        n = ctypes.c_int()
        ms = ctypes.c_double()
        _check(self.lib, self.lib.s2t_profile_read(self.h, kind, ctypes.byref(n), ctypes.byref(ms)))
        return n.value, ms.value

    def kernel_bytes(self, kind, stage=0):
        b = ctypes.c_double()
        _check(self.lib, self.lib.s2t_kernel_bytes(self.h, kind, stage, ctypes.byref(b)))
        return b.value

    def set_column_width(self, width):
        """s2t_set_column_width: a compiled column-group width for the steps (0: the default)."""
        _check(self.lib, self.lib.s2t_set_column_width(self.h, int(width)))

    def set_row_kernel(self, field_parallel):
        """s2t_set_row_kernel: the field-parallel row kernel (True, the default where compiled) or
        one thread group per row pair (False)."""
        _check(self.lib, self.lib.s2t_set_row_kernel(self.h, 1 if field_parallel else 0))

    def autotune(self, steps=5):
        """s2t_autotune: (chosen width, {width: ms per step}) from a scratch solver."""
        widths = column_widths(self.n)
        w = ctypes.c_int()
        ms = (ctypes.c_double * len(widths))()
        _check(self.lib, self.lib.s2t_autotune(self.h, int(steps), ctypes.byref(w), ms,
                                               len(widths)))
        return w.value, dict(zip(widths, list(ms)))


def copy_probe(nbytes, reps=10):
    """Average ms of one plain device copy of nbytes (HBM counter calibration)."""
    lib = load()
    ms = ctypes.c_double()
    _check(lib, lib.s2t_copy_probe(float(nbytes), int(reps), ctypes.byref(ms)))
    return ms.value


def _params(kw):
    p = dict(DEFAULTS)
    unknown = set(kw) - set(DEFAULTS)
    if unknown:
        raise KeyError("unknown parameters: %s" % sorted(unknown))
    p.update(kw)
    return Params(**{k: float(v) for k, v in p.items()})


def _diag_dict(d, spec=None, zonal=None):
    out = dict(energy=d.energy, enstrophy=d.enstrophy, dt=d.dt, maxSpeed=d.max_speed,
               time=d.time, stepIndex=d.step_index, nonfinite=d.nonfinite)
    if spec is not None:
        out["spectrum"] = spec
        out["zonal"] = zonal
    return out


def column_widths(n):
    """s2t_column_widths (host only): the compiled column-group widths for size n, default first."""
    lib = load()
    buf = (ctypes.c_int * 8)()
    cnt = ctypes.c_int()
    _check(lib, lib.s2t_column_widths(int(n), buf, 8, ctypes.byref(cnt)))
    return list(buf[:cnt.value])


def bandlimited_spectrum(n, kmin=8.0, kmax=16.0, amplitude=6.0, seed=2):
    """s2t_bandlimited_spectrum (host only): the band-limited random spectrum as complex64 n x n."""
    lib = load()
    a = np.empty(2 * n * n, np.float32)
    _check(lib, lib.s2t_bandlimited_spectrum(int(n), float(kmin), float(kmax), float(amplitude),
                                             int(seed) & 0xFFFFFFFF, _ptr(a)))
    return from_interleaved(a, n)


class DeviceBuffer:
    """Device memory for the Fft2D.register route (s2t_buffer_*): zero-filled at creation;
    write / read / copy_from are host-blocking and not ordered against a solver's stream."""

    def __init__(self, nbytes):
        self.lib = load()
        p = ctypes.c_void_p()
        _check(self.lib, self.lib.s2t_buffer_create(int(nbytes), ctypes.byref(p)))
        self.ptr = p
        self.nbytes = int(nbytes)

    def write(self, data, offset=0):
        a = np.ascontiguousarray(data)
        _check(self.lib, self.lib.s2t_buffer_write(self.ptr, int(offset), _ptr(a), a.nbytes))

    def read(self, dtype=np.float32, offset=0, nbytes=None):
        nbytes = self.nbytes - offset if nbytes is None else int(nbytes)
        out = np.empty(nbytes // np.dtype(dtype).itemsize, dtype)
        _check(self.lib, self.lib.s2t_buffer_read(self.ptr, int(offset), _ptr(out), out.nbytes))
        return out

    def copy_from(self, src, src_offset=0, dst_offset=0, nbytes=None):
        nbytes = src.nbytes - src_offset if nbytes is None else int(nbytes)
        _check(self.lib, self.lib.s2t_buffer_copy(self.ptr, int(dst_offset), src.ptr,
                                                  int(src_offset), nbytes))

    def close(self):
        if self.ptr:
            self.lib.s2t_buffer_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Ensemble of independent members of mixed grid sizes (s2t_batch_*)."""

    def __init__(self, sizes):
        self.lib = load()
        self.sizes = [int(n) for n in sizes]
        arr = (ctypes.c_int * len(self.sizes))(*self.sizes)
        h = ctypes.c_void_p()
        _check(self.lib, self.lib.s2t_batch_create(arr, len(self.sizes), ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.lib.s2t_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def configure(self, i, **kw):
        p = _params(kw)
        _check(self.lib, self.lib.s2t_batch_configure(self.h, i, ctypes.byref(p)))

    def seed_spectrum(self, i, k0=14.0, amplitude=6.0, seed=1):
        _check(self.lib, self.lib.s2t_batch_seed_spectrum(self.h, i, float(k0), float(amplitude),
                                                          int(seed) & 0xFFFFFFFF))

    def seed_physical(self, i, field):
        f = np.ascontiguousarray(np.asarray(field, np.float32).reshape(-1))
        _check(self.lib, self.lib.s2t_batch_seed_physical(self.h, i, _ptr(f)))

    def write_state(self, i, w):
        n = self.sizes[i]
        a = to_interleaved(np.asarray(w).reshape(n, n))
        _check(self.lib, self.lib.s2t_batch_write_state(self.h, i, _ptr(a)))

    def read_state(self, i):
        n = self.sizes[i]
        a = np.empty(2 * n * n, np.float32)
        _check(self.lib, self.lib.s2t_batch_read_state(self.h, i, _ptr(a)))
        return from_interleaved(a, n)

    def diagnostics(self, i):
        n = self.sizes[i]
        d = Diag()
        spec = np.empty(n // 2 + 1, np.float64)
        zonal = np.empty(n, np.float64)
        _check(self.lib, self.lib.s2t_batch_diagnostics(self.h, i, ctypes.byref(d), _ptr(spec),
                                                        _ptr(zonal)))
        return _diag_dict(d, spec, zonal)

    def use_graphs(self, enable):
        _check(self.lib, self.lib.s2t_batch_use_graphs(self.h, 1 if enable else 0))

    def use_grouped(self, enable):
        """s2t_batch_use_grouped: step the 256/512/1024 classes by grouped launches."""
        _check(self.lib, self.lib.s2t_batch_use_grouped(self.h, 1 if enable else 0))

    def autotune(self, steps=5):
        """s2t_batch_autotune: column-group width per size class; returns {size: width}."""
        _check(self.lib, self.lib.s2t_batch_autotune(self.h, int(steps)))
        return self.column_widths()

    def column_widths(self):
        """s2t_batch_column_width of each size class; returns {size: width}."""
        out = {}
        for i, n in enumerate(self.sizes):
            w = ctypes.c_int()
            _check(self.lib, self.lib.s2t_batch_column_width(self.h, i, ctypes.byref(w)))
            out[n] = w.value
        return out

    def step(self, nsteps=1):
        _check(self.lib, self.lib.s2t_batch_step(self.h, int(nsteps)))

    def sync(self):
        _check(self.lib, self.lib.s2t_batch_sync(self.h))


class Server:
    """Continuous-batching server (s2t_server_*): chunks of slots per grid size, allocated as
    requests need them within a device-memory budget; `classes` maps sizes to slots per chunk."""

    def __init__(self, classes, budget_bytes=0.0):
        self.lib = load()
        sizes = sorted(classes)
        self.sizes = sizes
        a = (ctypes.c_int * len(sizes))(*sizes)
        b = (ctypes.c_int * len(sizes))(*[int(classes[n]) for n in sizes])
        h = ctypes.c_void_p()
        _check(self.lib, self.lib.s2t_server_create(a, b, len(sizes), float(budget_bytes),
                                                    ctypes.byref(h)))
        self.h = h
        self.n_of = {}

    def close(self):
        if self.h:
            self.lib.s2t_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, n, steps, scenario="turbulence", scenario_param=0.0, k0=14.0, amplitude=6.0,
               seed=1, keep_state=False, **params):
        r = Request(n=int(n), steps=int(steps), scenario=SCENARIO[scenario],
                    keep_state=1 if keep_state else 0, scenario_param=float(scenario_param),
                    k0=float(k0), amplitude=float(amplitude), seed=int(seed) & 0xFFFFFFFF,
                    params=_params(params))
        rid = ctypes.c_uint64()
        _check(self.lib, self.lib.s2t_server_submit(self.h, ctypes.byref(r), ctypes.byref(rid)))
        self.n_of[rid.value] = int(n)
        return rid.value

    def poll(self, rid):
        st = ctypes.c_int()
        d = Diag()
        lat = ctypes.c_double()
        _check(self.lib, self.lib.s2t_server_poll(self.h, ctypes.c_uint64(rid), ctypes.byref(st),
                                                  ctypes.byref(d), ctypes.byref(lat)))
        return st.value, _diag_dict(d), lat.value

    def read_state(self, rid):
        n = self.n_of[rid]
        a = np.empty(2 * n * n, np.float32)
        _check(self.lib, self.lib.s2t_server_read_state(self.h, ctypes.c_uint64(rid), _ptr(a)))
        return from_interleaved(a, n)

    def cancel(self, rid):
        """s2t_server_cancel: drop a queued request, or stop a running one between steps."""
        _check(self.lib, self.lib.s2t_server_cancel(self.h, ctypes.c_uint64(rid)))

    def set_callback(self, fn):
        """fn(id, status) once per request at its final status, on the server's worker thread (or
        on the submitting / cancelling thread when final at that call); None removes it."""
        self._cb = SERVER_CALLBACK(lambda rid, status, user: fn(int(rid), int(status))) \
            if fn is not None else SERVER_CALLBACK()
        _check(self.lib, self.lib.s2t_server_set_callback(self.h, self._cb, None))

    def drain(self, timeout_s=0.0):
        _check(self.lib, self.lib.s2t_server_drain(self.h, float(timeout_s)))

    def inject_failure(self, after_admissions):
        """Make the worker fail at the admission after `after_admissions` admissions."""
        _check(self.lib, self.lib.s2t_server_inject_failure(self.h, int(after_admissions)))

    def stats(self):
        st = ServerStats()
        _check(self.lib, self.lib.s2t_server_get_stats(self.h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}
