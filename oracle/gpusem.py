"""Semantics that exist only in the reference's GPU path (test infrastructure only).

None of these has an FP64 counterpart in reference/spectral2d.py; each is restated here from the
WGSL/JS source it cites, in FP64 unless stated, and pinned as its docstring says.
"""

import numpy as np

from .spectral import TWO_PI, Grid

U32 = np.uint32
F32_TWO_PI = np.float32(6.28318530717958647)


# ---------------------------------------------------------------------------------------------
# integer hash
# ---------------------------------------------------------------------------------------------
def hash1(x):
    """32-bit integer hash of src/shaders/common.wgsl:76-80 (same as hash32, solver.js:43-47).

    multiply-add, data-dependent xorshift, multiply, xorshift, all modulo 2^32.
    Pinned by known answers computed with C uint32 arithmetic (tests/test_oracle_semantics.py)
    and by a uniformity check of the derived phases.
    """
    v = np.asarray(x, dtype=np.uint64) & 0xFFFFFFFF
    v = (v * 747796405 + 2891336453) & 0xFFFFFFFF
    v = (((v >> ((v >> 28) + 4)) ^ v) * 277803737) & 0xFFFFFFFF
    return ((v >> 22) ^ v).astype(np.uint32)


def hashf(a, b):
    """f32(hash1(a ^ hash1(b))) * 2^-32 evaluated in FP32 (common.wgsl:82-84)."""
    h = hash1(np.asarray(a, np.uint32) ^ hash1(b))
    return h.astype(np.float32) * np.float32(2.3283064365386963e-10)


def canonical_half(kx, ky):
    """The half of the modes that carries its own random phase: ky > 0, or ky = 0 and kx > 0
    (solver.wgsl:80-86 for the forcing, src/gpu/solver.js:401-405 for fillSpectrum); every other
    mode takes the conjugate of its mirror's value.

    Pinned by the known answers of FORCE_KAT / FILL_KAT (tests/test_oracle_pins.py), which include
    modes with ky > 0 and kx < 0 (canonical here, not under a kx-first rule) and modes with
    ky = 0.
    """
    return (ky > 0) | ((ky == 0) & (kx > 0))


def force_phase(idx, step_index):
    """Forcing phase of canonical index idx = row*N + col: TWO_PI * hashf(idx, stepIndex), the
    product in FP32 as in the WGSL (solver.wgsl:87).

    Pinned by known answers computed in C (tests/golden/hash_kat.c), by the zero means of
    e^{i phase} and e^{2 i phase} over the annulus and by zero step-to-step correlation
    (tests/test_oracle_pins.py).
    """
    return (F32_TWO_PI * hashf(np.asarray(idx, np.uint32), U32(step_index))).astype(np.float64)


def fill_phase(idx, seed):
    """fillSpectrum phase of canonical index idx = row*n + col:
    2 pi hash32(idx ^ hash32(seed)) / 2^32 in FP64 (src/gpu/solver.js:387, 404).

    Pinned by known answers computed in C (tests/golden/hash_kat.c) and by the zero means of
    e^{i phase} and e^{2 i phase} (tests/test_oracle_pins.py).
    """
    salt = hash1(np.uint32(int(seed) & 0xFFFFFFFF))
    h = hash1(np.asarray(idx, np.uint32) ^ salt).astype(np.float64)
    return TWO_PI * h / 4294967296.0


# ---------------------------------------------------------------------------------------------
# parameters
# ---------------------------------------------------------------------------------------------
def gpu_dt(speed, cfl, dx, dt_min, dt_max):
    """CFL step of src/shaders/reduce.wgsl:82-88: dt_max when speed <= 1e-8, clamped to dt_min."""
    dt = dt_max
    if speed > 1e-8:
        dt = min(dt_max, cfl * dx / speed)
    return max(dt, dt_min)


def force_norm(n):
    """FORCE_NORM = 0.01 N^2 (src/gpu/shaders.js:26)."""
    return 0.01 * n * n


# ---------------------------------------------------------------------------------------------
# stochastic forcing
# ---------------------------------------------------------------------------------------------
def annulus_envelope(g, k_f, width):
    """Gaussian annulus exp(-((|k| - k_f)/width)^2), in band, k != 0 (solver.wgsl:161-163).

    The WGSL writes the square as pow(x, 2.0), which WGSL leaves undefined for x < 0; the
    intended Gaussian is restated here for both signs.
    """
    kr = np.sqrt(g.k2)
    env = np.exp(-(((kr - k_f) / width) ** 2))
    return np.where(g.band & (g.k2 > 0), env, 0.0)


def stochastic_forcing(n, amp, k_f, width, step_index, dt):
    """White-in-time forcing spectrum added to the rhs (solver.wgsl assemble:161-176).

    For in-band modes with env > 1e-4: the canonical half is ky > 0, or ky = 0 and kx > 0; other
    modes take the hash of their mirror index with the sign of the sine flipped, so the field is
    Hermitian. phase = 2 pi hashf(row*N + col, stepIndex); magnitude
    amp * env * FORCE_NORM / sqrt(max(dt, 1e-6)).

    Pinned by: known answers of the phase composition for modes of both halves and of the ky = 0
    row, computed in C (``force_phase``, ``canonical_half``); zero means of e^{i phase} and
    e^{2 i phase} over the annulus and zero lag-1 / lag-2 step correlation (a pi u phase, a
    hash1(idx ^ step) or a hash-then-xor composition, each planted in this oracle, fails one of
    them: tests/test_oracle_pins.py); Hermitian symmetry, annulus support, 1/sqrt(dt) scaling and
    the closed-form energy injection rate of ``injection_rate`` (tests/test_oracle_semantics.py).
    """
    g = Grid(n)
    out = np.zeros((n, n), np.complex128)
    if amp <= 0:
        return out
    env = annulus_envelope(g, k_f, width)
    rows, cols = np.nonzero(env > 1e-4)
    kx = g.kx[rows, cols]
    ky = g.ky[rows, cols]
    canon = canonical_half(kx, ky)
    hr = np.where(canon, rows, (n - rows) % n)
    hc = np.where(canon, cols, (n - cols) % n)
    sgn = np.where(canon, 1.0, -1.0)
    ph = force_phase((hr * n + hc).astype(np.uint32), step_index)
    mag = amp * env[rows, cols] * force_norm(n) / np.sqrt(max(dt, 1e-6))
    out[rows, cols] = mag * (np.cos(ph) + 1j * sgn * np.sin(ph))
    return out


def injection_rate(n, amp, k_f, width):
    """Mean energy injection rate of ``stochastic_forcing``: 1/2 (0.01 amp)^2 sum env^2/k^2.

    Each step adds dt * f_hat with |f_hat| = amp env 0.01 N^2 / sqrt(dt) and a fresh phase, so
    E = 1/2 sum |w|^2/k^2/N^4 grows on average by dt * 1/2 (0.01 amp)^2 sum env^2/k^2.
    """
    g = Grid(n)
    env = annulus_envelope(g, k_f, width)
    env = np.where(env > 1e-4, env, 0.0)
    return 0.5 * (0.01 * amp) ** 2 * float(np.sum(env ** 2 * g.inv_k2))


def amp_for_rate(eps, n, k_f, width):
    """forceAmp that gives injection rate eps (inverse of ``injection_rate``)."""
    return float(np.sqrt(eps / injection_rate(n, 1.0, k_f, width)))


# ---------------------------------------------------------------------------------------------
# pointer forcing
# ---------------------------------------------------------------------------------------------
def _sep(a, b):
    d = a - b + 0.5
    return d - np.floor(d) - 0.5


def pointer_force_physical(n, splats):
    """Curl of Gaussian body forces at cell centres (src/shaders/force.wgsl:5-27).

    x = (col + 0.5)/N, y = (row + 0.5)/N; per splat with sigma^2 = max(radius^2, 1e-8) and
    periodic separation r: 2 amp exp(-r^2/sigma^2) (vx ry - vy rx)/sigma^2 where r^2 < 9 sigma^2.
    Splats are dicts with keys x, y, dx, dy, radius, amp (src/ui/pointer.js:58-67); at most 8 are
    used (src/gpu/solver.js:9, 327).

    Pinned by a finite-difference curl of the Gaussian body force and antisymmetry in (dx, dy).
    """
    c = (np.arange(n) + 0.5) / n
    y, x = np.meshgrid(c, c, indexing="ij")
    total = np.zeros((n, n))
    for s in list(splats)[:8]:
        rx = _sep(x, s["x"])
        ry = _sep(y, s["y"])
        r2 = rx * rx + ry * ry
        sig2 = max(s["radius"] ** 2, 1e-8)
        g = np.exp(-r2 / sig2)
        term = 2.0 * s["amp"] * g * (s["dx"] * ry - s["dy"] * rx) / sig2
        total += np.where(r2 < 9.0 * sig2, term, 0.0)
    return total


def pointer_force_spectrum(n, splats):
    """f_hat = FFT2 of the physical pointer forcing (solver.js:491-494), unscaled forward."""
    return np.fft.fft2(pointer_force_physical(n, splats))


# ---------------------------------------------------------------------------------------------
# seeding and scenarios
# ---------------------------------------------------------------------------------------------
def fill_amplitude(k2, k0):
    """Mode magnitude sqrt(k^2 / (1 + (|k|/k0)^8)) of fillSpectrum before the rms scale
    (src/gpu/solver.js:398-400). Pinned by magnitude ratios between modes
    (tests/test_oracle_pins.py)."""
    return np.sqrt(k2 / (1.0 + (np.sqrt(k2) / k0) ** 8))


def fill_spectrum(n, k0=14.0, amplitude=6.0, seed=1):
    """Seeded band-limited spectrum of src/gpu/solver.js:384-418, returned as complex64 N x N.

    In-band k != 0 modes get magnitude sqrt(k^2/(1+(|k|/k0)^8)) and phase
    2 pi hash32(canonicalIdx ^ hash32(seed)) / 2^32 (canonical half ky > 0 or ky = 0, kx > 0; the
    mirror takes the conjugate), stored as float32, then every value is rescaled (FP64 product,
    float32 store) so that sqrt(sum amp^2)/N^2 = amplitude.

    Pinned by: known answers of the phase for modes of both halves and of the ky = 0 row,
    computed in C (``fill_phase``), and magnitude ratios sqrt(env(k1)/env(k2)) between modes
    (a salt of seed instead of hash32(seed), a hash-then-xor salt, or env instead of sqrt(env),
    each planted in this oracle, fails them: tests/test_oracle_pins.py); Hermitian symmetry, band
    support and the rms normalisation (Parseval: the physical field has rms = amplitude).
    """
    g = Grid(n)
    ok = g.band & (g.k2 > 0)
    amp = np.where(ok, fill_amplitude(g.k2, k0), 0.0)
    rows, cols = np.indices((n, n))
    canon = canonical_half(g.kx, g.ky)
    hr = np.where(canon, rows, (n - rows) % n)
    hc = np.where(canon, cols, (n - cols) % n)
    sgn = np.where(canon, 1.0, -1.0)
    phase = fill_phase((hr * n + hc).astype(np.uint32), seed)
    re = np.where(ok, amp * np.cos(phase), 0.0).astype(np.float32)
    im = np.where(ok, sgn * amp * np.sin(phase), 0.0).astype(np.float32)
    total = float(np.sum(amp[ok] ** 2))
    rms = np.sqrt(total) / (n * n)
    if rms > 0:
        scale = amplitude / rms
        re = (re.astype(np.float64) * scale).astype(np.float32)
        im = (im.astype(np.float64) * scale).astype(np.float32)
    return re.astype(np.complex64) + 1j * im.astype(np.complex64)


def bandlimited_spectrum(n, kmin=8.0, kmax=16.0, amplitude=6.0, seed=2):
    """Band-limited random vorticity of s2t_bandlimited_spectrum, returned as complex64 N x N.

    This build's initial condition for the decaying-turbulence case (no counterpart in the
    reference, which seeds with fillSpectrum). Every in-band mode (2/3 rule, k != 0) with
    kmin <= |k| <= kmax is drawn per canonical mode (ky > 0, or ky = 0 and kx > 0; idx = row*n +
    col of the canonical mode): u_j = (hash32((3 idx + j) ^ hash32(seed)) + 1/2) / 2^32,
    g = sqrt(-2 ln u_0) cos(2 pi u_1) (Box-Muller), phase 2 pi u_2, value g e^{i phase} (the mirror
    takes the conjugate); FP64 throughout, scaled so that the physical field has rms `amplitude`
    (sqrt(sum g^2) / N^2), rounded once to float32.

    Pinned by: Hermitian symmetry, support, rms normalisation (Parseval), k = 0 zero, the
    Gaussian ratio E|g| / sqrt(E g^2) = sqrt(2/pi), and hash known answers.
    """
    g = Grid(n)
    kr = np.sqrt(g.k2)
    sel = g.band & (g.k2 > 0) & (kr >= kmin) & (kr <= kmax)
    rows, cols = np.indices((n, n))
    canon = (g.ky > 0) | ((g.ky == 0) & (g.kx > 0))
    hr = np.where(canon, rows, (n - rows) % n)
    hc = np.where(canon, cols, (n - cols) % n)
    idx = (hr * n + hc).astype(np.uint32)
    salt = hash1(np.uint32(seed & 0xFFFFFFFF))
    u = [(hash1((np.uint32(3) * idx + np.uint32(j)) ^ salt).astype(np.float64) + 0.5)
         / 4294967296.0 for j in range(3)]
    amp = np.sqrt(-2.0 * np.log(u[0])) * np.cos(TWO_PI * u[1])
    ph = TWO_PI * u[2]
    sgn = np.where(canon, 1.0, -1.0)
    re = np.where(sel, amp * np.cos(ph), 0.0)
    im = np.where(sel, sgn * amp * np.sin(ph), 0.0)
    total = float(np.sum(np.where(sel, amp * amp, 0.0)))
    rms = np.sqrt(total) / (float(n) * n)
    scale = amplitude / rms if rms > 0 else 0.0
    return (re * scale).astype(np.float32) + 1j * (im * scale).astype(np.float32)


def _grid(n, fn):
    h = TWO_PI / n
    y, x = np.meshgrid(np.arange(n) * h, np.arange(n) * h, indexing="ij")
    return fn(x, y)


def _wrap_delta(a, b):
    d = a - b
    d = np.where(d > np.pi, d - TWO_PI, d)
    return np.where(d < -np.pi, d + TWO_PI, d)


def blobs(n, items):
    """Sum of Gaussian vortices cut at r^2 < 25 a^2 (src/scenarios.js:30-44)."""
    def fn(x, y):
        acc = np.zeros_like(x)
        for b in items:
            dx = _wrap_delta(x, b["x"])
            dy = _wrap_delta(y, b["y"])
            r2 = dx * dx + dy * dy
            s2 = b["a"] * b["a"]
            acc = acc + np.where(r2 < 25 * s2, b["amp"] * np.exp(-r2 / s2), 0.0)
        return acc
    return _grid(n, fn)


def scenario_field(name, n, cfg=None):
    """Physical initial vorticity of the reference scenarios (src/scenarios.js:57-165).

    Pinned by what the mathematics fixes (tests/test_oracle_pins.py): Taylor-Green has exactly
    the four modes (+-2, +-2) of 6/4 N^2; the dipole's moments (circulation 12 per core, the
    separation, the sign, the column, the width; a sign swap, a doubled offset and a column
    misread, planted here, fail them); the merger's circulation, centroid and second moments;
    each polygon core's circulation and centroid on its jittered radius; the shear layer's
    velocity jump, width and perturbation amplitude.
    """
    cfg = cfg or {}
    if name == "taylor-green":
        return _grid(n, lambda x, y: 6.0 * np.cos(2 * x) * np.cos(2 * y))
    if name == "dipole":
        gamma, a, d = 12.0, 0.3, 0.9
        amp = gamma / (np.pi * a * a)
        cx = np.pi * 0.6
        return blobs(n, [dict(x=cx, y=np.pi - d / 2, a=a, amp=amp),
                         dict(x=cx, y=np.pi + d / 2, a=a, amp=-amp)])
    if name == "merger":
        d = cfg.get("mergerSep", 0.7)
        return blobs(n, [dict(x=np.pi - d / 2, y=np.pi, a=0.22, amp=30.0),
                         dict(x=np.pi + d / 2, y=np.pi, a=0.22, amp=30.0)])
    if name == "polygon":
        count = int(round(cfg.get("polygonN", 6)))
        mode = max(2, count // 2)
        items = []
        for j in range(count):
            th = TWO_PI * j / count
            r = 1.25 * (1 + 0.02 * np.cos(TWO_PI * mode * j / count))
            items.append(dict(x=np.pi + r * np.cos(th), y=np.pi + r * np.sin(th), a=0.19, amp=34.0))
        return blobs(n, items)
    if name == "shear layer":
        delta = max(TWO_PI / 30, 6 * TWO_PI / n)

        def fn(x, y):
            lower = y <= np.pi
            z = np.where(lower, (y - np.pi / 2) / delta, (1.5 * np.pi - y) / delta)
            dudy = np.where(lower, 1.0, -1.0) / np.cosh(z) ** 2 / delta
            return 0.05 * np.cos(x) - dudy
        return _grid(n, fn)
    raise KeyError(name)


# ---------------------------------------------------------------------------------------------
# dye
# ---------------------------------------------------------------------------------------------
def to_half_truncating(value):
    """float32 -> binary16 bits with mantissa truncation, as toHalf (solver.js:20-41)."""
    x = np.asarray(value, np.float32).view(np.uint32).astype(np.int64)
    sign = (x >> 16) & 0x8000
    exp = (x >> 23) & 0xFF
    mant = x & 0x7FFFFF
    out = np.empty_like(x)
    nan_inf = exp == 255
    e = exp - 127 + 15
    over = (~nan_inf) & (e >= 31)
    sub = (~nan_inf) & (~over) & (e <= 0)
    tiny = sub & (e < -10)
    norm = (~nan_inf) & (~over) & (~sub)
    out[nan_inf] = sign[nan_inf] | 0x7C00 | np.where(mant[nan_inf] != 0, 0x200, 0)
    out[over] = sign[over] | 0x7C00
    out[tiny] = sign[tiny]
    keep = sub & ~tiny
    m2 = (mant[keep] | 0x800000) >> (1 - e[keep])
    out[keep] = sign[keep] | (m2 >> 13)
    out[norm] = sign[norm] | (e[norm] << 10) | (mant[norm] >> 13)
    return out.astype(np.uint16)


def paint_bands(m):
    """'bands' dye pattern of solver.js:446-473 as an (m, m, 4) float16 array.

    Pinned by landmarks the formula fixes (tests/test_oracle_pins.py): the band level 1/2 on the
    zeros of sin(12 pi fy), bright and dark band centres, pure warm and cool colours at
    fx = 1/4 and 3/4, the 0.9 factor, truncation to binary16, alpha 1 and six bands; x / y or
    warm / cool swapped, planted in its output, fail them.
    """
    fy = np.arange(m) / m
    fx = np.arange(m) / m
    band = 0.5 + 0.5 * np.tanh(6 * np.sin(fy * np.pi * 12))
    mix = 0.5 + 0.5 * np.sin(fx * np.pi * 2)
    warm = np.array([1.0, 0.45, 0.22])
    cool = np.array([0.28, 0.62, 1.0])
    rgb = band[:, None, None] * (warm[None, None, :] * mix[None, :, None]
                                 + cool[None, None, :] * (1 - mix[None, :, None])) * 0.9
    bits = np.empty((m, m, 4), np.uint16)
    bits[..., :3] = to_half_truncating(rgb.astype(np.float32))
    bits[..., 3] = to_half_truncating(np.float32(1.0))
    return bits.view(np.float16)


def _bilinear(tex, p):
    """Linear filtering with repeat addressing, texel centres at (i + 0.5)/M (FP64 weights)."""
    m = tex.shape[0]
    g = p * m - 0.5
    b = np.floor(g)
    f = g - b
    ix = b[..., 0].astype(np.int64)
    iy = b[..., 1].astype(np.int64)
    fx = f[..., 0][..., None]
    fy = f[..., 1][..., None]
    x0, x1 = ix % m, (ix + 1) % m
    y0, y1 = iy % m, (iy + 1) % m
    top = tex[y0, x0] * (1 - fx) + tex[y0, x1] * fx
    bot = tex[y1, x0] * (1 - fx) + tex[y1, x1] * fx
    return top * (1 - fy) + bot * fy


def trace(vel, p, h):
    """RK2 midpoint back-trace p - v(p - v(p) h/2) h, wrapped to [0, 1) (dye.wgsl:27-45)."""
    v1 = _bilinear(vel, p) * h
    mid = np.mod(p - 0.5 * v1 + 1.0, 1.0)
    v2 = _bilinear(vel, mid) * h
    return np.mod(p - v2 + 1.0, 1.0)


def dye_step(dye, vel, dt, fade=1.0, splats=()):
    """One BFECC dye update with the MacCormack limiter (src/shaders/dye.wgsl:58-113).

    dye: (M, M, 4) float16 (rgba), vel: (N, N, 2) physical (u, v). Intermediate fields are
    stored as float16 like the reference's rgba16float textures; maths in FP64.
    fwd:  phi1 = phi0(trace(p, +h));  back: phihat = phi1(trace(p, -h)), h = dt / 2 pi
    corr: s = trace(p, +h); clamp(1.5 phi0(s) - 0.5 phihat(s), min/max of the 4 phi0 texels
          around s) * fade + sum splat colour * dyeAmp * exp(-r^2/sigma^2); alpha = 1.

    The reference's tests hold no dye vectors, so this is pinned by closed forms and invariants:
    for a uniform half-texel shift every Fourier mode of the field is multiplied by
    A (3/2 - |A|^2 / 2) fade, A = (1 + exp(-i theta)) / 2 (the BFECC symbol of bilinear filtering:
    pins the blend, both trace directions, h and the fade); a splat's mass is the Gaussian
    integral dyeAmp pi sigma^2 times its colour; constant fields stay constant, zero velocity is
    the identity, an integer-texel uniform shift is exact, and the output respects the limiter
    bounds.
    """
    m = dye.shape[0]
    phi0 = dye.astype(np.float64)
    c = (np.arange(m) + 0.5) / m
    py, px = np.meshgrid(c, c, indexing="ij")
    p = np.stack([px, py], axis=-1)
    vel = np.asarray(vel, np.float64)
    h = dt / TWO_PI
    s_fwd = trace(vel, p, h)
    phi1 = _bilinear(phi0, s_fwd).astype(np.float16).astype(np.float64)
    s_back = trace(vel, p, -h)
    phihat = _bilinear(phi1, s_back).astype(np.float16).astype(np.float64)
    s = s_fwd
    a = _bilinear(phi0, s)
    b = _bilinear(phihat, s)
    corrected = 1.5 * a - 0.5 * b
    g = s * m - 0.5
    base = np.floor(g).astype(np.int64)
    ix, iy = base[..., 0], base[..., 1]
    x0, x1 = ix % m, (ix + 1) % m
    y0, y1 = iy % m, (iy + 1) % m
    corners = np.stack([phi0[y0, x0], phi0[y0, x1], phi0[y1, x0], phi0[y1, x1]])
    lo = corners.min(axis=0)
    hi = corners.max(axis=0)
    res = np.clip(corrected, lo, hi) * fade
    for sp in list(splats)[:8]:
        rx = _sep(px, sp["x"])
        ry = _sep(py, sp["y"])
        sig2 = max(sp["radius"] ** 2, 1e-8)
        r2 = rx * rx + ry * ry
        w = np.where(r2 < 9 * sig2, sp["dyeAmp"] * np.exp(-r2 / sig2), 0.0)
        col = np.array(list(sp["color"]) + [0.0])
        res = res + w[..., None] * col[None, None, :]
    res[..., 3] = 1.0
    return res.astype(np.float16)


def gpu_spectrum_inband(w, n):
    """Reference GPU E(k) convention (spectrum.wgsl:19-39): in-band k != 0 modes only."""
    g = Grid(n)
    ok = g.band & (g.k2 > 0)
    nb = n // 2 + 1
    dens = np.where(ok, 0.5 * np.abs(w) ** 2 * g.inv_k2 / n ** 4, 0.0)
    idx = np.minimum(np.floor(np.sqrt(g.k2) + 0.5).astype(np.int64), nb - 1)
    return np.bincount(idx.ravel(), weights=dens.ravel(), minlength=nb)[:nb]
