"""FP64 restatement of the pseudo-spectral vorticity scheme (test infrastructure only).

The scheme (reference README.md "The scheme" and "Time integration"; reference/spectral2d.py):

* state: the vorticity spectrum w_hat on an N x N grid, row = ky, col = kx, both laid out as
  ``numpy.fft.fftfreq(N, 1/N)`` (Nyquist at -N/2)            -- spectral2d.py:16-18
* velocity: u_hat = i ky w_hat / k^2, v_hat = -i kx w_hat / k^2, zero at k = 0 -- spectral2d.py:40-43
* nonlinear term from the packed transforms A = w_hat + i u_hat, B = v_hat; the product
  Z = u w + i v w is transformed once and the two real spectra are unpacked by Hermitian
  symmetry; N_hat = i kx q1 + i ky q2, masked to the 2/3 band   -- spectral2d.py:45-58
* linear operator L = -mu - nu_h (k^2)^p + i beta kx / k^2, L(0) = -mu -- spectral2d.py:26-27
* integrating-factor SSP-RK3 with E(s) = exp(L s)                -- spectral2d.py:73-90
* CFL time step from max(|u| + |v|)                               -- spectral2d.py:66-71
* diagnostics E, Z, E(k), zonal-mean u                           -- spectral2d.py:92-108

Pinned by: tests/golden/fixtures.json (all five keys, to 1e-12) and the closed forms and
invariants of reference/spectral2d.py:132-229 (packed vs direct N_hat, div u = 0, curl u = w,
inviscid E/Z conservation, single-mode Rossby phase, Taylor-Green N = 0 and exp decay).
"""

import numpy as np

TWO_PI = 2.0 * np.pi

_WORKERS = 1
_POOL = None


def set_fft_workers(k):
    """Threads of the scheme: 1 (the default) runs everything on one thread with numpy.fft, the
    CPU baseline's setting; k > 1 runs the 2-D transforms with scipy.fft on k workers and the
    pointwise stages over k row bands on a thread pool (NumPy releases the GIL inside its
    loops), which the tests use to run long FP64 comparisons in reasonable time. The arithmetic
    of every element is the same either way; the transforms are pocketfft in both libraries and
    agree to FP64 rounding."""
    global _WORKERS, _POOL
    _WORKERS = max(1, int(k))
    if _POOL is not None:
        _POOL.shutdown()
        _POOL = None


def _rows(fn, n, count=1):
    """fn(r0, r1) -> `count` complex arrays of rows [r0, r1). On one thread fn runs once over all
    rows (the plain NumPy expressions); with k workers it runs over k row bands on a thread pool
    and the bands are written into complex128 outputs of n rows."""
    global _POOL
    if _WORKERS == 1 or n < 512:  # small grids: thread overhead exceeds the work
        return fn(0, n)
    if _POOL is None:
        from concurrent.futures import ThreadPoolExecutor
        _POOL = ThreadPoolExecutor(max_workers=_WORKERS)
    outs = [np.empty((n, n), np.complex128) for _ in range(count)]

    def band(b):
        res = fn(b[0], b[1])
        for o, r in zip(outs, res if count > 1 else (res,)):
            o[b[0]:b[1]] = r
    edges = np.linspace(0, n, _WORKERS + 1).astype(int)
    list(_POOL.map(band, zip(edges[:-1], edges[1:])))
    return tuple(outs) if count > 1 else outs[0]


def fft2(a):
    if _WORKERS > 1 and a.shape[0] >= 512:
        import scipy.fft
        return scipy.fft.fft2(a, workers=_WORKERS)
    return np.fft.fft2(a)


def ifft2(a):
    if _WORKERS > 1 and a.shape[0] >= 512:
        import scipy.fft
        return scipy.fft.ifft2(a, workers=_WORKERS)
    return np.fft.ifft2(a)


class Grid:
    """Wavenumbers, dealias mask and inverse Laplacian for an N x N periodic grid.

    Conventions follow reference/spectral2d.py:14-24 (kx varies along columns, ky along rows,
    dealias band |kx| < N/3 and |ky| < N/3, strict).
    """

    def __init__(self, n):
        self.n = int(n)
        freq = np.fft.fftfreq(self.n, 1.0 / self.n)
        self.ky, self.kx = np.meshgrid(freq, freq, indexing="ij")
        self.k2 = self.kx * self.kx + self.ky * self.ky
        safe = np.where(self.k2 > 0.0, self.k2, 1.0)
        self.inv_k2 = np.where(self.k2 > 0.0, 1.0 / safe, 0.0)
        self.kcut = self.n / 3.0
        self.band = (np.abs(self.kx) < self.kcut) & (np.abs(self.ky) < self.kcut)
        self.mirror = (-np.arange(self.n)) % self.n
        self.dx = TWO_PI / self.n


def nu_from_rate(rate, n, p):
    """Hyperviscosity coefficient from a damping rate at the cut-off wavenumber.

    nu_h = rate / kcut^(2p) with kcut = N/3 (src/gpu/solver.js:307-311); with rate = 1/tau this
    is reference/spectral2d.py:24. Pinned by fixtures.json ``step.nuH`` (tau = 1.5).
    """
    if rate <= 0:
        return 0.0
    return rate / (n / 3.0) ** (2 * p)


class Scheme:
    """Integrating-factor SSP-RK3 pseudo-spectral solver in FP64."""

    def __init__(self, n, beta=0.0, mu=0.0, nu=0.0, p=4, cfl=0.4, dt_max=0.05):
        self.g = Grid(n)
        self.n = self.g.n
        self.beta = float(beta)
        self.mu = float(mu)
        self.nu = float(nu)
        self.p = p
        self.cfl = float(cfl)
        self.dt_max = float(dt_max)
        g = self.g
        self.lin = (-self.mu - self.nu * g.k2 ** self.p) + 1j * (self.beta * g.kx * g.inv_k2)
        self.lin[0, 0] = -self.mu
        self.w = np.zeros((self.n, self.n), np.complex128)
        self.t = 0.0

    # -- state ---------------------------------------------------------------------------------
    def set_physical(self, field):
        """Transform a real field, mask it to the band and zero the mean (spectral2d.py:33-35)."""
        spec = fft2(np.asarray(field, np.float64))
        spec = np.where(self.g.band, spec, 0.0)
        spec[0, 0] = 0.0
        self.w = spec.astype(np.complex128)

    def physical(self, w=None):
        return ifft2(self.w if w is None else w).real

    # -- right-hand side -----------------------------------------------------------------------
    def packed_spectra(self, w):
        """A = w_hat + i u_hat and B = v_hat (solver.wgsl velocity_pack; spectral2d.py:45-50)."""
        g = self.g

        def band(r0, r1):
            ws = w[r0:r1]
            uh = 1j * g.ky[r0:r1] * g.inv_k2[r0:r1] * ws
            vh = -1j * g.kx[r0:r1] * g.inv_k2[r0:r1] * ws
            return ws + 1j * uh, vh
        return _rows(band, self.n, 2)

    def fields(self, w):
        """Physical omega, u, v from one packed inverse transform plus one for v."""
        a, b = self.packed_spectra(w)
        pa = ifft2(a)
        pb = ifft2(b)
        return pa.real, pa.imag, pb.real

    def velocity(self, w=None):
        _, u, v = self.fields(self.w if w is None else w)
        return u, v

    def unpack(self, z):
        """Hermitian unpack of a transform of two real fields (solver.wgsl assemble:152-155)."""
        m = self.g.mirror

        def band(r0, r1):
            zs = z[r0:r1]
            zm = np.conj(z[np.ix_(m[r0:r1], m)])
            return 0.5 * (zs + zm), -0.5j * (zs - zm)
        return _rows(band, self.n, 2)

    def product_transform(self, w):
        """Z_hat = FFT(u w + i v w): the one forward transform of the packed product."""
        omega, u, v = self.fields(w)

        def band(r0, r1):
            om = omega[r0:r1]
            return u[r0:r1] * om + 1j * (v[r0:r1] * om)
        return fft2(_rows(band, self.n))

    def nhat_from_z(self, z):
        """N_hat = (i kx q1 + i ky q2) masked, from the packed product spectrum."""
        q1, q2 = self.unpack(z)
        g = self.g

        def band(r0, r1):
            return np.where(g.band[r0:r1], 1j * g.kx[r0:r1] * q1[r0:r1]
                            + 1j * g.ky[r0:r1] * q2[r0:r1], 0.0)
        return _rows(band, self.n)

    def nonlinear(self, w):
        return self.nhat_from_z(self.product_transform(w))

    def nonlinear_direct(self, w):
        """Unpacked evaluation with two separate forward transforms (spectral2d.py:139)."""
        omega, u, v = self.fields(w)
        g = self.g
        nh = 1j * g.kx * fft2(u * omega) + 1j * g.ky * fft2(v * omega)
        return np.where(g.band, nh, 0.0)

    def rhs(self, w, forcing=None):
        f = -self.nonlinear(w)
        if forcing is not None:
            f = f + forcing
        return f

    # -- time stepping -------------------------------------------------------------------------
    def cfl_dt(self, w=None):
        """dt = min(dt_max, cfl dx / max(|u|+|v|)), dt_max when the flow is at rest (<= 1e-12)."""
        u, v = self.velocity(w)
        speed = float(np.max(np.abs(u) + np.abs(v)))
        if speed <= 1e-12:
            return self.dt_max
        return min(self.dt_max, self.cfl * self.g.dx / speed)

    def ifactor(self, s):
        return _rows(lambda r0, r1: np.exp(self.lin[r0:r1] * s), self.n)

    def _combine(self, ca, ea, w0, cb, eb, wc, dt, f):
        """ca ea w0 + cb eb (wc + dt f) over row bands (ca = 0: the first stage)."""

        def band(r0, r1):
            x = cb * eb[r0:r1] * (wc[r0:r1] + dt * f[r0:r1])
            return x if ca == 0.0 else ca * ea[r0:r1] * w0[r0:r1] + x
        return _rows(band, self.n)

    def step(self, dt=None, forcing=None):
        """One IF-SSP-RK3 step; k = 0 is zeroed at the end (spectral2d.py:76-90).

        Stages (RK coefficients of src/gpu/solver.js:11-15, [cA, a, cB, b]):
          w1 = E(dt) (w0 + dt f0)
          w2 = 3/4 E(dt/2) w0 + 1/4 E(-dt/2) (w1 + dt f1)
          w3 = 1/3 E(dt) w0 + 2/3 E(dt/2) (w2 + dt f2)
        """
        if dt is None:
            dt = self.cfl_dt()
        w0 = self.w
        w1 = self._combine(0.0, None, w0, 1.0, self.ifactor(dt), w0, dt, self.rhs(w0, forcing))
        w2 = self._combine(0.75, self.ifactor(0.5 * dt), w0, 0.25, self.ifactor(-0.5 * dt), w1, dt,
                           self.rhs(w1, forcing))
        w3 = self._combine(1.0 / 3.0, self.ifactor(dt), w0, 2.0 / 3.0, self.ifactor(0.5 * dt), w2,
                           dt, self.rhs(w2, forcing))
        w3[0, 0] = 0.0
        self.w = w3
        self.t += dt
        return dt

    # -- diagnostics ---------------------------------------------------------------------------
    def energy(self, w=None):
        w = self.w if w is None else w
        return 0.5 * float(np.sum(np.abs(w) ** 2 * self.g.inv_k2)) / self.n ** 4

    def enstrophy(self, w=None):
        w = self.w if w is None else w
        return 0.5 * float(np.sum(np.abs(w) ** 2)) / self.n ** 4

    def spectrum(self, w=None):
        """Radial E(k): bins round(|k|) clipped to N/2 (spectral2d.py:98-104)."""
        w = self.w if w is None else w
        nb = self.n // 2 + 1
        idx = np.clip(np.round(np.sqrt(self.g.k2)).astype(np.int64), 0, self.n // 2)
        dens = 0.5 * np.abs(w) ** 2 * self.g.inv_k2 / self.n ** 4
        return np.bincount(idx.ravel(), weights=dens.ravel(), minlength=nb)[:nb]

    def zonal_mean_u(self, w=None):
        u, _ = self.velocity(w)
        return u.mean(axis=1)


def random_field(n, k0=14.0, seed=0, amplitude=1.0):
    """Seeded random-phase field with envelope k^2/(1+(k/k0)^8), zero mean, std = amplitude.

    Same construction as reference/spectral2d.py:111-123 (numpy default_rng, one uniform phase
    per grid point in row-major order). Pinned by fixtures.json ``velocity.what`` (seed 3).
    """
    rng = np.random.default_rng(seed)
    g = Grid(n)
    kr = np.sqrt(g.k2)
    env = np.where(g.k2 > 0, g.k2 / (1.0 + (kr / k0) ** 8), 0.0)
    ph = rng.uniform(0.0, TWO_PI, (n, n))
    field = np.fft.ifft2(np.sqrt(env) * np.exp(1j * ph)).real
    field = field - field.mean()
    return amplitude * field / field.std()


def grid_coords(n):
    """Physical coordinates x_j = 2 pi j / N (src/scenarios.js:3-13)."""
    xs = np.arange(n) * (TWO_PI / n)
    gy, gx = np.meshgrid(xs, xs, indexing="ij")
    return gx, gy
