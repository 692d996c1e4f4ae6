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
        spec = np.fft.fft2(np.asarray(field, np.float64))
        spec = np.where(self.g.band, spec, 0.0)
        spec[0, 0] = 0.0
        self.w = spec.astype(np.complex128)

    def physical(self, w=None):
        return np.fft.ifft2(self.w if w is None else w).real

    # -- right-hand side -----------------------------------------------------------------------
    def packed_spectra(self, w):
        """A = w_hat + i u_hat and B = v_hat (solver.wgsl velocity_pack; spectral2d.py:45-50)."""
        g = self.g
        uh = 1j * g.ky * g.inv_k2 * w
        vh = -1j * g.kx * g.inv_k2 * w
        return w + 1j * uh, vh

    def fields(self, w):
        """Physical omega, u, v from one packed inverse transform plus one for v."""
        a, b = self.packed_spectra(w)
        pa = np.fft.ifft2(a)
        pb = np.fft.ifft2(b)
        return pa.real, pa.imag, pb.real

    def velocity(self, w=None):
        _, u, v = self.fields(self.w if w is None else w)
        return u, v

    def unpack(self, z):
        """Hermitian unpack of a transform of two real fields (solver.wgsl assemble:152-155)."""
        m = self.g.mirror
        zm = np.conj(z[np.ix_(m, m)])
        return 0.5 * (z + zm), -0.5j * (z - zm)

    def product_transform(self, w):
        """Z_hat = FFT(u w + i v w): the one forward transform of the packed product."""
        omega, u, v = self.fields(w)
        return np.fft.fft2(u * omega + 1j * (v * omega))

    def nhat_from_z(self, z):
        """N_hat = (i kx q1 + i ky q2) masked, from the packed product spectrum."""
        q1, q2 = self.unpack(z)
        g = self.g
        return np.where(g.band, 1j * g.kx * q1 + 1j * g.ky * q2, 0.0)

    def nonlinear(self, w):
        return self.nhat_from_z(self.product_transform(w))

    def nonlinear_direct(self, w):
        """Unpacked evaluation with two separate forward transforms (spectral2d.py:139)."""
        omega, u, v = self.fields(w)
        g = self.g
        nh = 1j * g.kx * np.fft.fft2(u * omega) + 1j * g.ky * np.fft.fft2(v * omega)
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
        return np.exp(self.lin * s)

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
        w1 = self.ifactor(dt) * (w0 + dt * self.rhs(w0, forcing))
        w2 = 0.75 * self.ifactor(0.5 * dt) * w0 + 0.25 * self.ifactor(-0.5 * dt) * (
            w1 + dt * self.rhs(w1, forcing))
        w3 = (1.0 / 3.0) * self.ifactor(dt) * w0 + (2.0 / 3.0) * self.ifactor(0.5 * dt) * (
            w2 + dt * self.rhs(w2, forcing))
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
