"""Initial conditions and workload shapes (numpy, host side).

Each function takes mesh coordinate arrays in the reference's original
numbering (x_cell, y_cell, x_edge, y_edge, angle_edge) and returns state
arrays in the reference's LayoutRight shapes: h[ncells, L], vn[nedges, L],
tr[ntracers, ncells, L].  The same arrays feed the HIP path and the CPU
oracle, so parity does not depend on how the initial fields are evaluated.
"""
from __future__ import annotations

import math

import numpy as np


def normal_velocity(angle_edge, vx, vy):
    """vn = n . (vx, vy) with n = (cos a, sin a), as every reference driver
    initialises vn (e.g. benchmark/benchmark.cpp:138-146)."""
    return np.cos(angle_edge)[:, None] * vx + np.sin(angle_edge)[:, None] * vy


# --------------------------------------------------------------------------
# The reference's own test cases
# --------------------------------------------------------------------------

class ManufacturedSolution:
    """test/nonlinear_swe_convergence.cpp:12-63."""

    grav = 9.80616
    f0 = 1e-4
    lx = 10000 * 1e3
    ly = math.sqrt(3) / 2 * lx
    eta0 = 1.0
    h0 = 1000.0
    mx = 2
    my = 2

    def __init__(self):
        self.kx = self.mx * (2 * math.pi / self.lx)
        self.ky = self.my * (2 * math.pi / self.ly)
        self.omega = math.sqrt(self.grav * self.h0 * (self.kx ** 2 + self.ky ** 2))

    def h(self, x, y, t):
        return self.h0 + self.eta0 * np.sin(self.kx * x + self.ky * y - self.omega * t)

    def vx(self, x, y, t):
        return self.eta0 * np.cos(self.kx * x + self.ky * y - self.omega * t)

    vy = vx

    def h_tend(self, x, y, t):
        phi = self.kx * x + self.ky * y - self.omega * t
        return self.eta0 * (-self.h0 * (self.kx + self.ky) * np.sin(phi) - self.omega * np.cos(phi)
                            + self.eta0 * (self.kx + self.ky) * np.cos(2 * phi))

    def vx_tend(self, x, y, t):
        phi = self.kx * x + self.ky * y - self.omega * t
        return self.eta0 * ((-self.f0 + self.grav * self.kx) * np.cos(phi) + self.omega * np.sin(phi)
                            - self.eta0 * (self.kx + self.ky) * np.sin(2 * phi) / 2)

    def vy_tend(self, x, y, t):
        phi = self.kx * x + self.ky * y - self.omega * t
        return self.eta0 * ((self.f0 + self.grav * self.ky) * np.cos(phi) + self.omega * np.sin(phi)
                            - self.eta0 * (self.kx + self.ky) * np.sin(2 * phi) / 2)

    def forcing(self, x_cell, y_cell, x_edge, y_edge, angle_edge, t):
        """Additional tendency of ManufacturedShallowWaterModel
        (nonlinear_swe_convergence.cpp:75-112), per cell and per edge."""
        fh = self.h_tend(x_cell, y_cell, t)
        fv = (np.cos(angle_edge) * self.vx_tend(x_edge, y_edge, t)
              + np.sin(angle_edge) * self.vy_tend(x_edge, y_edge, t))
        return fh, fv


class DoubleVortexTest:
    """test/double_vortex.cpp:8-77 (lx = ly) and benchmark/benchmark.cpp:12-81
    (ly = lx*sqrt(3)/2, selected with benchmark=True)."""

    g = 9.80616
    coriolis = 0.00006147
    h0 = 750.0
    ox = 0.1
    oy = 0.1
    dh = 75.0

    def __init__(self, benchmark: bool = False):
        self.lx = 5000000.0
        self.ly = 5000000.0 * math.sqrt(3) / 2 if benchmark else 5000000.0
        self.sigmax = 3.0 / 40.0 * self.lx
        self.sigmay = 3.0 / 40.0 * self.ly
        self.xc1 = (0.5 - self.ox) * self.lx
        self.yc1 = (0.5 - self.oy) * self.ly
        self.xc2 = (0.5 + self.ox) * self.lx
        self.yc2 = (0.5 + self.oy) * self.ly

    def _primes(self, x, y):
        pi = math.pi
        xp1 = self.lx / (pi * self.sigmax) * np.sin(pi / self.lx * (x - self.xc1))
        yp1 = self.ly / (pi * self.sigmay) * np.sin(pi / self.ly * (y - self.yc1))
        xp2 = self.lx / (pi * self.sigmax) * np.sin(pi / self.lx * (x - self.xc2))
        yp2 = self.ly / (pi * self.sigmay) * np.sin(pi / self.ly * (y - self.yc2))
        return xp1, yp1, xp2, yp2

    def h(self, x, y):
        xp1, yp1, xp2, yp2 = self._primes(x, y)
        pi = math.pi
        return self.h0 - self.dh * (np.exp(-0.5 * (xp1 * xp1 + yp1 * yp1))
                                    + np.exp(-0.5 * (xp2 * xp2 + yp2 * yp2))
                                    - 4.0 * pi * self.sigmax * self.sigmay / self.lx / self.ly)

    def vx(self, x, y):
        xp1, yp1, xp2, yp2 = self._primes(x, y)
        pi = math.pi
        ypp1 = self.ly / (2.0 * pi * self.sigmay) * np.sin(2.0 * pi / self.ly * (y - self.yc1))
        ypp2 = self.ly / (2.0 * pi * self.sigmay) * np.sin(2.0 * pi / self.ly * (y - self.yc2))
        return (-self.g * self.dh / self.coriolis / self.sigmay
                * (ypp1 * np.exp(-0.5 * (xp1 * xp1 + yp1 * yp1))
                   + ypp2 * np.exp(-0.5 * (xp2 * xp2 + yp2 * yp2))))

    def vy(self, x, y):
        xp1, yp1, xp2, yp2 = self._primes(x, y)
        pi = math.pi
        xpp1 = self.lx / (2.0 * pi * self.sigmax) * np.sin(2.0 * pi / self.lx * (x - self.xc1))
        xpp2 = self.lx / (2.0 * pi * self.sigmax) * np.sin(2.0 * pi / self.lx * (x - self.xc2))
        return (self.g * self.dh / self.coriolis / self.sigmax
                * (xpp1 * np.exp(-0.5 * (xp1 * xp1 + yp1 * yp1))
                   + xpp2 * np.exp(-0.5 * (xp2 * xp2 + yp2 * yp2))))

    def state(self, mesh, nlayers, ntracers=0, tracer_value=1000.0):
        x, y = np.asarray(mesh.x_cell), np.asarray(mesh.y_cell)
        h = np.repeat(self.h(x, y)[:, None], nlayers, axis=1)
        xe, ye = np.asarray(mesh.x_edge), np.asarray(mesh.y_edge)
        vx = np.repeat(self.vx(xe, ye)[:, None], nlayers, axis=1)
        vy = np.repeat(self.vy(xe, ye)[:, None], nlayers, axis=1)
        vn = normal_velocity(np.asarray(mesh.angle_edge), vx, vy)
        tr = np.full((ntracers, x.size, nlayers), tracer_value)
        return np.ascontiguousarray(h), np.ascontiguousarray(vn), tr


# --------------------------------------------------------------------------
# Benchmark workloads (the five configurations of the throughput metric)
# --------------------------------------------------------------------------

def dt_rule(dc, grav=9.81, h0=1000.0):
    """dt = 0.2*dc/sqrt(g*h0) (the inertia-gravity workload's rule, used for
    every workload that states no dt)."""
    return 0.2 * dc / math.sqrt(grav * h0)


def inertia_gravity_wave(mesh, nlayers, ntracers=1, seed=0, h0=1000.0, grav=9.81, f0=1e-4,
                         eta0=1.0):
    """Linear inertia-gravity wave, one wavelength per domain in x and y.

    eta = eta0*(1+0.01k)*cos(kx x + ky y) on layer k, the velocity from the
    linear plane-wave solution (u = eta0/(H|k|^2) (w kx cos - f ky sin),
    v = eta0/(H|k|^2) (w ky cos + f kx sin), w^2 = f^2 + gH|k|^2), seeded
    Gaussian noise of 1e-6*h0 on h, and tr = 1 + sin(2 pi x/Lx) sin(2 pi y/Ly).
    """
    lx, ly = mesh.period_x, mesh.period_y
    kx, ky = 2 * math.pi / lx, 2 * math.pi / ly
    k2 = kx * kx + ky * ky
    omega = math.sqrt(f0 * f0 + grav * h0 * k2)
    amp = eta0 * (1.0 + 0.01 * np.arange(nlayers))[None, :]
    xc, yc = np.asarray(mesh.x_cell), np.asarray(mesh.y_cell)
    th_c = (kx * xc + ky * yc)[:, None]
    rng = np.random.default_rng(seed)
    h = h0 + amp * np.cos(th_c) + 1e-6 * h0 * rng.standard_normal((xc.size, nlayers))
    xe, ye = np.asarray(mesh.x_edge), np.asarray(mesh.y_edge)
    th_e = (kx * xe + ky * ye)[:, None]
    s = amp / (h0 * k2)
    vx = s * (omega * kx * np.cos(th_e) - f0 * ky * np.sin(th_e))
    vy = s * (omega * ky * np.cos(th_e) + f0 * kx * np.sin(th_e))
    vn = normal_velocity(np.asarray(mesh.angle_edge), vx, vy)
    trf = 1.0 + np.sin(2 * math.pi * xc / lx) * np.sin(2 * math.pi * yc / ly)
    tr = np.repeat(np.repeat(trf[None, :, None], nlayers, axis=2), ntracers, axis=0)
    return (np.ascontiguousarray(h), np.ascontiguousarray(vn), np.ascontiguousarray(tr))


def _periodic_delta(a, c, period):
    d = a - c
    return d - period * np.round(d / period)


def gaussian_double_vortex(mesh, nlayers, ntracers=4, seed=0, h0=1000.0, dh=75.0, grav=9.81,
                           f0=1e-4, sigma_cells=8.0, centres=((0.4, 0.5), (0.6, 0.5)),
                           layer_scale=0.0):
    """Gaussian double vortex in geostrophic balance (the nx=512 workload).

    h = h0 - sum_c dH exp(-r^2/(2 sigma^2)) with periodic (minimum-image) r and
    sigma = 8*dc; u = -(g/f) dh/dy, v = (g/f) dh/dx.  layer_scale > 0 scales
    dH by (1 + layer_scale*k) on layer k.  Tracers are seeded sums of 16
    Fourier modes with |k| <= 8, rescaled to [0, 1] per layer.
    """
    lx, ly = mesh.period_x, mesh.period_y
    sigma = sigma_cells * mesh.dc
    amp = dh * (1.0 + layer_scale * np.arange(nlayers))[None, :]
    xc, yc = np.asarray(mesh.x_cell), np.asarray(mesh.y_cell)
    xe, ye = np.asarray(mesh.x_edge), np.asarray(mesh.y_edge)
    h = np.full((xc.size, nlayers), h0)
    vx = np.zeros((xe.size, nlayers))
    vy = np.zeros((xe.size, nlayers))
    for (fx, fy) in centres:
        cx, cy = fx * lx, fy * ly
        dxc, dyc = _periodic_delta(xc, cx, lx), _periodic_delta(yc, cy, ly)
        h -= amp * np.exp(-(dxc * dxc + dyc * dyc) / (2 * sigma * sigma))[:, None]
        dxe, dye = _periodic_delta(xe, cx, lx), _periodic_delta(ye, cy, ly)
        g = np.exp(-(dxe * dxe + dye * dye) / (2 * sigma * sigma))[:, None] * amp
        # dh/dx = dH*g*dx/sigma^2, dh/dy = dH*g*dy/sigma^2
        vx += -(grav / f0) * g * dye[:, None] / (sigma * sigma)
        vy += (grav / f0) * g * dxe[:, None] / (sigma * sigma)
    vn = normal_velocity(np.asarray(mesh.angle_edge), vx, vy)
    tr = fourier_tracers(mesh, nlayers, ntracers, seed)
    return (np.ascontiguousarray(h), np.ascontiguousarray(vn), tr)


def fourier_tracers(mesh, nlayers, ntracers, seed=0, nmodes=16, kmax=8):
    lx, ly = mesh.period_x, mesh.period_y
    xc, yc = np.asarray(mesh.x_cell), np.asarray(mesh.y_cell)
    rng = np.random.default_rng(seed + 1000)
    tr = np.empty((ntracers, xc.size, nlayers))
    cand = [(i, j) for i in range(-kmax, kmax + 1) for j in range(-kmax, kmax + 1)
            if 0 < i * i + j * j <= kmax * kmax]
    for l in range(ntracers):
        idx = rng.choice(len(cand), size=nmodes, replace=False)
        a = rng.uniform(0.2, 1.0, size=nmodes)
        ph = rng.uniform(0, 2 * math.pi, size=nmodes)
        base = np.zeros(xc.size)
        for m, q in enumerate(idx):
            i, j = cand[q]
            base += a[m] * np.cos(2 * math.pi * (i * xc / lx + j * yc / ly) + ph[m])
        lo, hi = base.min(), base.max()
        field = (base - lo) / (hi - lo)
        tr[l] = field[:, None]
    return np.ascontiguousarray(tr)


def gaussian_random_field(mesh, amplitude, corr_length, seed=0, cutoff=3.0):
    """Seeded stationary Gaussian random field at the cell centres of a doubly
    periodic mesh, with Gaussian covariance exp(-r^2/(2 l^2)) and standard
    deviation `amplitude` (spectral synthesis over the periodic Fourier
    modes with |k| l <= cutoff, rescaled to the requested deviation)."""
    lx, ly = mesh.period_x, mesh.period_y
    xc, yc = np.asarray(mesh.x_cell), np.asarray(mesh.y_cell)
    mmax = max(1, int(cutoff * lx / (2 * math.pi * corr_length)))
    nmax = max(1, int(cutoff * ly / (2 * math.pi * corr_length)))
    m = np.arange(-mmax, mmax + 1)
    n = np.arange(-nmax, nmax + 1)
    kx = 2 * math.pi * m / lx
    ky = 2 * math.pi * n / ly
    # spectral density of the Gaussian covariance ~ exp(-|k|^2 l^2 / 2);
    # amplitudes are its square root
    spec = np.exp(-0.25 * (kx[:, None] ** 2 + ky[None, :] ** 2) * corr_length ** 2)
    rng = np.random.default_rng(seed)
    coef = spec * (rng.standard_normal(spec.shape) + 1j * rng.standard_normal(spec.shape))
    ex = np.exp(1j * np.outer(xc, kx))            # [nc][M]
    ey = np.exp(1j * np.outer(yc, ky))            # [nc][N]
    field = np.real(np.einsum("cm,mn,cn->c", ex, coef, ey, optimize=True))
    field -= field.mean()
    return amplitude * field / field.std()


def adjoint_observations(mesh, cfg, seed=0):
    """Initial state of the observation run of the adjoint workload: the
    initial state with h perturbed by a Gaussian random field of 1e-2 m
    standard deviation and correlation length 5*dc (same on every layer)."""
    h, vn, tr = make_state(mesh, cfg, seed=seed)
    pert = gaussian_random_field(mesh, 1e-2, 5 * cfg["dc"], seed=seed + 17)
    return np.ascontiguousarray(h + pert[:, None]), vn, tr


def ensemble_members(nmembers=48, seed=0, nx_choices=(32, 48, 64, 96, 128, 192),
                     layer_choices=(16, 32, 64), max_tracers=4, dc=10e3):
    """Seeded member list of the ensemble workload: nx from nx_choices,
    layers from layer_choices, 1..max_tracers tracers, double-vortex centres
    offset by up to +-0.05 of the domain and dH uniform in [50, 100] m."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(nmembers):
        nx = int(rng.choice(nx_choices))
        out.append(dict(
            nx=nx, dc=dc, nlayers=int(rng.choice(layer_choices)),
            ntracers=int(rng.integers(1, max_tracers + 1)),
            centres=((0.4 + rng.uniform(-0.05, 0.05), 0.5 + rng.uniform(-0.05, 0.05)),
                     (0.6 + rng.uniform(-0.05, 0.05), 0.5 + rng.uniform(-0.05, 0.05))),
            dh=float(rng.uniform(50.0, 100.0)), seed=seed * 1000 + i))
    return out


# Workload table of the throughput metric -----------------------------------
CONFIGS = {
    "c1": dict(nx=64, dc=10e3, nlayers=64, ntracers=1, stepper="rk4", nsteps=10,
               init="inertia_gravity_wave", visc_del2=0.0, eddy_diff4=0.0, f0=1e-4, grav=9.81),
    "c2": dict(nx=512, dc=5e3, nlayers=64, ntracers=4, stepper="rk4", nsteps=100,
               init="gaussian_double_vortex", visc_del2=1e3, eddy_diff4=1e9, f0=1e-4, grav=9.81),
    "c3": dict(nx=256, dc=10e3, nlayers=32, ntracers=2, stepper="rk4", nsteps=200,
               init="inertia_gravity_wave", visc_del2=1e3, eddy_diff4=0.0, f0=1e-4, grav=9.81),
    # ensemble adjoint: 16 members on one mesh, shared del2 viscosity and
    # tracer del4, checkpoints capped at 8 GB
    "c5": dict(nx=128, dc=10e3, nlayers=64, ntracers=2, stepper="rk4", nsteps=500,
               init="inertia_gravity_wave", visc_del2=1e3, eddy_diff4=1e9, f0=1e-4, grav=9.81,
               nmembers=16, byte_cap=8_000_000_000),
}


def make_state(mesh, cfg, seed=0):
    if cfg["init"] == "inertia_gravity_wave":
        return inertia_gravity_wave(mesh, cfg["nlayers"], cfg["ntracers"], seed=seed,
                                    grav=cfg["grav"], f0=cfg["f0"])
    return gaussian_double_vortex(mesh, cfg["nlayers"], cfg["ntracers"], seed=seed,
                                  grav=cfg["grav"], f0=cfg["f0"])
