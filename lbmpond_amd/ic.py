"""Seeded initial conditions of the benchmark configurations, in lattice units.

The reference hard-codes its initial condition in the ``fill`` kernel (init.cu:2-108) and
loads it with ``Cell::set_base`` (init.cu:96).  Here each configuration is a host-side
generator returning the moment fields (rho, ux, uy, uz, T) as flat float64 arrays in the
reference's cell order (x fastest); ``Pond.set_state`` applies ``set_base`` on the device.

Lattice-unit mapping (one cell = dx, one step = dt, s = dt/dx in physical units):
``u_lat = s*u``, ``T_lat = s^2*T`` -- the scaling of the reference's TimeCorr kernel
(update.cu:22-25).  Ideal gas ``p = rho*T`` with ``gamma = (Dim+2)/Dim`` (init.cu:8).
Kinematic viscosity of the BGK step ``nu = (1/dtau - 1/2)*T`` (diagn.cu:19, init.cu:64).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

EC2 = 2.8569700138728056  # largest D1Q5 abscissa (consts.h:118)


@dataclass
class Case:
    """A configured run: lattice, grid, interpolator, relaxation and initial fields."""
    name: str
    lattice: str
    shape: tuple
    interp: str
    dtau: float
    steps: int
    fields: tuple  # (rho, ux, uy, uz, T) flat float64
    info: dict = field(default_factory=dict)


def _grid(shape):
    nx, ny, nz = shape
    z, y, x = np.meshgrid(np.arange(nz, dtype=np.float64), np.arange(ny, dtype=np.float64),
                          np.arange(nx, dtype=np.float64), indexing="ij")
    return x.reshape(-1), y.reshape(-1), z.reshape(-1)


def _uniform_noise(rng, n):
    return rng.uniform(-1.0, 1.0, n)


def sod_1d(nx=2000, cfl=0.5):
    """C1: D1Q5 Sod shock tube on [0,1]; left rho=1,u=0,p=1, right rho=0.125,u=0,p=0.1,
    split at x=0.5; gamma=3 (Dim=1); CFL defined on the fastest discrete speed."""
    xc = (np.arange(nx) + 0.5) / nx
    left = xc < 0.5
    rho = np.where(left, 1.0, 0.125)
    p = np.where(left, 1.0, 0.1)
    T = p / rho
    s = cfl / (0.0 + EC2 * np.sqrt(T.max() / 1.0))  # cells per unit time * dx ... dt/dx
    zeros = np.zeros(nx)
    info = dict(gamma=3.0, s=s, dx=1.0 / nx, dt=s / nx, T_scale=s * s)
    return Case("sod1d", "D1Q5", (nx, 1, 1), "weno5", 1.0 / 0.9, 1000,
                (rho, zeros.copy(), zeros.copy(), zeros.copy(), s * s * T), info)


def vortex_3d(n=256, seed=2, T0=0.01, mach=1.2, beta=5.0, radius=0.1, amp=1e-3, steps=200, lattice="D3Q27"):
    """C2: uniform Mach-1.2 flow along x plus a compressible isentropic vortex (tube along z,
    centre L/2, strength beta, radius 0.1 L) and a seeded uniform random density
    perturbation of amplitude 1e-3; gamma = 5/3."""
    g = 5.0 / 3.0
    x, y, z = _grid((n, n, n))
    R = radius * n
    xb = (x + 0.5 - 0.5 * n) / R
    yb = (y + 0.5 - 0.5 * n) / R
    r2 = xb * xb + yb * yb
    a = beta / (2 * np.pi) * np.exp(0.5 * (1 - r2))
    c0 = np.sqrt(T0)  # velocity unit of the nondimensional vortex
    ux = c0 * (mach * np.sqrt(g) - a * yb)
    uy = c0 * (a * xb)
    uz = np.zeros_like(ux)
    dT = -(g - 1) * beta * beta / (8 * g * np.pi * np.pi) * np.exp(1 - r2)
    T = T0 * (1 + dT)
    rho = (1 + dT) ** (1.0 / (g - 1))
    rng = np.random.default_rng(seed)
    rho = rho * (1 + amp * _uniform_noise(rng, rho.size))
    return Case("vortex3d", lattice, (n, n, n), "weno5", 1.0 / 0.9, steps, (rho, ux, uy, uz, T),
                dict(gamma=g, T0=T0, mach=mach, seed=seed))


def taylor_green(n=384, seed=3, T0=0.05, mach=0.5, re=1600.0, noise=1e-4, steps=500, lattice="D3Q27"):
    """C3: compressible Taylor-Green vortex, Mach 0.5, Re 1600, rho0=1, T0 (lattice units),
    L = n/(2 pi); seeded 1e-4 velocity noise; textbook pressure field."""
    g = 5.0 / 3.0
    x, y, z = _grid((n, n, n))
    k = 2 * np.pi / n
    X, Y, Z = k * (x + 0.5), k * (y + 0.5), k * (z + 0.5)
    V0 = mach * np.sqrt(g * T0)
    ux = V0 * np.sin(X) * np.cos(Y) * np.cos(Z)
    uy = -V0 * np.cos(X) * np.sin(Y) * np.cos(Z)
    uz = np.zeros_like(ux)
    p0 = 1.0 * T0
    p = p0 + 1.0 * V0 * V0 / 16.0 * (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2)
    T = np.full_like(ux, T0)
    rho = p / T0
    rng = np.random.default_rng(seed)
    ux = ux + noise * V0 * _uniform_noise(rng, ux.size)
    uy = uy + noise * V0 * _uniform_noise(rng, ux.size)
    uz = uz + noise * V0 * _uniform_noise(rng, ux.size)
    L = n / (2 * np.pi)
    nu = V0 * L / re
    dtau = 1.0 / (nu / T0 + 0.5)
    return Case("tgv", lattice, (n, n, n), "weno5", dtau, steps, (rho, ux, uy, uz, T),
                dict(gamma=g, T0=T0, V0=V0, nu=nu, re=re, seed=seed))


def rankine_hugoniot(mach, gamma, rho1=1.0, p1=1.0):
    """Post-shock state behind a shock of Mach `mach` moving into gas at rest
    (the closed form r-hugoniot.py evaluates numerically)."""
    c1 = np.sqrt(gamma * p1 / rho1)
    us = mach * c1
    rho2 = rho1 * (gamma + 1) * mach ** 2 / ((gamma - 1) * mach ** 2 + 2)
    p2 = p1 * (1 + 2 * gamma / (gamma + 1) * (mach ** 2 - 1))
    u2 = us * (1 - rho1 / rho2)
    return rho2, u2, p2, us


def shock_bubble(n=384, mach=2.0, radius=0.15, bubble_ratio=1.0 / 3.0, vmax=0.9, steps=500, lattice="D3Q27",
                 shock_x=0.25, bubble_x=0.5, width=1.0):
    """C4: planar Mach-2 shock (Rankine-Hugoniot post-shock state, gamma=5/3) moving along +x
    into rho=1, p=1, u=0, hitting a spherical light bubble (density ratio 1/3, radius 0.15 L,
    pressure equilibrium).  The lattice scale s keeps the fastest discrete speed at `vmax`
    cells per step.

    The post-shock slab occupies 0 <= x < shock_x of the periodic box, so it has a shock at
    shock_x and a second interface at the periodic boundary.  Every interface is a tanh profile
    `width` cells wide (width=0: sharp steps).  With sharp steps the gauge fixed point of cells
    on the interface is ill-conditioned in the first steps: some of them do not converge within
    the iteration cap of 100 even in the FP64 method."""
    g = 5.0 / 3.0
    rho2, u2, p2, us = rankine_hugoniot(mach, g)
    x, y, z = _grid((n, n, n))
    xc, yc, zc = (x + 0.5) / n, (y + 0.5) / n, (z + 0.5) / n
    rb = np.sqrt((xc - bubble_x) ** 2 + (yc - 0.5) ** 2 + (zc - 0.5) ** 2)
    if width > 0:
        w = width / n
        d = xc - np.round(xc)  # signed distance to the periodic boundary, in (-1/2, 1/2]
        post = 0.5 * (np.tanh(d / w) - np.tanh((d - shock_x) / w))
        inb = 0.5 * (1.0 - np.tanh((rb - radius) / w))
    else:
        post = (xc < shock_x).astype(np.float64)
        inb = (rb < radius).astype(np.float64)
    rho = 1.0 + (rho2 - 1.0) * post
    u = u2 * post
    p = 1.0 + (p2 - 1.0) * post
    rho = rho + (bubble_ratio - rho) * inb  # the bubble keeps the local pressure: T = p/rho rises
    T = p / rho
    tlat = 1.0 / 3.0 if lattice == "D3Q27" else 1.0
    emax = 1.0 if lattice == "D3Q27" else EC2
    s = vmax / (np.abs(u).max() + emax * np.sqrt(T.max() / tlat))
    zeros = np.zeros_like(xc)
    return Case("shock_bubble", lattice, (n, n, n), "weno5", 1.0 / 0.9, steps,
                (rho, s * u, zeros.copy(), zeros.copy(), s * s * T),
                dict(gamma=g, s=s, rho2=rho2, u2=u2, p2=p2, us=us, width=width))


def isotropic_turbulence(n=256, seed=5, k0=4.0, mach_t=0.5, T0=0.03, steps=300, lattice="D3Q125"):
    """C5: random-phase solenoidal velocity with E(k) ~ k^4 exp(-2 (k/k0)^2), scaled so that
    sqrt(<|u|^2>)/sqrt(gamma T0) = mach_t; uniform rho and T."""
    g = 5.0 / 3.0
    rng = np.random.default_rng(seed)
    k1 = np.fft.fftfreq(n, 1.0 / n)
    kr = np.fft.rfftfreq(n, 1.0 / n)
    KZ, KY, KX = np.meshgrid(k1, k1, kr, indexing="ij")
    K2 = KX * KX + KY * KY + KZ * KZ
    K = np.sqrt(K2)
    E = K ** 4 * np.exp(-2.0 * (K / k0) ** 2)
    amp = np.where(K > 0, np.sqrt(E / (4 * np.pi * np.maximum(K2, 1e-30))), 0.0)
    comps = []
    for _ in range(3):
        ph = rng.uniform(0, 2 * np.pi, K.shape)
        comps.append(amp * (rng.standard_normal(K.shape) + 1j * rng.standard_normal(K.shape)) * np.exp(1j * ph))
    # project onto divergence-free fields: u_hat -= k (k . u_hat) / |k|^2
    kd = (KX * comps[0] + KY * comps[1] + KZ * comps[2]) / np.where(K2 > 0, K2, 1.0)
    comps = [comps[0] - KX * kd, comps[1] - KY * kd, comps[2] - KZ * kd]
    u = [np.fft.irfftn(c, s=(n, n, n), axes=(0, 1, 2)) for c in comps]
    urms = np.sqrt(np.mean(u[0] ** 2 + u[1] ** 2 + u[2] ** 2))
    scale = mach_t * np.sqrt(g * T0) / urms
    ux, uy, uz = (scale * c.reshape(-1) for c in u)
    rho = np.ones(n ** 3)
    T = np.full(n ** 3, T0)
    return Case("turbulence", lattice, (n, n, n), "weno5", 1.0 / 0.9, steps, (rho, ux, uy, uz, T),
                dict(gamma=g, T0=T0, k0=k0, mach_t=mach_t, seed=seed))


CONFIGS = {
    "c1": sod_1d,
    "c2": vortex_3d,
    "c3": taylor_green,
    "c4": shock_bubble,
    "c5": isotropic_turbulence,
}
