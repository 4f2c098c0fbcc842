"""Shared test helpers: seeded input fields and error measures."""
import numpy as np


def rel_linf(a, b):
    """||a-b||_inf / ||b||_inf (NaN-aware: NaNs must coincide)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    na, nb = np.isnan(a), np.isnan(b)
    if (na != nb).any():
        return np.inf
    a, b = a[~na], b[~nb]
    if a.size == 0:
        return 0.0
    den = np.abs(b).max()
    num = np.abs(a - b).max()
    return float(num / den) if den > 0 else float(num)


STATE_ERRORS = ("rho", "u", "T")


def state_errors(got, ref):
    """Relative L-inf errors of density, velocity and temperature: rho and T per field, the
    velocity as a vector, max_k |u_k - u_k,ref| / max_k |u_k,ref| over all cells, so that a
    component that is zero in the reference (the axial velocity of a vortex tube) is measured
    against the flow speed rather than against its own round-off."""
    du = max(np.abs(np.asarray(got[k], np.float64) - np.asarray(ref[k], np.float64)).max() for k in ("ux", "uy", "uz"))
    um = max(np.abs(np.asarray(ref[k], np.float64)).max() for k in ("ux", "uy", "uz"))
    return {"rho": rel_linf(got["rho"], ref["rho"]), "u": float(du / um) if um > 0 else float(du),
            "T": rel_linf(got["T"], ref["T"])}


def smooth_random_state(shape, seed, T0=0.02, u_amp=0.05, rho_amp=0.1, dims=3):
    """Seeded smooth random (rho, u, T): a few random Fourier modes per field."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")

    def field(amp):
        out = np.zeros((nz, ny, nx))
        for _ in range(3):
            kx = rng.integers(0, 3)
            ky = rng.integers(0, 3) if ny > 1 else 0
            kz = rng.integers(0, 3) if nz > 1 else 0
            ph = rng.uniform(0, 2 * np.pi)
            out += amp * rng.uniform(-1, 1) * np.cos(2 * np.pi * (kx * x / nx + ky * y / ny + kz * z / nz) + ph)
        return out.reshape(-1)

    rho = 1.0 + field(rho_amp)
    ux = field(u_amp)
    uy = field(u_amp) if dims == 3 else np.zeros(nx * ny * nz)
    uz = field(u_amp) if dims == 3 else np.zeros(nx * ny * nz)
    T = T0 * (1.0 + field(0.2))
    return rho, ux, uy, uz, T


FIELDS = ("rho", "ux", "uy", "uz", "T")


def x_dominant_state(shape, seed, T0=0.01, yz_amp=1e-3):
    """Seeded state varying mainly along x, with a small longest-wavelength y/z component.
    Used for the as-written WENO5, whose y/z stencil amplifies short y/z wavelengths."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    z, y, x = [a.reshape(-1) for a in np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")]
    yz = yz_amp * np.cos(2 * np.pi * y / ny + 0.3) * np.cos(2 * np.pi * z / nz + 1.1)
    rho = 1 + 0.1 * np.cos(2 * np.pi * x / nx + rng.uniform(0, 6)) + yz
    ux = 0.05 * np.sin(2 * np.pi * x / nx + rng.uniform(0, 6))
    uy = 0.02 + yz
    uz = -0.01 + yz
    T = T0 * (1 + 0.1 * np.cos(2 * np.pi * x / nx + rng.uniform(0, 6)) + yz)
    return rho, ux, uy, uz, T
