"""Error measures for comparing two runs of the step on the same inputs.

`state_errors` is the relative L-inf measure of density, velocity and temperature.

`CapDependence` tracks the cells whose state may depend on an unconverged gauge iterate.  A cell
that reaches the iteration cap (steps.cu:422, 100 iterations) ends its step on whatever iterate
the fixed-point loop had reached.  Where the loop settles into a cycle (it does at some shock and
contact cells), that iterate depends on where the loop started, not only on the inputs, so two
runs that start the loop differently (the learned start, or FP32 rounding on the way) legitimately
differ there.  What a step writes to a cell depends on the previous state at the nodes of its
departure stencils only: departure points move at most about one cell per step (the lattice scale
keeps the fastest discrete speed under one cell per step) and the interpolation nodes reach two
cells beyond the nearest node, so a difference spreads by at most `reach` = 3 cells (periodic,
in each axis) per step.
"""
import numpy as np

STATE_ERRORS = ("rho", "u", "T")


def rel_linf(a, b):
    """||a-b||_inf / ||b||_inf; NaNs must coincide"""
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


def state_errors(got, ref, mask=None):
    """relative L-inf of rho and T per field and of u as a vector (max component error over the
    largest reference velocity component), over the cells where mask is true (all by default);
    the normalisation always uses the whole reference field"""
    sel = (lambda a: np.asarray(a, np.float64)) if mask is None else (lambda a: np.asarray(a, np.float64)[mask])
    if mask is not None and not mask.any():
        return {k: 0.0 for k in STATE_ERRORS}
    du = max(np.abs(sel(got[k]) - sel(ref[k])).max() for k in ("ux", "uy", "uz"))
    um = max(np.abs(np.asarray(ref[k], np.float64)).max() for k in ("ux", "uy", "uz"))
    out = {"u": float(du / um) if um > 0 else float(du)}
    for k in ("rho", "T"):
        den = np.abs(np.asarray(ref[k], np.float64)).max()
        num = np.abs(sel(got[k]) - sel(ref[k])).max()
        out[k] = float(num / den) if den > 0 else float(num)
    return {k: out[k] for k in STATE_ERRORS}


def dilate(mask, shape, reach):
    """cells within `reach` of a true cell (periodic Chebyshev distance); mask is flat, x fastest"""
    nx, ny, nz = shape
    m = mask.reshape(nz, ny, nx)
    for ax in range(3):
        if m.shape[ax] == 1:
            continue
        out = m.copy()
        for d in range(1, min(reach, m.shape[ax] // 2) + 1):
            out |= np.roll(m, d, axis=ax)
            out |= np.roll(m, -d, axis=ax)
        m = out
    return m.reshape(-1)


class CapDependence:
    """cells whose state after the steps seen so far depends on an iteration-capped cell"""

    def __init__(self, shape, max_iter, reach=3):
        self.shape = tuple(shape)
        self.max_iter = max_iter
        self.reach = reach
        self.mask = np.zeros(int(np.prod(shape)), bool)
        self.capped = []  # cells at the cap, per step

    def update(self, niter):
        """after a step: its per-cell iteration counts"""
        cap = np.asarray(niter) >= self.max_iter
        if self.mask.any():
            self.mask = dilate(self.mask, self.shape, self.reach)
        self.mask |= cap
        self.capped.append(int(cap.sum()))
