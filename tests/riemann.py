"""Exact solution of the 1D Riemann problem for an ideal gas (Toro, "Riemann Solvers and
Numerical Methods for Fluid Dynamics", ch. 4): Newton iteration for the star pressure, then
sampling of the self-similar solution.  Used to pin the composed PonD step physically."""
import numpy as np


def _f(p, rho, pk, ck, g):
    if p > pk:  # shock
        A = 2.0 / ((g + 1) * rho)
        B = (g - 1) / (g + 1) * pk
        s = np.sqrt(A / (p + B))
        return (p - pk) * s, s * (1 - (p - pk) / (2 * (B + p)))
    # rarefaction
    r = p / pk
    return 2 * ck / (g - 1) * (r ** ((g - 1) / (2 * g)) - 1), 1.0 / (rho * ck) * r ** (-(g + 1) / (2 * g))


def star_state(rl, ul, pl, rr, ur, pr, g):
    cl, cr = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)
    p = max(1e-8, 0.5 * (pl + pr))
    for _ in range(100):
        fl, dl = _f(p, rl, pl, cl, g)
        fr, dr = _f(p, rr, pr, cr, g)
        dp = (fl + fr + ur - ul) / (dl + dr)
        p_new = max(1e-10, p - dp)
        if abs(p_new - p) < 1e-14 * p:
            p = p_new
            break
        p = p_new
    fl, _ = _f(p, rl, pl, cl, g)
    fr, _ = _f(p, rr, pr, cr, g)
    u = 0.5 * (ul + ur) + 0.5 * (fr - fl)
    return p, u


def sample(xi, rl, ul, pl, rr, ur, pr, g):
    """rho, u, p at similarity coordinate xi = (x - x0)/t (arrays)."""
    ps, us = star_state(rl, ul, pl, rr, ur, pr, g)
    cl, cr = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)
    xi = np.asarray(xi, np.float64)
    rho, u, p = np.zeros_like(xi), np.zeros_like(xi), np.zeros_like(xi)
    for k, s in enumerate(xi):
        if s <= us:  # left of contact
            if ps > pl:  # left shock
                sl = ul - cl * np.sqrt((g + 1) / (2 * g) * ps / pl + (g - 1) / (2 * g))
                if s <= sl:
                    rho[k], u[k], p[k] = rl, ul, pl
                else:
                    rho[k] = rl * (ps / pl + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pl + 1)
                    u[k], p[k] = us, ps
            else:  # left rarefaction
                shl = ul - cl
                cls = cl * (ps / pl) ** ((g - 1) / (2 * g))
                stl = us - cls
                if s <= shl:
                    rho[k], u[k], p[k] = rl, ul, pl
                elif s >= stl:
                    rho[k], u[k], p[k] = rl * (ps / pl) ** (1 / g), us, ps
                else:
                    c = 2 / (g + 1) * (cl + (g - 1) / 2 * (ul - s))
                    u[k] = 2 / (g + 1) * (cl + (g - 1) / 2 * ul + s)
                    rho[k] = rl * (c / cl) ** (2 / (g - 1))
                    p[k] = pl * (c / cl) ** (2 * g / (g - 1))
        else:  # right of contact
            if ps > pr:  # right shock
                sr = ur + cr * np.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g))
                if s >= sr:
                    rho[k], u[k], p[k] = rr, ur, pr
                else:
                    rho[k] = rr * (ps / pr + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pr + 1)
                    u[k], p[k] = us, ps
            else:  # right rarefaction
                shr = ur + cr
                crs = cr * (ps / pr) ** ((g - 1) / (2 * g))
                str_ = us + crs
                if s >= shr:
                    rho[k], u[k], p[k] = rr, ur, pr
                elif s <= str_:
                    rho[k], u[k], p[k] = rr * (ps / pr) ** (1 / g), us, ps
                else:
                    c = 2 / (g + 1) * (cr - (g - 1) / 2 * (ur - s))
                    u[k] = 2 / (g + 1) * (-cr + (g - 1) / 2 * ur + s)
                    rho[k] = rr * (c / cr) ** (2 / (g - 1))
                    p[k] = pr * (c / cr) ** (2 * g / (g - 1))
    return rho, u, p
