"""Mirror boundaries (the reference's outflow switch: steps.cu:167-190 in PonDstepD3Q125,
1004-1022 in propagate<N>; Cell::mirrX/Y/Z, structs.cuh:552-578) on the GPU, against the
oracle (whose mirror path is pinned in test_oracle_pins.py against the doubled periodic domain).

FP64: bit for bit with identical iteration counts; split: within 1e-5 of FP64.  The reference
has no mirror path in its WENO5 interpolation, and a WENO5 context with a mirror is refused.
"""
import numpy as np
import pytest

import pyoracle as po
from helpers import FIELDS, STATE_ERRORS, smooth_random_state
from lbmpond_amd import Pond, PondError
from lbmpond_amd import ic
from lbmpond_amd.compare import state_errors


def wall_state(shape, T0=0.02, amp=0.05, seed=0):
    """A state that mirror walls fit: density and temperature even about each face (cosines of
    period twice the box) and every velocity component odd about the faces normal to it (sines),
    so no flow enters a wall.  Random mode amplitudes from the seed."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    z, y, x = [a.reshape(-1).astype(np.float64) for a in np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx),
                                                                    indexing="ij")]
    cx, cy, cz = (np.cos(np.pi * (c + 0.5) / n) for c, n in ((x, nx), (y, ny), (z, nz)))
    sx, sy, sz = (np.sin(np.pi * (c + 0.5) / n) for c, n in ((x, nx), (y, ny), (z, nz)))
    a = rng.uniform(0.5, 1.0, 6)
    rho = 1 + amp * a[0] * cx * cy * cz
    T = T0 * (1 + amp * a[1] * cx + amp * a[2] * cy * cz)
    ux = amp * np.sqrt(T0) * a[3] * sx * cy
    uy = amp * np.sqrt(T0) * a[4] * cx * sy if ny > 1 else np.zeros_like(x)
    uz = amp * np.sqrt(T0) * a[5] * sz * cy if nz > 1 else np.zeros_like(x)
    return rho, ux, uy, uz, T


def oracle_run(lattice, shape, interp, fields, steps, mirror, **kw):
    o = po.Oracle(lattice, shape, interp, mirror=mirror, **kw)
    o.set_state(*fields)
    o.step(steps)
    return o.fields()


def pond_run(lattice, shape, interp, precision, fields, steps, mirror, **kw):
    with Pond(lattice, shape, interp, precision, mirror=mirror, **kw) as p:
        p.set_state(*fields)
        p.step(steps)
        return p.fields()


def test_mirror_with_weno5_is_refused():
    for interp in ("weno5", "weno5c"):
        with pytest.raises(PondError, match="mirror"):
            Pond("D3Q27", (8, 6, 5), interp, "fp64", mirror=1)
    with pytest.raises(PondError, match="mirror"):
        Pond("D3Q27", (8, 6, 5), "lagr4", "fp64", mirror=8)


@pytest.mark.gpu
@pytest.mark.parametrize("lattice,shape,mirror,steps,kw", [
    ("D1Q5", (200, 1, 1), 1, 100, {}),
    ("D3Q27", (12, 8, 6), 7, 20, {}),
    ("D3Q27", (10, 7, 5), 2, 20, {}),
    ("D3Q27", (64, 8, 8), 5, 3, {}),
    ("D3Q27", (37, 9, 7), 7, 3, {"refill_chunk": 0, "pass1_iters": 2}),
    ("D3Q125", (5, 4, 3), 7, 20, {}),
])
def test_fp64_mirror_matches_oracle(gpu, lattice, shape, mirror, steps, kw):
    interp = "weno5" if lattice == "D3Q125" else "lagr4"
    f = wall_state(shape, seed=101)
    if lattice == "D1Q5":
        f = (f[0], f[1], 0 * f[1], 0 * f[1], f[4])
    print(f"[mirror f64] {lattice} {shape} mirror={mirror} steps={steps} {kw}", flush=True)
    ref = oracle_run(lattice, shape, interp, f, steps, mirror)
    got = pond_run(lattice, shape, interp, "fp64", f, steps, mirror, **kw)
    for k in FIELDS + ("niter",):
        assert np.array_equal(got[k], ref[k]), k
    assert np.isfinite(ref["rho"]).all() and (ref["niter"] < 100).all()
    # the walls change the result against the periodic run
    per = oracle_run(lattice, shape, interp, f, steps, 0)
    assert np.abs(per["rho"] - ref["rho"]).max() > 1e-8


@pytest.mark.gpu
def test_fp64_mirror_arbitrary_state_matches_oracle(gpu):
    """a state that does not fit the walls (flow into them): the same arithmetic, bit for bit,
    over the first steps (later the reference's own scheme heats the wall cells without bound)"""
    shape = (12, 8, 6)
    f = smooth_random_state(shape, seed=103, T0=0.02)
    ref = oracle_run("D3Q27", shape, "lagr4", f, 3, 7)
    got = pond_run("D3Q27", shape, "lagr4", "fp64", f, 3, 7)
    for k in FIELDS + ("niter",):
        assert np.array_equal(got[k], ref[k]), k


@pytest.mark.gpu
def test_sod_tube_between_mirror_walls(gpu):
    """D1Q5 Sod tube closed by mirror walls (Lagrange-4, 100 steps): bit for bit with the
    oracle; the walls reflect, so the mass in the tube stays within the interpolation's
    round-off of its initial value over the run"""
    case = ic.sod_1d(400)
    ref = oracle_run("D1Q5", case.shape, "lagr4", case.fields, 100, 1, dtau=case.dtau)
    got = pond_run("D1Q5", case.shape, "lagr4", "fp64", case.fields, 100, 1, dtau=case.dtau)
    for k in FIELDS + ("niter",):
        assert np.array_equal(got[k], ref[k]), k
    m0 = case.fields[0].sum()
    print(f"  mass drift {abs(got['rho'].sum() - m0) / m0:.2e}", flush=True)
    assert abs(got["rho"].sum() - m0) <= 1e-3 * m0


@pytest.mark.gpu
@pytest.mark.parametrize("lattice,shape,mirror", [("D3Q27", (16, 12, 10), 7), ("D3Q125", (6, 5, 4), 7)])
def test_split_mirror_matches_fp64(gpu, lattice, shape, mirror):
    """the split step with mirror walls against FP64: one step from the same FP64 state (after 5
    steps) within 1e-5 on every cell, and 5 steps from the same initial state within 1e-5"""
    interp = "weno5" if lattice == "D3Q125" else "lagr4"
    f = wall_state(shape, seed=102, T0=0.01)
    with Pond(lattice, shape, interp, "fp64", mirror=mirror) as a, \
            Pond(lattice, shape, interp, "split", mirror=mirror) as b, \
            Pond(lattice, shape, interp, "split", mirror=mirror) as c:
        a.set_state(*f)
        b.set_state(*f)
        a.step(5)
        b.step(5)
        traj = state_errors(b.fields(), a.fields())
        c.copy_state_from(a)
        a.step(1)
        c.step(1)
        ref, got = a.fields(), c.fields()
    one = state_errors(got, ref)
    print(f"  split vs fp64 with mirror walls: one step {one}, 5 steps {traj}", flush=True)
    assert (ref["niter"] < 100).all() and (got["niter"] < 100).all()
    for k in STATE_ERRORS:
        assert one[k] <= 1e-5, (k, one[k])
        assert traj[k] <= 1e-5, (k, traj[k])
