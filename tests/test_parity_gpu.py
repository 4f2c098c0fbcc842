"""GPU parity of the plain FP64 step against the CPU oracle, through the C ABI.

The bar: relative L-inf <= 1e-11 on rho, u, T after 100 steps on the same seeded
inputs; the plain path is built to reproduce the reference's rounding, so the tests also
require identical per-cell iteration counts.
"""
import numpy as np
import pytest

import pyoracle as po
from helpers import FIELDS, rel_linf, smooth_random_state, x_dominant_state
from lbmpond_amd import Pond
from lbmpond_amd import ic

pytestmark = pytest.mark.gpu

TOL = 1e-11


def run_pair(lattice, shape, interp, fields, steps, **kw):
    print(f"[parity] {lattice} {interp} shape={shape} steps={steps} kw={kw}", flush=True)
    o = po.Oracle(lattice, shape, interp, **kw)
    o.set_state(*fields)
    o.step(steps)
    ref = o.fields()
    with Pond(lattice, shape, interp, "fp64", **kw) as p:
        p.set_state(*fields)
        p.step(steps)
        got = p.fields()
        fpop = p.populations()
    return ref, got, o.populations(), fpop


def check(ref, got, fref=None, fgot=None, tol=TOL, exact_niter=True):
    errs = {k: rel_linf(got[k], ref[k]) for k in FIELDS}
    print("  rel Linf:", {k: f"{v:.3e}" for k, v in errs.items()},
          " bitwise-equal cells:", int(np.sum(got["rho"] == ref["rho"])), "/", ref["rho"].size,
          " niter mean:", float(ref["niter"].mean()), flush=True)
    for k, v in errs.items():
        assert v <= tol, (k, v)
    if exact_niter:
        assert np.array_equal(got["niter"], ref["niter"])
    if fref is not None:
        assert rel_linf(fgot, fref) <= tol


def test_d1q5_sod_weno5_100_steps(gpu):
    case = ic.sod_1d(2000)
    ref, got, fr, fg = run_pair("D1Q5", case.shape, "weno5", case.fields, 100, dtau=case.dtau)
    check(ref, got, fr, fg)


def test_d1q5_random_lagr4(gpu):
    f = smooth_random_state((96, 1, 1), seed=11, T0=0.03, dims=1)
    ref, got, fr, fg = run_pair("D1Q5", (96, 1, 1), "lagr4", f, 100)
    check(ref, got, fr, fg)


def test_d3q27_weno5c_100_steps(gpu):
    shape = (8, 6, 5)
    f = smooth_random_state(shape, seed=21, T0=0.01)
    ref, got, fr, fg = run_pair("D3Q27", shape, "weno5c", f, 100)
    assert np.isfinite(ref["rho"]).all()
    check(ref, got, fr, fg)


def test_d3q27_lagr4_100_steps(gpu):
    shape = (8, 6, 5)
    f = smooth_random_state(shape, seed=22, T0=0.02)
    ref, got, fr, fg = run_pair("D3Q27", shape, "lagr4", f, 100)
    assert np.isfinite(ref["rho"]).all()
    check(ref, got, fr, fg)


def test_d3q27_weno5_as_written_15_steps(gpu):
    """propagateWENO<3,5> as written grows short y/z wavelengths every step (its y/z stencil
    extrapolates downwind); from an x-dominant state the trajectory stays finite for 15 steps
    and is compared bit for bit."""
    shape = (8, 6, 5)
    f = x_dominant_state(shape, seed=24, T0=0.01)
    ref, got, fr, fg = run_pair("D3Q27", shape, "weno5", f, 15)
    assert np.isfinite(ref["rho"]).all()
    check(ref, got, fr, fg)


def test_d1q5_sod_generic_weno5_100_steps(gpu):
    """propagateWENO<5,M> (the generic template, sub-stencils numbered from the right) on the
    1D Sod tube, bit for bit against the oracle; its rounding differs from propagateWENO<3,5>
    (the two are compared as well, to show that the kernel really is the other operation
    order: equal to 1e-9, not bitwise)."""
    case = ic.sod_1d(2000)
    ref, got, fr, fg = run_pair("D1Q5", case.shape, "weno5g", case.fields, 100, dtau=case.dtau)
    check(ref, got, fr, fg)
    with Pond("D1Q5", case.shape, "weno5", "fp64", dtau=case.dtau) as p:
        p.set_state(*case.fields)
        p.step(100)
        other = p.fields()
    assert rel_linf(got["rho"], other["rho"]) < 1e-9
    assert not np.array_equal(got["rho"], other["rho"])


def test_d3q27_generic_weno5_15_steps(gpu):
    """the generic WENO5 in 3D, from an x-dominant state (its y/z stencil extrapolates downwind
    like propagateWENO<3,5>'s), bit for bit against the oracle, on the production tile order
    (64 x 8 x 4: row patches and whole-row lane refill) and on a ragged grid"""
    for shape, seed in (((8, 6, 5), 25), ((64, 8, 4), 26)):
        f = x_dominant_state(shape, seed=seed, T0=0.01)
        ref, got, fr, fg = run_pair("D3Q27", shape, "weno5g", f, 15)
        assert np.isfinite(ref["rho"]).all()
        check(ref, got, fr, fg)


def test_d3q27_weno5c_ragged_grid(gpu):
    # N not a multiple of the 64-lane wave; rows shorter than a wave; Ny != Nz
    shape = (13, 9, 7)
    f = smooth_random_state(shape, seed=23, T0=0.015)
    ref, got, fr, fg = run_pair("D3Q27", shape, "weno5c", f, 5)
    check(ref, got, fr, fg)


def test_d3q27_vortex_ic(gpu):
    case = ic.vortex_3d(n=12, seed=2)
    ref, got, fr, fg = run_pair("D3Q27", case.shape, "weno5c", case.fields, 10)
    check(ref, got, fr, fg)


def test_d3q125_random(gpu):
    shape = (4, 4, 4)
    f = smooth_random_state(shape, seed=31, T0=0.03)
    ref, got, fr, fg = run_pair("D3Q125", shape, "weno5", f, 8)
    check(ref, got, fr, fg)


def test_d3q125_ragged(gpu):
    shape = (5, 4, 3)
    f = smooth_random_state(shape, seed=32, T0=0.02)
    ref, got, fr, fg = run_pair("D3Q125", shape, "weno5", f, 4)
    check(ref, got, fr, fg)


def test_edge_zero_cells_and_cap(gpu):
    """all-zero cells are reset to (rho=1,u=0,T=TLat) (steps.cu:363); tol=0 forces the
    iteration cap (steps.cu:422)."""
    shape = (10, 4, 3)
    rho, ux, uy, uz, T = smooth_random_state(shape, seed=41, T0=0.02)
    for a in (rho, ux, uy, uz, T):
        a[[3, 17, 50]] = 0.0
    ref, got, fr, fg = run_pair("D3Q27", shape, "weno5c", (rho, ux, uy, uz, T), 3, tol_abs=0.0, tol_rel=0.0,
                                max_iter=4)
    assert (ref["niter"] == 4).all()
    check(ref, got, fr, fg)


def test_edge_nan_neighbour_and_hot_spot(gpu):
    """a NaN population in one cell: NaN transfer terms are dropped (structs.cuh:542); a
    large temperature jump exercises the T<0 clamp path (steps.cu:475-478)."""
    shape = (12, 5, 4)
    rho, ux, uy, uz, T = smooth_random_state(shape, seed=42, T0=0.02)
    T[7] = 0.6
    o_fields = (rho, ux, uy, uz, T)
    o = po.Oracle("D3Q27", shape, "weno5c")
    o.set_state(*o_fields)
    f = o.populations()
    f[5, 30] = np.nan
    print("[parity] nan neighbour", flush=True)
    o.set_populations(f, *o_fields)
    o.step(3)
    ref = o.fields()
    with Pond("D3Q27", shape, "weno5c", "fp64") as p:
        p.set_populations(f, *o_fields)
        p.step(3)
        got = p.fields()
    check(ref, got)


@pytest.mark.parametrize("interp", ["weno5c", "weno5", "lagr4"])
def test_hot_cells_take_the_one_target_path(gpu, interp):
    """Cells whose thermal speed sqrt(T/TLat) exceeds 1.5 (here T = 1.2 and 0.9 in a T = 0.02
    state) have their three ex targets' windows too far apart for the grouped D3Q27 step and
    take the one-target path, inside a grid whose other cells take the grouped one: bit for bit
    against the oracle, on a row-patch grid and on a ragged one."""
    for shape, seed in (((64, 4, 4), 43), ((13, 6, 5), 44)):
        rho, ux, uy, uz, T = x_dominant_state(shape, seed=seed, T0=0.02) if interp == "weno5" else \
            smooth_random_state(shape, seed=seed, T0=0.02)
        T = T.copy()
        T[[9, 40, 41]] = [1.2, 0.9, 1.2]
        print(f"[parity] hot cells {interp} {shape}", flush=True)
        o = po.Oracle("D3Q27", shape, interp)
        o.set_state(rho, ux, uy, uz, T)
        o.step(3)
        ref = o.fields()
        with Pond("D3Q27", shape, interp, "fp64") as p:
            p.set_state(rho, ux, uy, uz, T)
            p.step(3)
            got = p.fields()
        for k in FIELDS:
            assert np.array_equal(np.isnan(got[k]), np.isnan(ref[k])), k
            fin = np.isfinite(ref[k])
            assert np.array_equal(got[k][fin], ref[k][fin]), k
        assert np.array_equal(got["niter"], ref["niter"])


def test_run_twice_bitwise(gpu):
    shape = (16, 8, 4)
    f = smooth_random_state(shape, seed=51, T0=0.01)
    outs = []
    for _ in range(2):
        with Pond("D3Q27", shape, "weno5c", "fp64") as p:
            p.set_state(*f)
            p.step(5)
            outs.append(p.fields())
    for k in FIELDS + ("niter",):
        assert np.array_equal(outs[0][k], outs[1][k])


def test_nonfinite_and_absurd_gauges_stay_in_bounds(gpu):
    """Regression: departure points of NaN / huge gauges used to be converted to int without a
    check (undefined), giving arbitrary stencil addresses.  They now map through floor_int
    (NaN -> 0, clamp to +-2^30 cells) in the kernels and the oracle alike, so the step stays in
    bounds and the result (including where NaNs appear) matches the oracle."""
    shape = (12, 6, 5)
    rho, ux, uy, uz, T = smooth_random_state(shape, seed=71, T0=0.02)
    T[5] = np.nan
    ux[40] = 1e300
    uy[77] = -1e12
    T[101] = np.inf
    for interp in ("weno5c", "lagr4"):
        print(f"[parity] non-finite gauges {interp}", flush=True)
        o = po.Oracle("D3Q27", shape, interp, max_iter=5)
        o.set_state(rho, ux, uy, uz, T)
        o.step(2)
        ref = o.fields()
        with Pond("D3Q27", shape, interp, "fp64", max_iter=5) as p:
            p.set_state(rho, ux, uy, uz, T)
            p.step(2)
            got = p.fields()
        for k in FIELDS:
            assert np.array_equal(np.isnan(got[k]), np.isnan(ref[k])), k
            fin = np.isfinite(ref[k])
            assert np.array_equal(got[k][fin], ref[k][fin]), k


def test_two_pass_fp64_is_independent_of_first_pass_length(gpu):
    """The FP64 step's second pass resumes cells bit for bit: any first-pass length gives the
    oracle's result."""
    shape = (16, 12, 10)
    f = smooth_random_state(shape, seed=71, T0=0.01)
    o = po.Oracle("D3Q27", shape, "weno5c")
    o.set_state(*f)
    o.step(3)
    ref = o.fields()
    for k in (1, 2):
        with Pond("D3Q27", shape, "weno5c", "fp64", pass1_iters=k) as p:
            p.set_state(*f)
            p.step(3)
            got = p.fields()
        for key in ("rho", "ux", "uy", "uz", "T", "niter"):
            assert np.array_equal(got[key], ref[key]), (k, key)


@pytest.mark.parametrize("interp", ["weno5", "lagr4", "weno5c"])
def test_fp64_lane_refill_matches_oracle(gpu, interp):
    """The FP64 first pass with lane refill (a wave owns several 64-cell tiles and its lanes take
    new cells as theirs finish), with and without a second pass, gives the oracle's result bit for
    bit, on a grid whose last tile is ragged; so does the two-pass step with the node loads
    through 64-bit global addresses (global_loads)."""
    shape = (37, 9, 7)
    f = smooth_random_state(shape, seed=72, T0=0.01)
    o = po.Oracle("D3Q27", shape, interp)
    o.set_state(*f)
    o.step(3)
    ref = o.fields()
    for chunk, k1, gl in ((1, 8, 0), (3, 2, 0), (8, 1000, 0), (3, 2, 1)):
        print(f"[refill f64] {interp} chunk {chunk} first pass {k1} global loads {gl}", flush=True)
        with Pond("D3Q27", shape, interp, "fp64", refill_chunk=chunk, pass1_iters=k1, global_loads=gl) as p:
            p.set_state(*f)
            p.step(3)
            got = p.fields()
        for key in ("rho", "ux", "uy", "uz", "T", "niter"):
            assert np.array_equal(got[key], ref[key]), (interp, chunk, k1, gl, key)


@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_checkpoint_restart_is_exact(gpu, tmp_path, precision):
    """A run restored from a checkpoint continues bit for bit like the uninterrupted run (state,
    iteration counts and step count); a checkpoint of another grid is refused."""
    shape = (20, 9, 7)
    f = smooth_random_state(shape, seed=73, T0=0.01)
    path = tmp_path / "state.ckpt"
    with Pond("D3Q27", shape, "weno5c", precision) as p:
        p.set_state(*f)
        p.step(3)
        p.checkpoint(path)
        p.step(2)
        ref = p.fields()
        ref_pop = p.populations()
        ref_step = p.diagnostics()["step"]
    with Pond("D3Q27", shape, "weno5c", precision) as q:
        q.restore(path)
        q.step(2)
        got = q.fields()
        assert np.array_equal(q.populations(), ref_pop)
        assert q.diagnostics()["step"] == ref_step
    for k in ("rho", "ux", "uy", "uz", "T", "niter"):
        assert np.array_equal(got[k], ref[k]), k
    with Pond("D3Q27", (20, 9, 8), "weno5c", precision) as r:
        with pytest.raises(Exception):
            r.restore(path)


# ---- the work distribution the large grids run: 4x4 (y, z) row patches of 64-cell row segments
# (used where nx % 64 == 0 and 4 divides ny and nz) with the default lane-refill chunks


def _row_patch_grid(shape):
    nx, ny, nz = shape
    return nx % 64 == 0 and ny % 4 == 0 and nz % 4 == 0


@pytest.mark.parametrize("shape,interp,steps", [((64, 8, 8), "weno5c", 3), ((128, 8, 4), "lagr4", 2)])
def test_fp64_row_patch_order_matches_oracle(gpu, shape, interp, steps):
    """FP64 on a grid that takes the row-patch order and the default refill chunks and first-pass
    length (the configuration of every 256^3 and 384^3 run) equals the oracle bit for bit, and so
    do plain row order with one tile per wave and the node loads through 64-bit global addresses
    (global_loads, the path of grids whose planes do not fit a 32-bit buffer offset)."""
    assert _row_patch_grid(shape)
    f = smooth_random_state(shape, seed=81, T0=0.01)
    o = po.Oracle("D3Q27", shape, interp)
    o.set_state(*f)
    o.step(steps)
    ref = o.fields()
    for kw in ({}, {"tile_patch": 0, "refill_chunk": 0}, {"global_loads": 1}):
        print(f"[row patch f64] {shape} {interp} {kw}", flush=True)
        with Pond("D3Q27", shape, interp, "fp64", **kw) as p:
            p.set_state(*f)
            p.step(steps)
            got = p.fields()
        for key in FIELDS + ("niter",):
            assert np.array_equal(got[key], ref[key]), (kw, key)


def test_d3q125_100_steps(gpu):
    """D3Q125 (PonDstepD3Q125, steps.cu:151-250) against the oracle for 100 steps."""
    shape = (4, 4, 4)
    f = smooth_random_state(shape, seed=33, T0=0.02)
    ref, got, fr, fg = run_pair("D3Q125", shape, "weno5", f, 100)
    assert np.isfinite(ref["rho"]).all()
    check(ref, got, fr, fg)
