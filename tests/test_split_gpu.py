"""FP32/FP64-split step against the plain FP64 step (the FP64 oracle run: the plain path is
bit-identical to the CPU oracle, see test_parity_gpu.py) on the same seeded inputs.

The bar is <= 1e-5 relative L-inf on rho, u (as a vector) and T; every test prints the
measured error and the mean gauge iterations of both paths.
"""
import numpy as np
import pytest

from helpers import FIELDS, STATE_ERRORS, rel_linf, smooth_random_state, x_dominant_state
from lbmpond_amd import Pond
from lbmpond_amd.compare import CapDependence, state_errors
from lbmpond_amd import ic

pytestmark = pytest.mark.gpu


def run(lattice, shape, interp, precision, fields, steps, **kw):
    with Pond(lattice, shape, interp, precision, **kw) as p:
        p.set_state(*fields)
        iters = []
        dep = CapDependence(shape, p.params.max_iter)
        for _ in range(steps):
            p.step(1)
            iters.append(p.diagnostics()["niter_mean"])
            dep.update(p.niter())
        out = p.fields()
        out["niter_hist"] = np.array(iters)
        out["cap_dependence"] = dep
        return out


def compare(shape, interp, fields, steps, tol, dtau=1.0 / 0.9, lattice="D3Q27", max_dependent=0.1):
    """Split against FP64 after `steps` steps.  A cell that hit the iteration cap in either run
    ends its step on an unconverged iterate that depends on where its loop started, so the bar
    applies outside the cells such a cell can have influenced (lbmpond_amd.compare), which may
    cover at most `max_dependent` of the grid."""
    print(f"[split] {lattice} {interp} {shape} steps={steps}", flush=True)
    ref = run(lattice, shape, interp, "fp64", fields, steps, dtau=dtau)
    got = run(lattice, shape, interp, "split", fields, steps, dtau=dtau)
    dep = ref["cap_dependence"].mask | got["cap_dependence"].mask
    errs = state_errors(got, ref, ~dep if dep.any() else None)
    print("  split vs fp64 rel Linf:", {k: f"{v:.2e}" for k, v in errs.items()},
          f" mean iters fp64 {ref['niter_hist'].mean():.2f} split {got['niter_hist'].mean():.2f}", flush=True)
    if dep.any():
        print(f"  capped cells per step: fp64 {ref['cap_dependence'].capped} split {got['cap_dependence'].capped};"
              f" {dep.mean():.3f} of the cells excluded", flush=True)
    assert dep.mean() <= max_dependent, dep.mean()
    for k in STATE_ERRORS:
        assert errs[k] <= tol, (k, errs[k])
    return errs, ref, got


def test_split_weno5c_random(gpu):
    shape = (16, 12, 10)
    compare(shape, "weno5c", smooth_random_state(shape, seed=61, T0=0.01), 20, 1e-5)


def test_split_generic_weno5_runs_the_as_written_kernel(gpu):
    """POND_WENO5G in the split path is POND_WENO5's kernel (the same interpolant; the split
    arithmetic does not keep the FP64 operation order anyway): bitwise equal results"""
    shape = (16, 8, 4)
    f = x_dominant_state(shape, seed=62, T0=0.01)
    a = run("D3Q27", shape, "weno5g", "split", f, 3)
    b = run("D3Q27", shape, "weno5", "split", f, 3)
    for k in FIELDS:
        assert np.array_equal(a[k], b[k]), k


def test_split_lagr4_random(gpu):
    shape = (16, 12, 10)
    compare(shape, "lagr4", smooth_random_state(shape, seed=62, T0=0.02), 20, 1e-5)


def test_split_taylor_green_small(gpu):
    """20 steps: in step 18 a few cells reach the iteration cap in both runs; the cells they can
    have influenced by step 20 (about a fifth of this small grid) are left out."""
    case = ic.taylor_green(n=32, steps=20)
    compare(case.shape, "weno5c", case.fields, 20, 1e-5, dtau=case.dtau, max_dependent=0.3)


def test_split_taylor_green_error_tracks_fp64_sensitivity(gpu):
    """Re 1600 Taylor-Green (dtau near 2, little dissipation) over 17 steps, the last step
    before the FP64 run first leaves a cell at the iteration cap on this grid: bounded like the
    vortex case by the FP64 path's own response to a 1e-7 density perturbation."""
    case = ic.taylor_green(n=32, steps=17)
    sensitivity_bound(case, 17)


def test_split_vortex_small(gpu):
    case = ic.vortex_3d(n=32)
    compare(case.shape, "weno5c", case.fields, 10, 1e-5)


def test_split_vortex_error_tracks_fp64_sensitivity(gpu):
    """Over 30 steps the vortex case amplifies any perturbation about a hundredfold: its seeded
    1e-3 density noise puts the WENO smoothness indicators near their 1e-6 regularisation, and
    the nonlinear weights respond to that."""
    sensitivity_bound(ic.vortex_3d(n=32), 30)


def sensitivity_bound(case, steps):
    """The split path's FP32 rounding (~1e-7 per step) enters at every step, and what enters at
    step k is amplified over the remaining steps at most as much as a perturbation of the
    initial state, so the split error after n steps is bounded by n times the FP64 path's own
    response to a 1e-7 perturbation of the initial density (with a factor 2 for the noise
    floor, up to 2e-7)."""
    _, ref, got = compare(case.shape, "weno5c", case.fields, steps, np.inf, dtau=case.dtau, max_dependent=0.0)
    rho, ux, uy, uz, T = case.fields
    rng = np.random.default_rng(17)
    pert = run("D3Q27", case.shape, "weno5c", "fp64", (rho * (1 + 1e-7 * rng.uniform(-1, 1, rho.shape)), ux, uy, uz, T),
               steps, dtau=case.dtau)
    e_split = state_errors(got, ref)
    e_pert = state_errors(pert, ref)
    print("  fp64 response to a 1e-7 density perturbation:", {k: f"{v:.2e}" for k, v in e_pert.items()})
    for k in STATE_ERRORS:
        assert e_split[k] <= 2 * steps * max(e_pert[k], 1e-7), (k, e_split[k], e_pert[k])


def test_split_fewer_or_equal_iterations(gpu):
    shape = (16, 16, 16)
    f = smooth_random_state(shape, seed=63, T0=0.01)
    _, ref, got = compare(shape, "weno5c", f, 10, 1e-5)
    assert got["niter_hist"].mean() <= ref["niter_hist"].mean()


def test_split_state_roundtrip(gpu):
    shape = (8, 4, 4)
    rho, ux, uy, uz, T = smooth_random_state(shape, seed=64, T0=0.02)
    with Pond("D3Q27", shape, "weno5c", "split") as p:
        p.set_state(rho, ux, uy, uz, T)
        out = p.fields()
        f = p.populations()
    assert rel_linf(out["rho"], rho) < 1e-7 and rel_linf(out["ux"], ux) < 1e-7
    assert np.array_equal(out["T"], T)  # T is kept in FP64
    w = np.array([8 / 27] + [0] * 26)
    assert abs(f[0] - (w[0] * rho).astype(np.float32)).max() < 1e-6


def test_split_d3q125_random(gpu):
    """D3Q125 split step (block per cell, fixed 3x3x3 stencil) against the FP64 D3Q125 step."""
    shape = (6, 5, 4)
    compare(shape, "lagr4", smooth_random_state(shape, seed=65, T0=0.05, u_amp=0.05), 10, 1e-5, lattice="D3Q125")


def test_split_d3q125_turbulence(gpu):
    case = ic.isotropic_turbulence(n=12)
    errs, ref, got = compare(case.shape, "lagr4", case.fields, 10, 1e-5, dtau=case.dtau, lattice="D3Q125")
    assert got["niter_hist"].mean() <= ref["niter_hist"].mean()


def test_two_pass_split_is_independent_of_first_pass_length(gpu):
    """Cells left unconverged by the first pass resume in the second with the same state: the
    result is bitwise the same for any first-pass length (pond_params.pass1_iters)."""
    case = ic.shock_bubble(n=24)
    outs = []
    for k in (1, 3, 100):
        outs.append(run("D3Q27", case.shape, "weno5c", "split", case.fields, 3, dtau=case.dtau, pass1_iters=k))
    for o in outs[:2]:
        for key in ("rho", "ux", "uy", "uz", "T", "niter"):
            assert np.array_equal(o[key], outs[2][key]), key


def test_refill_matches_one_tile_per_wave(gpu):
    """The lane-refill single pass (a wave owns several 64-cell tiles and its lanes take new cells
    as theirs converge) gives bitwise the results of one tile per wave, on a shock-bubble state
    and on a grid whose rows are shorter than a wave and whose last tile is ragged.  The WENO5C
    waves are persistent and claim chunks from eight per-XCD slabs: the small grids give fewer
    chunks than slabs (8 tiles: chunks of 3 make 3 chunks, one wave each) and slabs of unequal
    length (9 tiles in chunks of 1: slab 0 holds two chunks, so its wave claims twice)."""
    cases = ((ic.shock_bubble(n=40).fields, (40, 40, 40), 4),
             (smooth_random_state((37, 20, 11), seed=71, T0=0.01), (37, 20, 11), 3),
             (smooth_random_state((16, 8, 4), seed=72, T0=0.01), (16, 8, 4), 3),
             (smooth_random_state((24, 8, 3), seed=73, T0=0.01), (24, 8, 3), 3))
    for fields, shape, steps in cases:
        outs = []
        for chunk in (0, 1, 3, 8):
            print(f"[refill] {shape} chunk {chunk}", flush=True)
            with Pond("D3Q27", shape, "weno5c", "split", refill_chunk=chunk) as p:
                p.set_state(*fields)
                p.step(steps)
                outs.append(p.fields())
        for o in outs[1:]:
            for k in ("rho", "ux", "uy", "uz", "T", "niter"):
                assert np.array_equal(o[k], outs[0][k]), (shape, k)


def test_split_absurd_gauges_terminate(gpu):
    """NaN, huge and infinite gauges in a few cells: the grouped WENO5C step bounds its x window
    (falling back to target-by-target interpolation when the three departure windows lie far
    apart), so the step ends promptly; cells away from the absurd ones match the FP64 step."""
    shape = (32, 12, 10)
    nx, ny, nz = shape
    rho, ux, uy, uz, T = smooth_random_state(shape, seed=74, T0=0.02)
    def cell(x, y, z):
        return x + nx * (y + ny * z)
    T[cell(1, 2, 3)] = np.nan
    ux[cell(2, 5, 5)] = 1e300
    uy[cell(1, 8, 2)] = -1e12
    T[cell(2, 3, 7)] = 1e12
    T[cell(3, 9, 1)] = np.inf
    ref = run("D3Q27", shape, "weno5c", "fp64", (rho, ux, uy, uz, T), 2)
    got = run("D3Q27", shape, "weno5c", "split", (rho, ux, uy, uz, T), 2)
    x = np.arange(nx * ny * nz) % nx
    far = (x >= 14) & (x <= 20)
    assert np.isfinite(ref["rho"][far]).all()
    # normalised over the far cells too (the absurd values would otherwise set the scale)
    fields = ("rho", "ux", "uy", "uz", "T")
    errs = state_errors({k: got[k][far] for k in fields}, {k: ref[k][far] for k in fields})
    print("  split vs fp64 away from the absurd cells:", {k: f"{v:.2e}" for k, v in errs.items()})
    for k in STATE_ERRORS:
        assert errs[k] <= 1e-5, (k, errs[k])


@pytest.mark.parametrize("lattice,interp,shape", [("D3Q27", "weno5c", (32, 12, 10)), ("D3Q125", "lagr4", (12, 10, 8))])
def test_split_nonfinite_nodes_match_fp64(gpu, lattice, interp, shape):
    """Non-finite data in a few cells (NaN or infinite temperature, density or velocity, and a
    velocity beyond FP32 range).  The reference transfer drops non-finite terms
    (structs.cuh:542), which confines such a cell to itself: its neighbours interpolate it as a
    node that adds only finite terms.  The split step must confine it the same way: outside the
    influence of iteration-capped cells the non-finite cells coincide with the FP64 step's, and
    the finite ones match it to 1e-5."""
    nx, ny, nz = shape
    rho, ux, uy, uz, T = smooth_random_state(shape, seed=75, T0=0.02)
    def cell(x, y, z):
        return x + nx * (y + ny * z)
    T[cell(1, 2, 3)] = np.nan
    ux[cell(2, 5, 5)] = 1e300
    T[cell(3, 9 % ny, 1)] = np.inf
    uz[cell(nx // 2, 4, 4)] = -np.inf
    rho[cell(nx - 5, 7, nz - 2)] = np.nan
    fields = (rho, ux, uy, uz, T)
    steps = 2
    ref = run(lattice, shape, interp, "fp64", fields, steps)
    got = run(lattice, shape, interp, "split", fields, steps)
    dep = ref["cap_dependence"].mask | got["cap_dependence"].mask
    bad_ref, bad_got = ~np.isfinite(ref["rho"]), ~np.isfinite(got["rho"])
    print(f"  non-finite cells fp64 {np.flatnonzero(bad_ref)} split {np.flatnonzero(bad_got)};"
          f" {dep.mean():.3f} of the cells cap-dependent", flush=True)
    assert bad_ref.sum() <= 5, "the reference confines each non-finite cell"
    assert np.array_equal(bad_ref & ~dep, bad_got & ~dep)
    good = ~dep & ~bad_ref
    assert good.mean() > 0.5
    keys = ("rho", "ux", "uy", "uz", "T")
    errs = state_errors({k: got[k][good] for k in keys}, {k: ref[k][good] for k in keys})
    print("  split vs fp64 on the finite cells:", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k in STATE_ERRORS:
        assert errs[k] <= 1e-5, (k, errs[k])


@pytest.mark.parametrize("learned", [False, True])
def test_split_row_patch_order_is_bitwise_neutral(gpu, learned):
    """The split step's row-patch order with whole-row refill chunks (the default wherever
    nx % 64 == 0 and 4 divides ny and nz: every 256^3 and 384^3 run) gives the same bits as plain
    row order with one tile per wave, with the plain and the learned start."""
    shape = (128, 8, 8)
    f = smooth_random_state(shape, seed=82, T0=0.01)
    outs = []
    for kw in ({}, {"tile_patch": 0, "refill_chunk": 0}, {"tile_patch": 0}):
        print(f"[row patch split] learned={learned} {kw}", flush=True)
        with Pond("D3Q27", shape, "weno5c", "split", use_mlp=learned, **kw) as p:
            if learned:
                p.mlp_init(seed=5)
                params, scales = p.mlp_params()
                p.mlp_set_params(0.2 * params, scales)
            p.set_state(*f)
            p.step(3)
            outs.append(p.fields())
    for o in outs[1:]:
        for k in ("rho", "ux", "uy", "uz", "T", "niter"):
            assert np.array_equal(o[k], outs[0][k]), k


def test_split_against_oracle(gpu):
    """The split step against the CPU oracle itself (not through the FP64 GPU path): 10 steps
    of a smooth random state, relative L-inf <= 1e-5 on rho, u and T."""
    import pyoracle as po
    shape = (16, 12, 10)
    f = smooth_random_state(shape, seed=65, T0=0.01)
    o = po.Oracle("D3Q27", shape, "weno5c")
    o.set_state(*f)
    o.step(10)
    ref = o.fields()
    with Pond("D3Q27", shape, "weno5c", "split") as p:
        p.set_state(*f)
        p.step(10)
        got = p.fields()
    errs = state_errors(got, ref)
    print("  split vs oracle rel Linf:", {k: f"{v:.2e}" for k, v in errs.items()},
          f" mean iters oracle {ref['niter'].mean():.2f} split {got['niter'].mean():.2f}", flush=True)
    assert (ref["niter"] < 100).all() and (got["niter"] < 100).all()
    for k in STATE_ERRORS:
        assert errs[k] <= 1e-5, (k, errs[k])


def test_copy_state_between_precisions(gpu):
    """pond_copy_state: an FP64 state copied into a split context steps exactly like the same
    state uploaded with pond_set_populations (FP64 -> FP32 rounding to nearest either way), and
    an FP64 -> FP64 copy continues bit for bit like the source"""
    shape = (16, 12, 10)
    f = smooth_random_state(shape, seed=66, T0=0.01)
    with Pond("D3Q27", shape, "weno5c", "fp64") as a:
        a.set_state(*f)
        a.step(2)
        pop = a.populations()
        fl = a.fields()
        with Pond("D3Q27", shape, "weno5c", "split") as b, Pond("D3Q27", shape, "weno5c", "split") as c, \
                Pond("D3Q27", shape, "weno5c", "fp64") as d:
            b.copy_state_from(a)
            c.set_populations(pop, fl["rho"], fl["ux"], fl["uy"], fl["uz"], fl["T"])
            d.copy_state_from(a)
            assert b.diagnostics()["step"] == 2
            for p in (a, b, c, d):
                p.step(2)
            gb, gc, ga, gd = b.fields(), c.fields(), a.fields(), d.fields()
    for k in ("rho", "ux", "uy", "uz", "T", "niter"):
        assert np.array_equal(gb[k], gc[k]), k
        assert np.array_equal(ga[k], gd[k]), k
