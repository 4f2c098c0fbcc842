"""Global diagnostics and the .arr writer on the GPU, through the C ABI.

pond_diagnostics replaces calcTotMoments (update.cu:52-69: mass, x momentum, energy),
diagn (diagn.cu:3-22: k) and NiterMax (steps.cu:513); it is checked against the oracle's
totals of the same trajectory and against numpy sums of the fields the library returns.
pond_drop replaces drop() (drop.cu:2-34); its files are read back byte for byte.
"""
import os

import numpy as np
import pytest

import pyoracle as po
from helpers import smooth_random_state
from lbmpond_amd import Pond, read_arr
from lbmpond_amd import pond as pond_mod

pytestmark = pytest.mark.gpu


# (Dim, TLat, largest |e| component) per lattice
LAT = {"D1Q5": (1.0, 1.0, 2.8569700138728056), "D3Q27": (3.0, 1.0 / 3.0, 1.0), "D3Q125": (3.0, 1.0, 2.8569700138728056)}


def numpy_totals(fl, lattice, shape):
    dim, tlat, emax = LAT[lattice]
    rho, ux, uy, uz, T = (np.asarray(fl[k], np.float64) for k in ("rho", "ux", "uy", "uz", "T"))
    nx, ny, nz = shape
    d = (2 * np.pi / nx) * (2 * np.pi / ny) * (2 * np.pi / nz)
    u2 = ux * ux + uy * uy + uz * uz
    terms = {"mass": rho, "momentum_x": rho * ux, "energy": rho * (u2 + dim * T), "k": 0.5 / (8 * np.pi ** 3) * u2 * d}
    out = {k: float(v.sum()) for k, v in terms.items()}
    # the size of a sum's rounding is set by the sum of its terms' magnitudes (momentum_x cancels)
    out["scale"] = {k: float(np.abs(v).sum()) for k, v in terms.items()}
    out.update(niter_max=int(fl["niter"].max()), niter_mean=float(fl["niter"].mean()))
    out["max_departure"] = float((np.max([np.abs(ux), np.abs(uy), np.abs(uz)], axis=0) +
                                  np.sqrt(np.abs(T) / tlat) * emax).max())
    return out


def check_totals(d, ref, scale, tol=1e-12):
    for k in ("mass", "momentum_x", "energy", "k"):
        assert abs(d[k] - ref[k]) <= tol * scale[k], (k, d[k], ref[k])


@pytest.mark.parametrize("lattice,shape,interp", [("D3Q27", (16, 12, 10), "weno5c"), ("D1Q5", (200, 1, 1), "weno5"),
                                                  ("D3Q125", (5, 4, 3), "weno5")])
def test_fp64_diagnostics_match_oracle_totals(gpu, lattice, shape, interp):
    dims = 1 if lattice == "D1Q5" else 3
    f = smooth_random_state(shape, seed=91, T0=0.02, dims=dims)
    o = po.Oracle(lattice, shape, interp)
    o.set_state(*f)
    o.step(3)
    ot = o.totals()
    ofl = o.fields()
    with Pond(lattice, shape, interp, "fp64") as p:
        p.set_state(*f)
        p.step(3)
        d = p.diagnostics()
        fl = p.fields()
    print(f"[diag] {lattice} gpu {d} oracle {ot}", flush=True)
    ref = numpy_totals(fl, lattice, shape)
    check_totals(d, ot, ref["scale"])
    check_totals(d, ref, ref["scale"])
    assert d["max_departure"] == pytest.approx(ref["max_departure"], rel=1e-14)
    assert d["niter_max"] == int(ofl["niter"].max())
    assert d["niter_mean"] == pytest.approx(float(ofl["niter"].mean()), rel=1e-15)
    assert d["step"] == 3


@pytest.mark.parametrize("use_mlp", [False, True])
def test_split_diagnostics_match_fields(gpu, use_mlp):
    shape = (16, 12, 10)
    f = smooth_random_state(shape, seed=92, T0=0.01)
    with Pond("D3Q27", shape, "weno5c", "split", use_mlp=use_mlp) as p:
        if use_mlp:
            p.mlp_init(seed=1)
            params, scales = p.mlp_params()
            p.mlp_set_params(0.2 * params, scales)
        p.set_state(*f)
        d0 = p.diagnostics()
        assert d0["step"] == 0 and d0["niter_max"] == 0
        p.step(2)
        d = p.diagnostics()
        fl = p.fields()
    print(f"[diag split] {d}", flush=True)
    ref = numpy_totals(fl, "D3Q27", shape)
    check_totals(d, ref, ref["scale"])
    assert d["niter_max"] == ref["niter_max"]
    assert d["niter_mean"] == pytest.approx(ref["niter_mean"], rel=1e-15)
    assert d["max_departure"] == pytest.approx(ref["max_departure"], rel=1e-14)
    # conserved to the split path's rounding: BGK and the gauge transfer keep mass and momentum
    o = po.Oracle("D3Q27", shape, "weno5c")
    o.set_state(*f)
    o.step(2)
    ot = o.totals()
    assert abs(d["mass"] - ot["mass"]) <= 1e-6 * ot["mass"]


def test_diagnostics_tile_reduction_on_a_large_grid(gpu):
    """more cells than one pass of the reduction grid covers (1024 blocks x 256 lanes)"""
    shape = (128, 64, 48)
    f = smooth_random_state(shape, seed=93, T0=0.01)
    with Pond("D3Q27", shape, "weno5c", "split") as p:
        p.set_state(*f)
        p.step(1)
        d = p.diagnostics()
        fl = p.fields()
    ref = numpy_totals(fl, "D3Q27", shape)
    check_totals(d, ref, ref["scale"])
    assert d["niter_max"] == ref["niter_max"]


@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_drop_writes_the_reference_arr_files(gpu, tmp_path, precision):
    """drop(): Rho/Vx/Vy/Vz/T_%07d.arr (step number), header {comment length 0, dim 3,
    sizeof(double), Nx, Ny, Nz}, then the field z-outer / x-fastest -- equal to pond_get_fields
    bit for bit, and read back by pond_read_arr."""
    shape = (20, 9, 7)
    f = smooth_random_state(shape, seed=94, T0=0.01)
    with Pond("D3Q27", shape, "weno5c", precision) as p:
        p.set_state(*f)
        p.step(2)
        p.drop(tmp_path)
        fl = p.fields()
    names = {"Rho": "rho", "Vx": "ux", "Vy": "uy", "Vz": "uz", "T": "T"}
    assert sorted(os.listdir(tmp_path)) == sorted(f"{k}_0000002.arr" for k in names)
    for fname, key in names.items():
        path = tmp_path / f"{fname}_0000002.arr"
        raw = path.read_bytes()
        assert len(raw) == 24 + 8 * fl[key].size
        hdr = np.frombuffer(raw[:24], np.int32)
        assert hdr.tolist() == [0, 3, 8, *shape]
        body = np.frombuffer(raw[24:], np.float64)
        assert np.array_equal(body, fl[key]), fname
        back = read_arr(str(path))
        assert back.shape == (shape[2], shape[1], shape[0])
        assert np.array_equal(back.reshape(-1), fl[key])


def test_reference_named_facade(gpu, tmp_path):
    """init / calcStep / drop / diagnostics with the reference's names and LBMParams fields"""
    shape = (16, 8, 4)
    f = smooth_random_state(shape, seed=95, T0=0.01)
    pars = pond_mod.LBMParams(Nx=shape[0], Ny=shape[1], Nz=shape[2], lattice="D3Q27", interp="weno5c")
    pond_mod.init(pars, *f)
    for _ in range(3):
        pond_mod.calcStep()
    assert pars.iStep == 3
    pond_mod.drop(tmp_path)
    d = pond_mod.diagnostics()
    assert d["step"] == 3
    o = po.Oracle("D3Q27", shape, "weno5c")
    o.set_state(*f)
    o.step(3)
    ref = o.fields()
    assert np.array_equal(read_arr(str(tmp_path / "Rho_0000003.arr")).reshape(-1), ref["rho"])
    check_totals(d, o.totals(), numpy_totals(ref, "D3Q27", shape)["scale"])
    pond_mod._current["pond"].close()
    pond_mod._current["pond"] = None
