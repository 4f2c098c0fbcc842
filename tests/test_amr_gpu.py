"""Switched-off cells and the 1D cell activation of amr.cu on the GPU, through the C ABI,
against the oracle.

A cell whose Cinfo::set is 0 is not stepped (LBMpull, steps.cu:359): its records stay as they
were and it remains an interpolation node.  With pond_params.amr the library runs AMR_add and
AMR_remove (amr.cu:1-55) after every step, as calcStep would with the launches of
update.cu:102-105 enabled.  Everything here is integer flags and FP64 records moved or averaged,
so the bar is bitwise: fields, populations, iteration counts and the flags.
"""
import os

import numpy as np
import pytest

import pyoracle as po
from helpers import FIELDS, smooth_random_state
from lbmpond_amd import Pond
from lbmpond_amd import ic

pytestmark = pytest.mark.gpu


def assert_same(o, p, label):
    ref, got = o.fields(), p.fields()
    for k in FIELDS:
        assert np.array_equal(ref[k], got[k]), (label, k)
    assert np.array_equal(ref["niter"], got["niter"]), label
    assert np.array_equal(o.populations(), p.populations()), label
    assert np.array_equal(o.active(), p.active()), label


@pytest.mark.parametrize("interp", ["lagr4", "weno5", "weno5g"])
def test_switched_off_cells_match_oracle(gpu, interp):
    n = 96
    f = smooth_random_state((n, 1, 1), seed=41, T0=0.03, dims=1)
    act = np.ones(n, np.int32)
    act[[0, 7, 8, 9, 50, 95]] = 0
    print(f"[amr] switched-off cells {interp}", flush=True)
    o = po.Oracle("D1Q5", (n, 1, 1), interp)
    o.set_state(*f)
    o.set_active(act)
    with Pond("D1Q5", (n, 1, 1), interp, "fp64") as p:
        p.set_state(*f)
        p.set_active(act)
        for s in range(4):
            o.step(5)
            p.step(5)
            assert_same(o, p, f"step {5 * (s + 1)}")


def sod_pair(steps, n=2000, interp="lagr4"):
    case = ic.sod_1d(n)
    o = po.Oracle("D1Q5", case.shape, interp, dtau=case.dtau, amr=1)
    o.set_state(*case.fields)
    p = Pond("D1Q5", case.shape, interp, "fp64", dtau=case.dtau, amr=1)
    p.set_state(*case.fields)
    return case, o, p


@pytest.mark.parametrize("interp", ["lagr4", "weno5"])
def test_amr_sod_tube_matches_oracle(gpu, interp):
    """The 2000-cell Sod tube with cell activation, 120 steps, bit for bit: the uniform states
    thin out to every fourth cell in the first step, the cells around the waves stay on and cells
    are switched back on in the gaps the waves cross"""
    case, o, p = sod_pair(0, interp=interp)
    counts, added = [], 0
    prev = np.ones(case.shape[0], np.int32)
    with p:
        for s in range(12):
            print(f"[amr] sod {interp} steps {10 * s}..{10 * s + 10}", flush=True)
            for _ in range(10):
                o.step(1)
                p.step(1)
                a = o.active()
                added += int(((a == 1) & (prev == 0)).sum())
                prev = a
            assert_same(o, p, f"step {10 * (s + 1)}")
            counts.append(int(prev.sum()))
    print("  active cells every 10 steps:", counts, " cells switched back on:", added, flush=True)
    assert counts[0] < 0.4 * case.shape[0]
    assert added > 0


def test_amr_checkpoint_restarts_bit_for_bit(gpu, tmp_path):
    """the checkpoint of a context with switched-off cells holds both buffers and the flags"""
    case, o, p = sod_pair(0)
    path = os.path.join(str(tmp_path), "amr.ckpt")
    with p:
        p.step(15)
        act15 = p.active()
        p.checkpoint(path)
        p.step(10)
        ref = p.fields()
        ref_act = p.active()
        ref_f = p.populations()
        with Pond("D1Q5", case.shape, "lagr4", "fp64", dtau=case.dtau, amr=1) as q:
            q.restore(path)
            assert np.array_equal(q.active(), act15)
            q.step(10)
            got = q.fields()
            assert np.array_equal(q.active(), ref_act)
            assert np.array_equal(q.populations(), ref_f)
    assert (act15 == 0).any()
    for k in FIELDS:
        assert np.array_equal(ref[k], got[k]), k


def test_diagnostics_cover_the_active_cells(gpu):
    """calcTotMoments sums the active cells (update.cu:58), diagn's k every cell (diagn.cu:8-18),
    the iteration figures the stepped cells"""
    case, o, p = sod_pair(0)
    o.step(20)
    with p:
        p.step(20)
        fl = p.fields()
        act = p.active().astype(bool)
        d = p.diagnostics()
    assert not act.all()
    tot = o.totals()
    for k in ("mass", "momentum_x", "energy", "k"):
        assert abs(d[k] - tot[k]) <= 1e-12 * max(abs(tot[k]), 1e-300), (k, d[k], tot[k])
    u2 = fl["ux"] ** 2 + fl["uy"] ** 2 + fl["uz"] ** 2
    ref = {"mass": fl["rho"][act].sum(), "momentum_x": (fl["rho"] * fl["ux"])[act].sum(),
           "energy": (fl["rho"] * (u2 + fl["T"]))[act].sum(), "k": (0.5 / (8 * np.pi ** 3) * u2 * (2 * np.pi) ** 3 / act.size).sum()}
    for k, v in ref.items():
        assert abs(d[k] - v) <= 1e-12 * max(abs(v), 1e-300), (k, d[k], v)
    assert d["niter_max"] == int(fl["niter"][act].max())
    assert abs(d["niter_mean"] - fl["niter"][act].mean()) < 1e-12


def test_set_active_is_refused_where_cells_cannot_be_switched_off(gpu):
    with Pond("D3Q27", (8, 4, 4), "lagr4", "fp64") as p:
        assert (p.active() == 1).all()
        with pytest.raises(Exception):
            p.set_active(np.zeros(p.n, np.int32))
