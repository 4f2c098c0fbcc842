"""Accuracy of the FP32/FP64-split step and of the learned start against the plain FP64 step on
the shock-bubble (Mach-2 shock, light bubble; the benchmark's configuration on a 64^3 grid).

The gate (<= 1e-5 relative L-inf on rho, u, T) is stated on one step from the same FP64 state,
over all cells, iteration-capped ones included: the error the split arithmetic adds per step.
Along a trajectory the split path's per-step rounding (~2e-7) is amplified by the flow itself
(the shock and the bubble's contact layer); the FP64 path's own response to a 1e-7 perturbation
of its initial state bounds that growth, as in test_split_gpu.py's sensitivity tests.
"""
import numpy as np
import pytest

from helpers import STATE_ERRORS
from lbmpond_amd import Pond, ic
from lbmpond_amd.compare import CapDependence, state_errors

pytestmark = pytest.mark.gpu

N = 64
COLLECT = 4
STEPS = 16
CHECKS = (4, 8, 12)


def mk(case, precision, **kw):
    p = Pond("D3Q27", case.shape, "weno5c", precision, dtau=case.dtau, **kw)
    return p


@pytest.fixture(scope="module")
def runs(gpu):
    case = ic.shock_bubble(n=N)
    ref, split, learned = mk(case, "fp64"), mk(case, "split"), mk(case, "split", use_mlp=True)
    pert = mk(case, "fp64")
    rho, ux, uy, uz, T = case.fields
    rng = np.random.default_rng(17)
    U = lambda: rng.uniform(-1, 1, rho.shape)
    pert.set_state(rho * (1 + 1e-7 * U()), ux + 1e-7 * np.sqrt(T) * U(), uy, uz, T * (1 + 1e-7 * U()))
    for p in (ref, split, learned):
        p.set_state(*case.fields)
    one_step, departure = [], []
    dep_ref, dep_learned = CapDependence(case.shape, 100), CapDependence(case.shape, 100)
    for s in range(STEPS):
        if s in CHECKS:
            out = {}
            for name in ("split", "learned"):
                q = mk(case, "split", use_mlp=name == "learned")
                if name == "learned":
                    q.mlp_set_params(*learned.mlp_params())
                q.copy_state_from(ref)
                q.step(1)
                out[name] = q.fields()
                q.close()
        for p in (ref, split, learned, pert):
            p.step(1)
        if s < COLLECT:
            learned.mlp_collect(N ** 3 // 8, seed=100 + s)
        if s == COLLECT - 1:
            learned.mlp_train(epochs=4, batch=4096, lr=2e-3, seed=5)
        departure.append(ref.diagnostics()["max_departure"])
        dep_ref.update(ref.niter())
        dep_learned.update(learned.niter())
        if s in CHECKS:
            fr = ref.fields()
            one_step.append((s, fr, out))
            print(f"[accuracy] one step from FP64 step {s}: "
                  + ", ".join(f"{k} {state_errors(v, fr)}" for k, v in out.items()), flush=True)
    res = {"ref": ref.fields(), "split": split.fields(), "learned": learned.fields(), "pert": pert.fields(),
           "one_step": one_step, "departure": departure, "dep_ref": dep_ref, "dep_learned": dep_learned}
    for p in (ref, split, learned, pert):
        p.close()
    return res


@pytest.mark.parametrize("path", ["split", "learned"])
def test_one_step_from_the_fp64_state_within_1e5_on_all_cells(runs, path):
    for s, fr, out in runs["one_step"]:
        errs = state_errors(out[path], fr)
        capped_here = int(((out[path]["niter"] >= 100) & (fr["niter"] < 100)).sum())
        print(f"  step {s} {path}: {errs}, mean iterations {out[path]['niter'].mean():.3f} "
              f"(FP64 {fr['niter'].mean():.3f}), capped here only {capped_here}", flush=True)
        for k in STATE_ERRORS:
            assert errs[k] <= 1e-5, (s, path, k, errs[k])


def test_split_trajectory_within_the_fp64_sensitivity_bound(runs):
    """after 16 steps the split path on its own trajectory: bounded by the number of steps times
    the FP64 path's response to a 1e-7 perturbation of its initial state (twice that for the
    noise floor of the split residual).  (No such bound holds for the learned start along a
    trajectory: where the gauge map cycles, the FP64 step ends on its 100th iterate, a predicted
    start can converge instead, and on this grid the cells such a cell can influence cover the
    whole grid within 16 steps; its per-step error is tested above.)"""
    e_pert = state_errors(runs["pert"], runs["ref"])
    e = state_errors(runs["split"], runs["ref"])
    dep = runs["dep_ref"].mask | runs["dep_learned"].mask
    print(f"  split trajectory {e}; learned {state_errors(runs['learned'], runs['ref'])}; FP64 response to 1e-7: "
          f"{e_pert}; cells a capped cell can have influenced {dep.mean():.3f}", flush=True)
    for k in STATE_ERRORS:
        assert e[k] <= 2 * STEPS * max(e_pert[k], 1e-7), (k, e[k], e_pert[k])


def test_departures_stay_within_the_cap_dependence_reach(runs):
    """lbmpond_amd.compare takes a difference to spread at most 3 cells per step: the WENO5C
    stencil reaches 2 cells beyond the nearest node of the departure point, so this holds while
    every population moves less than 1.5 cells per step"""
    print(f"  largest departure per step: {max(runs['departure']):.4f} cells", flush=True)
    assert max(runs["departure"]) < 1.5


def test_d3q125_learned_and_split_against_fp64(gpu):
    """D3Q125 (PonDstepD3Q125's fixed 3x3x3 stencil) on the turbulence configuration, 16^3: the
    split step with the learned start (network trained on the solver's first steps) against the
    FP64 step, one step from the same FP64 state and over the whole run, within 1e-5"""
    case = ic.isotropic_turbulence(n=16)
    kw = dict(dtau=case.dtau)
    with Pond("D3Q125", case.shape, "weno5", "fp64", **kw) as ref, \
            Pond("D3Q125", case.shape, "weno5", "split", use_mlp=True, **kw) as lr:
        ref.set_state(*case.fields)
        lr.set_state(*case.fields)
        for s in range(6):
            if s == 4:
                with Pond("D3Q125", case.shape, "weno5", "split", use_mlp=True, **kw) as q:
                    q.mlp_set_params(*lr.mlp_params())
                    q.copy_state_from(ref)
                    q.step(1)
                    one = q.fields()
                    used_mlp = q.timing()["last_mlp_ms"] > 0
            ref.step(1)
            lr.step(1)
            if s < 3:
                lr.mlp_collect(2048, seed=40 + s)
            if s == 2:
                lr.mlp_train(epochs=6, batch=512, lr=2e-3, seed=3)
            if s == 4:
                e_one = state_errors(one, ref.fields())
        e_traj = state_errors(lr.fields(), ref.fields())
    print(f"  D3Q125 learned: one step {e_one}, 6 steps {e_traj}, inference ran {used_mlp}", flush=True)
    assert used_mlp
    for k in STATE_ERRORS:
        assert e_one[k] <= 1e-5, (k, e_one[k])
        assert e_traj[k] <= 1e-5, (k, e_traj[k])
