#!/usr/bin/env python3
"""PonD step benchmark: prints ONE JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c4] [--interp I] [--n N]
                    [--no-cpu] [--no-sensitivity]

Headline (the metric of the shock-bubble configuration, c4): MLUPS and mean gauge iterations
per step of the learned-initializer + FP32/FP64-split PonD step against the plain FP64 step on
D3Q27 384^3, with the error of the split paths against the FP64 path stated.

Each path starts from the same seeded initial condition and takes C + W + K steps: C "collect"
steps (the learned path samples its training data from them and trains its network after them;
the other paths just step), W untimed warmup steps, then K timed steps.  A step is one PonD step
over the whole grid (gauge fixed-point iteration, pull streaming, BGK collide; for the learned
path also the network inference that predicts the starting gauge).  Inputs are resident in HBM
before the timed region; every timed step is bracketed by device synchronisation; MLUPS =
cells * K / wall time of the timed steps.  The step kernels' own times come from HIP events
recorded on the library's stream around every launch (pond_timing_info).

Accuracy (split paths against FP64), three figures:
  * one_step: from the FP64 path's own state at two points of its run (after the warmup and
    before its last timed step), one step of each split path and of FP64, compared over all
    cells, iteration-capped ones included.  This is the error the split arithmetic adds per
    step, and what the <= 1e-5 gate is stated on.
  * trajectory: the states after all C + W + K steps, each path on its own trajectory.
  * fp64_sensitivity: the FP64 path's own response over the same steps to a 1e-7 relative
    perturbation of its initial state (run in lockstep with the FP64 path).  A trajectory
    difference that the split arithmetic seeds at ~1e-7 per step grows like this.

The CPU baseline is the oracle (the restatement of the reference's step, as written, no
memoisation) on the host threads, on a bounded sample of the same workload taken at the same
point of the run (the state after C + W steps, computed by the bit-identical FP64 GPU path on a
reduced grid; the D1Q5 tube in full).  Only one GPU is used; --gpus above 1 exits with a message.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from lbmpond_amd.compare import CapDependence, state_errors  # noqa: E402

METRIC = "MLUPS + mean gauge iters/step, learned+FP32 vs plain FP64, at stated L∞ err"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6
VALU_ISSUE_PEAK = 256 * 4 * 16 * 2.4e9 / 1e12  # T lane-instructions/s
ACCURACY_GATE = 1e-5

# algorithmic bytes per lattice update: read one state, write one state, int32 iterations
BYTES_PER_LUP = {("D3Q27", "fp64"): 2 * (27 * 8 + 5 * 8) + 4, ("D3Q27", "split"): 2 * (27 * 4 + 4 * 4 + 8) + 4,
                 ("D3Q125", "fp64"): 2 * (125 * 8 + 5 * 8) + 4, ("D3Q125", "split"): 2 * (125 * 4 + 4 * 4 + 8) + 4,
                 ("D1Q5", "fp64"): 2 * (5 * 8 + 5 * 8) + 4}

# per configuration: default steps, warmup, collect steps, paths, CPU sample grid edge
CONFIGS = {
    "c1": dict(steps=1000, warmup=10, collect=0, paths="fp64", cpu_n=None),
    "c2": dict(steps=3, warmup=1, collect=0, paths="split,fp64", cpu_n=32),
    "c3": dict(steps=3, warmup=1, collect=0, paths="split,fp64", cpu_n=32),
    "c4": dict(steps=3, warmup=1, collect=4, paths="learned,split,fp64", cpu_n=32),
    "c5": dict(steps=2, warmup=1, collect=2, paths="learned,split,fp64", cpu_n=8),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_case(name, n):
    from lbmpond_amd import ic
    if name == "c1":
        return ic.sod_1d(nx=n or 2000)
    if name == "c2":
        return ic.vortex_3d(n=n or 256)
    if name == "c3":
        return ic.taylor_green(n=n or 384)
    if name == "c4":
        return ic.shock_bubble(n=n or 384)
    if name == "c5":
        return ic.isotropic_turbulence(n=n or 256)
    raise SystemExit(f"unknown config {name}")


def default_interp(case):
    # the as-written WENO5 is unstable in 3D (its y/z stencil extrapolates downwind); the 3D
    # D3Q27 configurations use the centred WENO5; D3Q125 has its fixed 3x3x3 stencil
    return "weno5c" if case.lattice == "D3Q27" else case.interp


def kernel_name(lattice, prec, interp):
    if lattice == "D3Q125":
        return "k_step_q125_f64" if prec == "fp64" else "k_step_split125"
    if lattice == "D1Q5":
        return "k_step_f64<0,1,false>" if interp == "lagr4" else "k_step_f64<0,0,false>"
    code = {"weno5": 0, "lagr4": 1, "weno5c": 2}[interp]
    if prec == "fp64":
        return f"k_step_f64_refill<1,{code},true>"
    return f"k_step_split27_refill<{code}>" if code != 0 else "k_step_split27<0>"


class Run:
    """One path's context, stepped through collect, warmup and timed steps."""

    def __init__(self, case, interp, name, train_cfg):
        from lbmpond_amd import Pond
        self.name = name
        self.prec = "fp64" if name.startswith("fp64") else "split"
        self.learned = name == "learned"
        self.p = Pond(case.lattice, case.shape, interp, self.prec, dtau=case.dtau, use_mlp=self.learned)
        self.train_cfg = train_cfg
        self.dep = CapDependence(case.shape, self.p.params.max_iter)
        self.iters, self.wall, self.res = [], 0.0, {"path": name}
        self.max_departure = 0.0

    def step(self, phase, k, timed=False):
        if timed:
            self.p.sync()
            w0 = time.perf_counter()
            self.p.step(1)
            self.p.sync()
            self.wall += time.perf_counter() - w0
        else:
            self.p.step(1)
            self.p.sync()
        if phase == "collect" and self.learned:
            self.p.mlp_collect(self.train_cfg["samples_per_step"], seed=1000 + k)
        d = self.p.diagnostics()
        self.max_departure = max(self.max_departure, d["max_departure"])
        if timed:
            self.iters.append(d["niter_mean"])
        self.dep.update(self.p.niter())

    def train(self):
        cfg = self.train_cfg
        t = time.perf_counter()
        losses = self.p.mlp_train(epochs=cfg["epochs"], batch=cfg["batch"], lr=cfg["lr"], seed=11)
        self.res.update(train_s=round(time.perf_counter() - t, 3), train_samples=self.p.mlp_dataset_size(),
                        train_loss=[round(float(x), 5) for x in losses])
        log(f"[bench] {self.name}: trained on {self.res['train_samples']} samples in {self.res['train_s']} s, "
            f"losses {self.res['train_loss']}")


def run_path(case, interp, name, collect, warmup, steps, train_cfg, hooks=None, keep=False):
    """C + W + K steps of one path; hooks(step_index, run) runs before each step (outside the
    timed region).  Returns the run (its context closed unless keep)."""
    r = Run(case, interp, name, train_cfg)
    r.p.set_state(*case.fields)
    total = collect + warmup + steps
    t0 = r.p.timing()
    last_log = time.time()
    for s in range(total):
        if hooks:
            hooks(s, r)
        phase = "collect" if s < collect else "warmup" if s < collect + warmup else "timed"
        if phase == "timed" and s == collect + warmup:
            t0 = r.p.timing()
        r.step(phase, s, timed=phase == "timed")
        if phase == "collect" and s == collect - 1 and r.learned:
            r.train()
        if total <= 40 or (s + 1) % 100 == 0 or s + 1 == total or time.time() - last_log > 30:
            log(f"[bench] {name}: {phase} step {s + 1}/{total}")
            last_log = time.time()
    t1 = r.p.timing()
    cells = int(np.prod(case.shape))
    r.res.update(cells=cells, wall_s=r.wall, kernel_ms=t1["total_ms"] - t0["total_ms"], mlp_ms=t1["last_mlp_ms"],
                 mlups=cells * steps / r.wall / 1e6, niter_mean=float(np.mean(r.iters)),
                 niter_max=int(r.p.diagnostics()["niter_max"]), max_departure=r.max_departure)
    r.res["fields"] = r.p.fields()
    it = r.res["fields"]["niter"].astype(np.float64)
    pad = (-it.size) % 64
    r.res["niter_wave_mean"] = float(np.pad(it, (0, pad)).reshape(-1, 64).max(axis=1).mean())
    if not keep:
        r.p.close()
    return r


def errs(got, ref, mask=None):
    return {k: float(f"{v:.3e}") for k, v in state_errors(got, ref, mask).items()}


def cpu_baseline(case_name, interp, collect, warmup, cfg):
    """The oracle on a bounded sample of the same workload at the same point of the run."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from lbmpond_amd import Pond
    cpu = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as fh:
            cpu = next((ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name")), cpu)
    except OSError:
        pass
    if cfg["cpu_n"] is None:
        # the whole workload: the D1Q5 tube from its initial condition, every step
        case = make_case(case_name, None)
        o = pyoracle.Oracle(case.lattice, case.shape, interp, dtau=case.dtau)
        o.set_state(*case.fields)
        nsteps = case.steps
        t = time.perf_counter()
        o.step(nsteps)
        dt = time.perf_counter() - t
        sample = f"the whole workload: {case.shape[0]} cells x {nsteps} steps from the initial condition"
        niter = float(o.fields()["niter"].mean())
    else:
        # the state after collect + warmup steps of a reduced grid of the same initial-condition
        # family, computed by the FP64 GPU path (bit-identical to the oracle), then oracle steps
        n = cfg["cpu_n"]
        case = make_case(case_name, n)
        with Pond(case.lattice, case.shape, interp, "fp64", dtau=case.dtau) as p:
            p.set_state(*case.fields)
            p.step(collect + warmup)
            fl = p.fields()
            f = p.populations()
        o = pyoracle.Oracle(case.lattice, case.shape, interp, dtau=case.dtau)
        o.set_populations(f, fl["rho"], fl["ux"], fl["uy"], fl["uz"], fl["T"])
        nsteps, dt = 0, 0.0
        its = []
        while nsteps < 1 or (dt < 10.0 and nsteps < 4):
            t = time.perf_counter()
            o.step(1)
            dt += time.perf_counter() - t
            its.append(float(o.fields()["niter"].mean()))
            nsteps += 1
        niter = float(np.mean(its))
        sample = (f"{n}^3 cells of the {case_name} initial-condition family, steps {collect + warmup + 1}.."
                  f"{collect + warmup + nsteps} (the state after {collect + warmup} steps from the FP64 GPU path)")
    cells = int(np.prod(case.shape))
    return {"value": round(cells * nsteps / dt / 1e6, 6), "unit": "MLUPS", "cores": pyoracle.threads(),
            "kind": "port", "cpu": cpu,
            "cell_iterations_per_s": round(cells * nsteps * niter / dt, 1), "niter_mean": round(niter, 4),
            "sample": f"oracle (the reference's step as written, no memoisation), {case.lattice} {interp}, "
                      f"{sample}: {cells * nsteps} cell-steps in {dt:.1f} s"}


def profiled(tag, kernel, workload, key):
    """A per-launch figure of `kernel` on `workload` from the committed rocprofv3 summaries: a
    line 'TAG: kernel=K workload=W KEY=V ...' in profiles/*.md.  Returns (V, file)."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        with open(path) as fh:
            for line in fh:
                line = line.strip().strip("`")
                if not line.startswith(tag + ":"):
                    continue
                kv = dict(t.split("=", 1) for t in line[len(tag) + 1:].split() if "=" in t)
                if kv.get("kernel") == kernel and kv.get("workload") == workload and key in kv:
                    return float(kv[key]), os.path.relpath(path, ROOT)
    return None, None


def copy_calibration(nbytes):
    """One device-to-device copy of nbytes through the HIP runtime: the reference point that
    scales rocprof's FETCH_SIZE / WRITE_SIZE readings (a plain copy reads and writes nbytes)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for ptr in (a, b):
        if hip.hipMalloc(ctypes.byref(ptr), ctypes.c_size_t(nbytes)) != 0:
            raise SystemExit("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    t = time.perf_counter()
    if hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) != 0:  # hipMemcpyDeviceToDevice
        raise SystemExit("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    dt = time.perf_counter() - t
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_bytes": nbytes, "seconds": dt, "GBps_rw": 2 * nbytes / dt / 1e9}), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--collect", type=int, default=None, help="steps whose cells train the network")
    ap.add_argument("--config", default="c4", choices=sorted(CONFIGS))
    ap.add_argument("--interp", default=None, choices=["weno5", "weno5c", "lagr4"])
    ap.add_argument("--n", type=int, default=None, help="grid edge override (testing only)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-sensitivity", action="store_true", help="skip the perturbed FP64 run")
    ap.add_argument("--paths", default=None)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run one device copy of BYTES (rocprof traffic calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (the decomposition over several GPUs is not built); "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)
    cfg = CONFIGS[args.config]
    steps = cfg["steps"] if args.steps is None else args.steps
    warmup = cfg["warmup"] if args.warmup is None else args.warmup
    collect = cfg["collect"] if args.collect is None else args.collect
    paths = (args.paths or cfg["paths"]).split(",")
    # the FP64 path runs last: the one-step checks start the split paths from its state
    paths = [p for p in paths if p != "fp64"] + (["fp64"] if "fp64" in paths else [])
    if "learned" in paths and collect < 1:
        raise SystemExit("the learned path needs at least one collect step")

    case = make_case(args.config, args.n)
    interp = args.interp or default_interp(case)
    cells = int(np.prod(case.shape))
    log(f"[bench] initial condition ready: {case.name} {case.lattice} {case.shape} {interp}")
    train_cfg = {"samples_per_step": max(2048, min(500_000, cells // 8)), "epochs": 3, "batch": 4096, "lr": 2e-3}
    total = collect + warmup + steps
    results = {}
    mlp_params = None
    for name in paths:
        if name == "fp64":
            continue
        keep = name == "learned"
        r = run_path(case, interp, name, collect, warmup, steps, train_cfg, keep=keep)
        if keep:
            mlp_params = r.p.mlp_params()
            r.p.close()
        results[name] = r

    accuracy = {"gate": ACCURACY_GATE, "one_step": []}
    if "fp64" in paths:
        from lbmpond_amd import Pond
        split_paths = [n for n in results]
        sens = None
        if not args.no_sensitivity and split_paths:
            rng = np.random.default_rng(17)
            rho, ux, uy, uz, T = case.fields
            U = lambda: rng.uniform(-1.0, 1.0, rho.shape)
            sens = Pond(case.lattice, case.shape, interp, "fp64", dtau=case.dtau)
            sens.set_state(rho * (1 + 1e-7 * U()), ux + 1e-7 * np.sqrt(np.abs(T)) * U(), uy, uz, T * (1 + 1e-7 * U()))
        checks = sorted({collect + warmup, total - 1}) if split_paths else []
        pending = {}

        def hooks(s, run):
            # before FP64 step s+1: one step of every split path from the FP64 state; the FP64
            # step itself is the reference (compared after it, in the next call)
            if pending:
                ref = run.p.fields()
                entry = {"from_step": pending["from"], "fp64_capped_cells": int((ref["niter"] >= 100).sum())}
                for name, got in pending["fields"].items():
                    # where either step stopped at the cap its value is an unconverged iterate
                    # (of a cycle, in the FP64 step too), not a fixed point to compare
                    conv = (got["niter"] < 100) & (ref["niter"] < 100)
                    entry[name] = {"all_cells": errs(got, ref),
                                   "converged_cells": errs(got, ref, conv) if not conv.all() else errs(got, ref),
                                   "capped_cells_here_not_fp64": int(((got["niter"] >= 100) & (ref["niter"] < 100)).sum()),
                                   "niter_mean": round(float(got["niter"].mean()), 4)}
                accuracy["one_step"].append(entry)
                pending.clear()
            if sens is not None and s > 0:
                sens.step(1)
                sens.sync()  # not running beside the FP64 path's timed step
            if s in checks:
                out = {}
                for name in split_paths:
                    q = Pond(case.lattice, case.shape, interp, "split", dtau=case.dtau, use_mlp=name == "learned")
                    if name == "learned":
                        q.mlp_set_params(*mlp_params)
                    q.copy_state_from(run.p)
                    q.step(1)
                    out[name] = q.fields()
                    q.close()
                pending.update({"from": s, "fields": out})
                log(f"[bench] one-step checks from FP64 step {s}")

        r = run_path(case, interp, "fp64", collect, warmup, steps, train_cfg, hooks=hooks, keep=True)
        hooks(total, r)  # the last check's comparison
        if sens is not None:
            accuracy["fp64_sensitivity"] = {"perturbation": 1e-7, "steps": total,
                                            "rel_linf": errs(sens.fields(), r.res["fields"])}
            sens.close()
        r.p.close()
        results["fp64"] = r

    ref = results.get("fp64")
    head = results.get("learned") or results.get("split") or ref
    head_name = head.name
    summary = {}
    for name, r in results.items():
        res = r.res
        s = {"mlups": round(res["mlups"], 4), "niter_mean": round(res["niter_mean"], 4), "niter_max": res["niter_max"],
             "niter_wave_mean_last": round(res["niter_wave_mean"], 4),
             "ms_per_step": round(res["wall_s"] * 1e3 / steps, 3),
             "kernel_ms_per_step": round(res["kernel_ms"] / steps, 3),
             "max_departure_cells_per_step": round(res["max_departure"], 4)}
        if ref is not None and name != "fp64":
            s["trajectory_rel_linf_vs_fp64"] = errs(res["fields"], ref.res["fields"])
            # cells a capped cell (whose result is an unconverged iterate of a cycle, in the FP64
            # path too) can have influenced, at the measured departure reach
            dep = r.dep.mask | ref.dep.mask
            s["capped_cells_per_step"] = {"fp64": ref.dep.capped, name: r.dep.capped}
            s["cap_dependence_fraction"] = round(float(dep.mean()), 5)
            if not dep.all():
                s["trajectory_rel_linf_vs_fp64_outside_cap_dependence"] = errs(res["fields"], ref.res["fields"], ~dep)
        for k in ("train_s", "train_samples", "train_loss"):
            if k in res:
                s[k] = res[k]
        if res.get("mlp_ms"):
            s["mlp_infer_ms"] = round(res["mlp_ms"], 3)
        summary[name] = s

    prec = head.prec
    launch_s = head.res["kernel_ms"] * 1e-3 / steps
    bpl = BYTES_PER_LUP[(case.lattice, prec)]
    achieved = bpl * cells / launch_s / 1e9
    kernel = kernel_name(case.lattice, prec, interp)
    workload = f"{case.name}-{case.lattice}-{'x'.join(map(str, case.shape))}"
    traffic, traffic_src = profiled("traffic", kernel, workload, "bytes_per_launch")
    one = accuracy["one_step"]
    if one:
        worst = {k: max(e[head_name]["all_cells"][k] for e in one) for k in ("rho", "u", "T")}
        worst_conv = {k: max(e[head_name]["converged_cells"][k] for e in one) for k in ("rho", "u", "T")}
        accuracy["gate_rel_linf"] = worst
        accuracy["gate_definition"] = ("max over the one-step checks of the relative L-inf error of the headline "
                                       "path against the FP64 step from the same FP64 state, all cells")
        accuracy["gate_met"] = all(v <= ACCURACY_GATE for v in worst.values())
        accuracy["gate_rel_linf_converged_cells"] = worst_conv
        accuracy["gate_met_converged_cells"] = all(v <= ACCURACY_GATE for v in worst_conv.values())
    if ref is not None and head is not ref:
        accuracy["trajectory"] = summary[head_name]["trajectory_rel_linf_vs_fp64"]
    max_dep = max(r.res["max_departure"] for r in results.values())
    accuracy["max_departure_cells_per_step"] = round(max_dep, 4)
    out = {
        "metric": METRIC,
        "value": round(head.res["mlups"], 4),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(head.res["wall_s"] * 1e3 / steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if prec == "split" else "f64",
        "data": "synthetic",
        "config": {"workload": f"{case.name}: {case.lattice} {'x'.join(map(str, case.shape))} {interp} PonD step, "
                               f"path {head_name}", "lattice": case.lattice, "grid": list(case.shape),
                   "interp": interp, "dtau": case.dtau, "collect_steps": collect, "paths": paths,
                   "initial_condition": {k: v for k, v in case.info.items() if isinstance(v, (int, float, str))}},
        "niter_mean": round(head.res["niter_mean"], 4),
        "rel_linf_vs_fp64": accuracy.get("gate_rel_linf"),
        "accuracy": accuracy,
        "speedup_vs_fp64": round(head.res["mlups"] / ref.res["mlups"], 3) if ref is not None and head is not ref else None,
        "paths": summary,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic, "bytes_per_lup": bpl,
                     "algorithmic_bytes_per_launch": bpl * cells, "kernel": kernel, "traffic_source": traffic_src},
    }
    # executed VALU work of the headline kernel per cell-iteration (rocprofv3 counters of the same
    # kernel and workload, in profiles/), at this run's cell-iterations and launch time
    cell_iters = cells * head.res["niter_mean"]
    f32, src32 = profiled("flops", kernel, workload, "f32_flop_per_cell_iter")
    f64, _ = profiled("flops", kernel, workload, "f64_flop_per_cell_iter")
    insts, src_i = profiled("valu", kernel, workload, "valu_lane_insts_per_cell_iter")
    if f32 is not None and f64 is not None:
        r32 = f32 * cell_iters / launch_s / 1e12
        r64 = f64 * cell_iters / launch_s / 1e12
        out["valu_flops"] = {"f32_tflops": round(r32, 3), "f64_tflops": round(r64, 3),
                             "frac_of_peak": round(r32 / FP32_PEAK_TFLOPS + r64 / FP64_PEAK_TFLOPS, 4),
                             "peaks_tflops": {"f32": FP32_PEAK_TFLOPS, "f64": FP64_PEAK_TFLOPS}, "kernel": kernel,
                             "source": src32}
    if insts is not None:
        rate = insts * cell_iters / launch_s / 1e12
        out["valu_issue"] = {"achieved": round(rate, 3), "peak": VALU_ISSUE_PEAK, "unit": "T lane-instructions/s",
                             "frac": round(rate / VALU_ISSUE_PEAK, 4), "kernel": kernel, "source": src_i}
    if not args.no_cpu:
        log("[bench] cpu baseline")
        cb = cpu_baseline(args.config, interp, collect, warmup, cfg)
        cb["gpu_cell_iterations_per_s"] = round(cells * head.res["niter_mean"] * steps / head.res["wall_s"], 1)
        out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
