#!/usr/bin/env python3
"""PonD step benchmark: prints ONE JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c4] [--n N] [--no-cpu]

Headline (BASELINE metric): MLUPS and mean gauge iterations per step of the learned-initializer
+ FP32/FP64-split PonD step against the plain FP64 step on D3Q27 384^3, with the relative
L-inf error of the split paths against the FP64 run on the same inputs after the same steps.

Each path starts from the same seeded initial condition and takes C + W + K steps: C
"collect" steps (the learned path samples its training data from them and trains its
network after them; the other paths just step), W untimed warmup steps, then K timed steps.
A step is one PonD step over the whole grid (gauge fixed-point iteration, pull streaming,
BGK collide; for the learned path also the network inference that predicts the starting
gauge).  Inputs are resident in HBM before the timed region, which is bracketed by device
synchronisation; MLUPS = cells * K / wall time.  The step kernels' own times come from HIP
events recorded on the library's stream around every launch (pond_timing_info).

The CPU baseline is the oracle (the restatement of the reference's step, as written, no
memoisation) timed on the host threads on a bounded sample of the same initial-condition
family.  Only one GPU is used; --gpus above 1 exits with a message.
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
FP64_PEAK_TFLOPS = 78.6
VALU_ISSUE_PEAK = 256 * 4 * 16 * 2.4e9 / 1e12  # T lane-instructions/s

# algorithmic bytes per lattice update: read one state, write one state, int32 iterations
BYTES_PER_LUP = {("D3Q27", "fp64"): 2 * (27 * 8 + 5 * 8) + 4, ("D3Q27", "split"): 2 * (27 * 4 + 4 * 4 + 8) + 4,
                 ("D3Q125", "fp64"): 2 * (125 * 8 + 5 * 8) + 4, ("D3Q125", "split"): 2 * (125 * 4 + 4 * 4 + 8) + 4,
                 ("D1Q5", "fp64"): 2 * (5 * 8 + 5 * 8) + 4}
# the step kernel each path launches for the bench configurations
KERNEL_NAME = {("D3Q27", "fp64"): "k_step_f64<1,2,true>", ("D3Q27", "split"): "k_step_split27_refill<2>",
               ("D3Q125", "fp64"): "k_step_q125_f64", ("D3Q125", "split"): "k_step_split125"}
# algorithmic FP64 flops per cell-iteration of the reference's algorithm with de-duplicated
# transfers: 19 flops per D3Q27 transfer term, 27 targets x 50 transfers x 27 terms (WENO5)
FLOP_PER_CELL_ITER = {"D3Q27": 27 * 50 * 27 * 19, "D3Q125": 125 * 27 * 125 * 29, "D1Q5": 5 * 5 * 5 * 30}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_case(name, n):
    from lbmpond_amd import ic
    if name == "c2":
        return ic.vortex_3d(n=n or 256)
    if name == "c3":
        return ic.taylor_green(n=n or 384)
    if name == "c4":
        return ic.shock_bubble(n=n or 384)
    if name == "c5":
        return ic.isotropic_turbulence(n=n or 256)
    raise SystemExit(f"unknown config {name}")


def interp_for(case):
    # the as-written WENO5 is unstable in 3D; the 3D configurations use the centred variant
    return "weno5c" if case.lattice == "D3Q27" else case.interp


def run_path(case, precision, learned, collect, warmup, steps, train_cfg):
    from lbmpond_amd import Pond
    interp = interp_for(case)
    tag = f"{precision}{'+learned' if learned else ''}"
    p = Pond(case.lattice, case.shape, interp, precision, dtau=case.dtau, use_mlp=learned)
    p.set_state(*case.fields)
    res = {"path": tag}
    dep = CapDependence(case.shape, p.params.max_iter)  # cells downstream of an iteration-capped cell
    for c in range(collect):
        p.step(1)
        if learned:
            p.mlp_collect(train_cfg["samples_per_step"], seed=1000 + c)
        p.sync()
        dep.update(p.niter())
        log(f"[bench] {tag}: collect step {c + 1}/{collect}")
    if learned:
        t = time.perf_counter()
        losses = p.mlp_train(epochs=train_cfg["epochs"], batch=train_cfg["batch"], lr=train_cfg["lr"], seed=11)
        res["train_s"] = round(time.perf_counter() - t, 3)
        res["train_samples"] = p.mlp_dataset_size()
        res["train_loss"] = [round(float(x), 5) for x in losses]
        log(f"[bench] {tag}: trained on {res['train_samples']} samples in {res['train_s']} s, losses {res['train_loss']}")
    for w in range(warmup):
        p.step(1)
        p.sync()
        dep.update(p.niter())
        log(f"[bench] {tag}: warmup step {w + 1}/{warmup}")
    t0 = p.timing()
    iters = []
    p.sync()
    wall = 0.0
    for k in range(steps):
        # each timed step is bracketed by device synchronisation; the per-cell iteration counts
        # are read between steps, outside the timed region
        w0 = time.perf_counter()
        p.step(1)
        p.sync()
        wall += time.perf_counter() - w0
        iters.append(p.diagnostics()["niter_mean"])
        dep.update(p.niter())
        log(f"[bench] {tag}: timed step {k + 1}/{steps}")
    t1 = p.timing()
    cells = int(np.prod(case.shape))
    res.update(cells=cells, wall_s=wall, kernel_ms=t1["total_ms"] - t0["total_ms"], mlp_ms=t1["last_mlp_ms"],
               mlups=cells * steps / wall / 1e6, niter_mean=float(np.mean(iters)),
               niter_max=int(p.diagnostics()["niter_max"]))
    res["fields"] = p.fields()
    res["cap_dependence"] = dep
    # a wave of 64 consecutive cells runs as many iterations as its slowest cell
    it = res["fields"]["niter"].astype(np.float64)
    pad = (-it.size) % 64
    res["niter_wave_mean"] = float(np.pad(it, (0, pad)).reshape(-1, 64).max(axis=1).mean())
    p.close()
    return res


def cpu_baseline(case_name, interp):
    """oracle on a bounded sample of the same IC family, all host threads"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    sample_n = {"c2": 32, "c3": 32, "c4": 32, "c5": 8}[case_name]
    case = make_case(case_name, sample_n)
    o = pyoracle.Oracle(case.lattice, case.shape, interp, dtau=case.dtau)
    o.set_state(*case.fields)
    t = time.perf_counter()
    o.step(1)
    dt = time.perf_counter() - t
    cells = int(np.prod(case.shape))
    niter = float(o.fields()["niter"].mean())
    cpu = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as fh:
            cpu = next((ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name")), cpu)
    except OSError:
        pass
    return {"value": round(cells / dt / 1e6, 6), "unit": "MLUPS", "cores": pyoracle.threads(), "kind": "port",
            "cpu": cpu,
            "sample": f"oracle (reference algorithm as written, no memoisation), {case.lattice} {interp} "
                      f"{sample_n}^3 cells x 1 step of the {case_name} initial-condition family "
                      f"({cells} cell-steps in {dt:.1f} s, mean gauge iterations {niter:.2f}, "
                      f"{cells * niter / dt / 1e6:.5f} M cell-iterations/s)"}


def profiled(tag, kernel, workload, key):
    """A per-launch figure of `kernel` on `workload` from the committed rocprofv3 counter
    summaries: a line 'TAG: kernel=K workload=W KEY=V' in profiles/*.md.  Returns (V, file)."""
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


def profiled_traffic(kernel, workload):
    """HBM bytes per launch (FETCH_SIZE and WRITE_SIZE per dispatch, scaled by the same counters'
    reading of a plain device copy)"""
    return profiled("traffic", kernel, workload, "bytes_per_launch")


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
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--collect", type=int, default=4, help="steps whose cells train the network")
    ap.add_argument("--config", default="c4")
    ap.add_argument("--n", type=int, default=None, help="grid edge override (testing only)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--paths", default="fp64,split,learned")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run one device copy of BYTES (rocprof traffic calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only; how the step would shard is described in DESIGN.md; "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)

    case = make_case(args.config, args.n)
    cells = int(np.prod(case.shape))
    log(f"[bench] initial condition ready: {case.name} {case.lattice} {case.shape}")
    train_cfg = {"samples_per_step": max(2048, min(500_000, cells // 8)), "epochs": 3, "batch": 4096, "lr": 2e-3}
    paths = args.paths.split(",")
    results = {}
    for name in paths:
        prec = "fp64" if name == "fp64" else "split"
        results[name] = run_path(case, prec, name == "learned", args.collect, args.warmup, args.steps, train_cfg)

    ref = results.get("fp64")
    head = results.get("learned") or results.get("split") or ref
    summary = {}
    for name, r in results.items():
        s = {"mlups": round(r["mlups"], 4), "niter_mean": round(r["niter_mean"], 4), "niter_max": r["niter_max"],
             "niter_wave_mean_last": round(r["niter_wave_mean"], 4),
             "ms_per_step": round(r["wall_s"] * 1e3 / args.steps, 3),
             "kernel_ms_per_step": round(r["kernel_ms"] / args.steps, 3)}
        if ref is not None and name != "fp64":
            s["rel_linf_vs_fp64"] = {k: float(f"{v:.3e}") for k, v in state_errors(r["fields"], ref["fields"]).items()}
            # a cell that hit the iteration cap in either run ends its step on an unconverged
            # iterate that depends on where its loop started; the error over the cells outside
            # the region such cells can have influenced (3 cells per step)
            dep = r["cap_dependence"].mask | ref["cap_dependence"].mask
            s["capped_cells_per_step"] = {"fp64": ref["cap_dependence"].capped, name: r["cap_dependence"].capped}
            s["cells_in_cap_dependence"] = int(dep.sum())
            if not dep.all():
                s["rel_linf_vs_fp64_outside_cap_dependence"] = {
                    k: float(f"{v:.3e}") for k, v in state_errors(r["fields"], ref["fields"], ~dep).items()}
        for k in ("train_s", "train_samples", "train_loss"):
            if k in r:
                s[k] = r[k]
        if r.get("mlp_ms"):
            s["mlp_infer_ms"] = round(r["mlp_ms"], 3)
        summary[name] = s
    head_name = [k for k, v in results.items() if v is head][0]
    prec = "fp64" if head_name == "fp64" else "split"
    launch_s = head["kernel_ms"] * 1e-3 / args.steps
    bpl = BYTES_PER_LUP[(case.lattice, prec)]
    achieved = bpl * cells / launch_s / 1e9
    flops = FLOP_PER_CELL_ITER[case.lattice] * head["niter_mean"] * cells
    err = summary[head_name].get("rel_linf_vs_fp64")
    kernel = KERNEL_NAME[(case.lattice, prec)]
    workload = f"{case.name}-{case.lattice}-{'x'.join(map(str, case.shape))}"
    traffic, traffic_src = profiled_traffic(kernel, workload)
    valu_insts, valu_src = profiled("valu", kernel, workload, "valu_wave_insts_per_launch")
    out = {
        "metric": METRIC,
        "value": round(head["mlups"], 4),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["wall_s"] * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if prec == "split" else "f64",
        "data": "synthetic",
        "config": {"workload": f"{case.name}: {case.lattice} {'x'.join(map(str, case.shape))} {interp_for(case)} "
                               f"PonD step, path {head_name}", "lattice": case.lattice, "grid": list(case.shape),
                   "interp": interp_for(case), "dtau": case.dtau, "collect_steps": args.collect,
                   "paths": paths},
        "niter_mean": round(head["niter_mean"], 4),
        "rel_linf_vs_fp64": err,
        # the same error outside the cells that a cell at the iteration cap (which ends its step
        # on an unconverged iterate of a cycle, in the FP64 reference too) can have influenced
        "rel_linf_vs_fp64_outside_cap_dependence": summary[head_name].get("rel_linf_vs_fp64_outside_cap_dependence"),
        "cap_dependence_fraction": (round(summary[head_name]["cells_in_cap_dependence"] / cells, 5)
                                    if "cells_in_cap_dependence" in summary[head_name] else None),
        "speedup_vs_fp64": round(head["mlups"] / ref["mlups"], 3) if ref is not None and head is not ref else None,
        "paths": summary,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic, "bytes_per_lup": bpl,
                     "algorithmic_bytes_per_launch": bpl * cells, "kernel": kernel, "traffic_source": traffic_src},
        "valu_roofline": {"bound": "fp64-valu", "achieved": round(flops / launch_s / 1e12, 4), "peak": FP64_PEAK_TFLOPS,
                          "unit": "TFLOP/s", "frac": round(flops / launch_s / 1e12 / FP64_PEAK_TFLOPS, 5),
                          "flop_per_cell_iter": FLOP_PER_CELL_ITER[case.lattice],
                          "note": "the reference algorithm's FP64 flops (19 per transfer term, de-duplicated "
                                  "transfers) at the path's mean iterations; the split path computes the same "
                                  "result with sum-factorised FP32 transfers, so its figure can exceed the peak"},
    }
    if valu_insts is not None:
        # executed VALU instructions of the kernel (rocprofv3 SQ_INSTS_VALU per launch) over this
        # run's launch time, against the issue rate: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz
        rate = valu_insts * 64 / launch_s / 1e12
        out["valu_issue"] = {"bound": "valu-issue", "achieved": round(rate, 3), "peak": VALU_ISSUE_PEAK,
                             "unit": "T lane-instructions/s", "frac": round(rate / VALU_ISSUE_PEAK, 4),
                             "kernel": kernel, "source": valu_src}
    if not args.no_cpu:
        log("[bench] cpu baseline")
        out["cpu_baseline"] = cpu_baseline(args.config, interp_for(case))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
