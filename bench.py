"""Benchmark of the momentum-loop integrators on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload selfenergy|threshold|finite_T|uneven]

A step is one evaluation of the workload's whole call set (for the default workload, the 64
two-angle self-energy integrals {64 q, 32 cos1, 32 cos2} of the 64 external momenta), with inputs
resident in HBM.  Both precisions run in the same process: FP64 (plain) and mixed (FP32 integrand,
FP64 nodes/denominators/accumulation); `value` is the mixed path's integrals/s, accepted only if
its relative error against the FP64 oracle is <= 1e-8.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_VALU = 78.6e12   # MI355X FP64 vector, FLOP/s
PEAK_FP32_VALU = 157.3e12  # MI355X FP32 vector (packed), FLOP/s
MIXED_TOL = 1e-8


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--workload", default="selfenergy", choices=["selfenergy", "threshold", "finite_T", "uneven"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="bound on the CPU baseline sample")
    return ap.parse_args()


def time_steps(D, fn, steps, warmup):
    """Run fn() warmup times, then time `steps` calls between two HIP events on the library's
    stream of this thread (the stream every launch below is enqueued on)."""
    for _ in range(warmup):
        fn()
    D.stream_sync()
    e0, e1 = D.Event(), D.Event()
    t0 = time.perf_counter()
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    e1.synchronize()
    wall = time.perf_counter() - t0
    return e0.elapsed_ms(e1) / steps, wall


# Floating-point operations per integrand evaluation of the self-energy kernels, counted from
# the gfx950 ISA of the inner loop ("FLOP accounting").
FLOP_PER_EVAL = {"selfenergy": {0: None, 1: None}}


def run_selfenergy(D, args, out_json):
    import diffrg_amd.workloads as W
    c = W.selfenergy()
    qp = D.QuadratureProvider()
    n = len(c["p"])
    b = D.SEBatch(qp, c["p"], c["nq"], c["na"], [c["Z"]] * n, *[[v] * n for v in c["coord"]],
                  x_extent=c["x_extent"], k=c["k"])
    evals = b.evals
    out = D.DeviceArray(n)
    res = {}
    for mode, name in ((D.FP64, "fp64"), (D.MIXED, "mixed")):
        tune = b.autotune(mode, reps=5)
        ms, wall = time_steps(D, lambda: b.run(out, mode), args.steps, args.warmup)
        vals = out.to_host()
        res[name] = {"ms": ms, "wall_s": wall, "shape": tune, "out": vals}
    # parity of the mixed path against the FP64 oracle on the same nodes (and the FP64 GPU path)
    from oracle import oracle as O
    xq = (qp.get_points(c["nq"]), qp.get_weights(c["nq"]))
    xa = (qp.get_points(c["na"]), qp.get_weights(c["na"]))
    coord = O.logcoord(c["n_tab"], *c["coord"])
    ref = O.c1_selfenergy(c["p"], [c["Z"]] * n, [coord] * n, [xq] * n, [xa] * n, c["x_extent"], c["k"])
    err_fp64 = float(np.max(np.abs(res["fp64"]["out"] - ref) / np.abs(ref)))
    err_mixed = float(np.max(np.abs(res["mixed"]["out"] - ref) / np.abs(ref)))
    ms_mixed, ms_fp64 = res["mixed"]["ms"], res["fp64"]["ms"]
    out_json.update({
        "metric": "integrals/sec at rel.err<=1e-8 vs FP64 (mixed vs plain)",
        "value": n / (ms_mixed * 1e-3) if err_mixed <= MIXED_TOL else 0.0,
        "unit": "integrals/s",
        "ms_per_step": ms_mixed,
        "dtype": "fp32-eval/fp64-acc",
        "config": {"workload": "selfenergy: 64 p on log grid [1e-3,40], 2-angle {64,32,32}, 128-pt Z table",
                   "integrals_per_step": n, "evals_per_step": evals},
        "plain_fp64": {"integrals_per_s": n / (ms_fp64 * 1e-3), "ms_per_step": ms_fp64,
                       "evals_per_s": evals / (ms_fp64 * 1e-3), "max_rel_err_vs_oracle": err_fp64,
                       "shape": res["fp64"]["shape"]},
        "mixed": {"integrals_per_s": n / (ms_mixed * 1e-3), "evals_per_s": evals / (ms_mixed * 1e-3),
                  "max_rel_err_vs_oracle": err_mixed, "shape": res["mixed"]["shape"]},
        "mixed_over_plain": ms_fp64 / ms_mixed,
    })
    return {"evals": evals, "n": n, "ms_mixed": ms_mixed, "ms_fp64": ms_fp64, "ref_fn": lambda: O.c1_selfenergy(
        c["p"], [c["Z"]] * n, [coord] * n, [xq] * n, [xa] * n, c["x_extent"], c["k"])}


def cpu_baseline(info, seconds):
    from oracle import oracle as O
    fn = info["ref_fn"]
    fn()  # warm-up
    times = []
    t_start = time.perf_counter()
    while len(times) < 5 or (time.perf_counter() - t_start < seconds and len(times) < 50):
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
        if time.perf_counter() - t_start > 4 * seconds:
            break
    med = statistics.median(times)
    cpu_model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": info["n"] / med, "unit": "integrals/s", "cores": O.num_threads(), "kind": "port",
            "sample": f"{len(times)} runs of the full call set ({info['n']} integrals, {info['evals']} evals) "
                      f"on the C restatement of the reference TBB loop, median {med * 1e3:.1f} ms",
            "evals_per_s": info["evals"] / med, "cpu_model": cpu_model}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU sharding is described in "
                                   "DESIGN.md and not implemented", "n_gpus": args.gpus}))
        sys.exit(2)
    import diffrg_amd as D
    D.init(0)
    out = {"n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True, "scaling": "strong",
           "vs_baseline": None, "data": "synthetic (SplitMix64 seed 0xD1FF0000+c)"}
    if args.workload != "selfenergy":
        raise SystemExit("only the selfenergy workload is wired into bench.py yet")
    info = run_selfenergy(D, args, out)
    fpe = FLOP_PER_EVAL["selfenergy"][1]
    achieved = (fpe * info["evals"] / (info["ms_mixed"] * 1e-3)) if fpe else None
    out["roofline"] = {"bound": "valu", "achieved": achieved, "peak": PEAK_FP64_VALU, "unit": "FLOP/s",
                       "frac": (achieved / PEAK_FP64_VALU) if achieved else None, "traffic": None,
                       "kernel": "se_kernel<MIXED>"}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(info, args.cpu_seconds)
    out["device"] = D.device_props()["name"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
