"""Benchmark of the momentum-loop integrators on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--only-headline] [--no-cpu-baseline]

Headline (the JSON line's metric/value): the two-angle self-energy call set -- 64 external
momenta on a log grid in [1e-3, 40], nodes {64 q^2, 32 cos1, 32 cos2}, 128-point dressing table --
evaluated in mixed precision (FP32 hardware seeds, FP64 nodes/denominators/corrections/
accumulation), accepted only when its relative error against the FP64 oracle is <= 1e-8.  A step
is one evaluation of the whole call set with inputs resident in HBM; K steps are timed between
two HIP events on the stream the kernels run on.  The plain FP64 path is timed in the same run.

The same run also measures the other workloads (reported under "workloads"): the same call set
through the drop-in template API (Integrator4D2AngGPU + request_data, and as 64 get() calls), LPA
threshold integrals and finite-T integrals (launched one by one and from a hipGraph of 10 call
sets), the uneven 4096-integral persistent batch (hipGraph-captured) and the flow server under
Poisson arrivals (p50/p99 latency), each checked against the oracle.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X published FP64 vector peak (FLOP/s).  Packed FP32 is disabled in this build, so plain FP32
# and FP64 VALU instructions issue at the same rate and this peak bounds both precisions.
PEAK_FP64_VALU = 78.6e12
# one wave64 VALU instruction per clock per CU (256 CUs x 4 SIMD16 x 2.4 GHz / 4 clocks)
PEAK_VALU_WAVE_INSTR = 256 * 2.4e9
MIXED_TOL = 1e-8
GRAPH_SETS = 10  # call sets per hipGraph launch

# (FP operations with FMA = 2, VALU instructions) per integrand evaluation in the innermost
# loops of se_kernel<MODE, BLOCK, UNROLL>, counted from the gfx950 assembly of
# csrc/integrals.hip by `python -m diffrg_amd.isa_count`: "pairs" when a lane owns two angular
# pairs (na^2 >= 2 BLOCK), "pairs2" the same loop taking two radial nodes per step (the mixed
# path's unroll 2, SEBatch.unroll), "single" when a lane owns one pair and two radial nodes per
# iteration.  tests/test_isa_counts.py checks these figures against the current source.
# Keyed by the integrand form (SEBatch.variants): 0 FP64, 1 mixed, 2 mixed on an integral that
# admits the narrow exponent rounding.
SE_OPS = {0: {"pairs": (223.5, 203.0), "single": (223.5, 205.0)},
          1: {"pairs": (46.0, 44.0), "pairs2": (46.0, 43.5), "single": (46.0, 46.0)},
          2: {"pairs": (46.0, 42.0), "pairs2": (46.0, 41.5), "single": (46.0, 44.0)}}
SE_OPS_BLOCK = {}  # (mode, block, loop) whose schedule differs from the others


BASELINE_NOTES = "BASELINE.md"


def frozen_weights():
    """FLOP (and VALU) per evaluation, measured once from the loops' assembly and frozen in the
    repository's baseline notes (their ```json flop-weights block)."""
    text = open(os.path.join(ROOT, BASELINE_NOTES)).read()
    start = text.index("```json flop-weights") + len("```json flop-weights")
    return json.loads(text[start:text.index("```", start)])


FROZEN = frozen_weights()


def ceiling(w):
    """Instruction-mix bound of a loop: FLOP / (2 VALU) per evaluation."""
    return w["flop"] / (2.0 * w["valu"])


def se_loop(na, block, unroll=1):
    if 2 * min(na * na, block) > na * na:
        return "single"
    return "pairs2" if unroll == 2 else "pairs"


def se_ops(variant, na, block, unroll=1):
    """(FLOP, VALU) per evaluation of the loop se_kernel<., block, unroll> runs with integrand
    form `variant` for na angular nodes."""
    loop = se_loop(na, block, unroll)
    return SE_OPS_BLOCK.get((variant, block, loop), SE_OPS[variant][loop])


def se_variant(b, mode):
    """The integrand form of a batch whose integrals all use the same one."""
    v = np.unique(b.variants(mode))
    assert len(v) == 1, v
    return int(v[0])


def valu_issue_frac(valu_per_eval, evals, seconds):
    """Fraction of the chip's VALU issue slots the innermost loop alone fills at this rate."""
    return valu_per_eval * evals / 64 / seconds / PEAK_VALU_WAVE_INSTR


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=1000,
                    help="untimed launches before each timed run (1000: the clocks settle after the autotuner)")
    ap.add_argument("--only-headline", action="store_true", help="skip the other workloads")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="bound on the CPU baseline sample")
    ap.add_argument("--flows", type=int, default=1000, help="flows in the steady-state serving run")
    ap.add_argument("--flow-check-frac", type=float, default=0.1,
                    help="fraction of the served flows checked against the CPU reference (stratified over the "
                         "grid sizes, one flow per CPU thread); 0: six flows spanning the grid sizes")
    ap.add_argument("--workloads", default="template,threshold,finite_T,uneven,flows",
                    help="comma-separated workloads measured after the headline")
    ap.add_argument("--shape", action="append", default=[], metavar="MODE:BLOCK,TILE,GRID",
                    help="fixed launch shape instead of the autotuner for the headline (MODE fp64|mixed); "
                         "for counter runs, where serialised dispatches distort the tuner's timings")
    ap.add_argument("--eager-only", action="store_true",
                    help="headline: time eager launches only (no hipGraph, multi-stream or batched runs), so a "
                         "kernel trace ends with exactly the timed launches")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device copy (calibrates the HBM byte counters)")
    return ap.parse_args()


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "se_kernel_selfenergy_traffic.txt")


def measured_traffic():
    """HBM bytes per se_kernel<MIXED> launch from the committed counter summary, or None."""
    try:
        for line in open(TRAFFIC_FILE):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def copy_calibration():
    """Copies a 256 MiB device buffer 5 times with dfrg_memcpy_d2d: 256 MiB read + 256 MiB written."""
    import diffrg_amd as D
    D.init(0)
    n = 1 << 25
    src = D.DeviceArray.from_host(np.ones(n))
    dst = D.DeviceArray(n)
    for _ in range(5):
        D.memcpy_d2d(dst, src)
    D.stream_sync()
    print(json.dumps({"copy_bytes_each_way": n * 8, "copies": 5}))


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def time_steps(D, fn, steps, warmup):
    """Warm up, then time `steps` calls of fn() between two HIP events on this thread's stream."""
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


def time_streams(D, fns, steps, warmup):
    """Throughput of independent call sets issued round-robin on len(fns) streams: fns[s]() enqueues
    one call set on stream s (with its own output and scratch).  Host clock between two full
    synchronisations; returns ms per call set."""
    for i in range(warmup):
        fns[i % len(fns)][0]()
    for _, st in fns:
        st.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        fns[i % len(fns)][0]()
    for _, st in fns:
        st.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


N_STREAMS = 4
BATCH = 8  # call sets per launch in the batched measurements of the small workloads


def eager_and_graph(D, fn, args):
    """ms per call of fn() launched eagerly from Python, and replayed from a hipGraph of GRAPH_SETS
    captured calls (no host cost per call: a kernel of a few microseconds launched from Python is
    otherwise timed at the host's launch rate)."""
    ms_e, _ = time_steps(D, fn, args.steps, args.warmup)
    g = D.Graph(fn, times=GRAPH_SETS)
    ms_g, _ = time_steps(D, g.launch, max(1, args.steps // GRAPH_SETS), max(1, args.warmup // GRAPH_SETS))
    return ms_e, ms_g / GRAPH_SETS


def ft_best_block(D, run):
    """Threads per integral of ft_batch with the shortest device time on this batch (hipGraph of
    GRAPH_SETS calls per candidate, so the host's launch rate does not decide)."""
    best = None
    for block in (64, 128, 256, 512):
        D.ft_set_block(block)
        g = D.Graph(run, times=GRAPH_SETS)
        t, _ = time_steps(D, g.launch, 20, 5)
        if best is None or t < best[0]:
            best = (t, block)
    return best[1]


def finite(v):
    return v if math.isfinite(v) and v != 0.0 else None


def rel_err(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b)) / np.abs(np.asarray(b))))


def cpu_timed(fn, seconds, min_runs=3, max_runs=50):
    fn()
    times = []
    t_start = time.perf_counter()
    while len(times) < min_runs or (time.perf_counter() - t_start < seconds and len(times) < max_runs):
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
        if time.perf_counter() - t_start > 4 * seconds:
            break
    return statistics.median(times), len(times)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return ""


# --------------------------------------------------------------------------------- workloads

def selfenergy(D, qp, args):
    import diffrg_amd.workloads as W
    from oracle import oracle as O
    c = W.selfenergy()
    n = len(c["p"])
    b = D.SEBatch(qp, c["p"], c["nq"], c["na"], [c["Z"]] * n, *[[v] * n for v in c["coord"]],
                  x_extent=c["x_extent"], k=c["k"])
    out = D.DeviceArray(n)
    xq = (qp.get_points(c["nq"]), qp.get_weights(c["nq"]))
    xa = (qp.get_points(c["na"]), qp.get_weights(c["na"]))
    coord = O.logcoord(c["n_tab"], *c["coord"])
    ref_fn = lambda: O.c1_selfenergy(c["p"], [c["Z"]] * n, [coord] * n, [xq] * n, [xa] * n, c["x_extent"], c["k"])
    ref = ref_fn()
    res = {}
    ms_objs = []
    for _ in range(0 if args.eager_only else N_STREAMS):
        ms_objs.append((D.SEBatch(qp, c["p"], c["nq"], c["na"], [c["Z"]] * n, *[[v] * n for v in c["coord"]],
                                  x_extent=c["x_extent"], k=c["k"]), D.DeviceArray(n), D.Stream()))
    fixed = {m.split(":")[0]: [int(v) for v in m.split(":")[1].split(",")] for m in args.shape}
    for mode, name in ((D.FP64, "fp64"), (D.MIXED, "mixed")):
        if name in fixed:
            b.set_shape(*fixed[name])
            shape = b.shape()
        else:
            shape = b.autotune(mode, reps=5)
        ms_eager, _ = time_steps(D, lambda: b.run(out, mode), args.steps, args.warmup)
        err_eager = rel_err(out.to_host(), ref)
        err = err_eager
        ms_graph = ms_streams = float("inf")
        if not args.eager_only:
            # the same call sets replayed from a hipGraph of GRAPH_SETS launches
            per = GRAPH_SETS
            ms_graph, _ = time_steps(D, lambda: b.graph_launch(out, mode, launches=per), max(1, args.steps // per),
                                     max(1, args.warmup // per))
            ms_graph /= per
            err = max(err, rel_err(out.to_host(), ref))
            # independent call sets on N_STREAMS streams, one batch object (own scratch) per stream
            fns = [(lambda bb=bb, o=o, st=st: bb.run(o, mode, stream=st.ptr), st) for bb, o, st in ms_objs]
            for bb, _, _ in ms_objs:
                bb.set_shape(*[shape[k] for k in ("block", "tile_evals", "grid")])
            ms_streams = time_streams(D, fns, args.steps, args.warmup)
            for _, o, _ in ms_objs:
                err = max(err, rel_err(o.to_host(), ref))
        ms = min(ms_eager, ms_graph)
        res[name] = {"ms": ms, "ms_eager": ms_eager, "ms_graph": ms_graph, "ms_streams": ms_streams, "shape": shape,
                     "err": err, "launch": "hipGraph" if ms_graph < ms_eager else "eager", "unroll": b.unroll(mode)}
        log(f"selfenergy {name}: eager {ms_eager * 1e3:.1f} us, graph {ms_graph * 1e3:.1f} us, "
            f"{N_STREAMS} streams {ms_streams * 1e3:.1f} us per call set, "
            f"err {err:.2e}, shape {shape}")
    if args.eager_only:
        return {"n": n, "evals": b.evals, "na": c["na"], "res": res, "ref_fn": ref_fn,
                "variant": {D.FP64: se_variant(b, D.FP64), D.MIXED: se_variant(b, D.MIXED)}}
    # BATCH call sets in one launch (the 64 integrals repeated): the kernel's throughput on C1's
    # integrand without the fixed per-launch cost
    nb = BATCH * n
    bb = D.SEBatch(qp, np.tile(c["p"], BATCH), c["nq"], c["na"], [c["Z"]] * nb, *[[v] * nb for v in c["coord"]],
                   x_extent=c["x_extent"], k=c["k"])
    outb = D.DeviceArray(nb)
    refb = np.tile(ref, BATCH)
    for mode, name in ((D.FP64, "fp64"), (D.MIXED, "mixed")):
        shape = bb.autotune(mode, reps=3)
        ms, _ = time_steps(D, lambda: bb.run(outb, mode), max(1, args.steps // BATCH), max(1, args.warmup // BATCH))
        va = se_ops(se_variant(bb, mode), c["na"], shape["block"], bb.unroll(mode))[1]
        fl = FROZEN["C1_selfenergy"][name]["flop"]
        res[name][f"batched_x{BATCH}"] = {
            "integrals_per_s": nb / (ms * 1e-3), "us_per_launch": ms * 1e3, "shape": shape,
            "max_rel_err": rel_err(outb.to_host(), refb),
            "roofline_frac": fl * bb.evals / (ms * 1e-3) / PEAK_FP64_VALU,
            "valu_issue_frac": valu_issue_frac(va, bb.evals, ms * 1e-3)}
        log(f"selfenergy {name} x{BATCH}: {ms * 1e3:.1f} us per launch, shape {shape}")
    return {"n": n, "evals": b.evals, "na": c["na"], "res": res, "ref_fn": ref_fn,
            "variant": {D.FP64: se_variant(b, D.FP64), D.MIXED: se_variant(b, D.MIXED)}}


def threshold(D, qp, args):
    import diffrg_amd.workloads as W
    from oracle import oracle as O
    c = W.threshold()
    x, w = qp.get_points(64), qp.get_weights(64)
    ref_fn = lambda: O.c0_threshold(c["m2"], x, w, c["x_extent"], c["k"], 1.0, c["d"])
    ref = ref_fn()
    d_m2, d_x, d_w = (D.DeviceArray.from_host(a) for a in (c["m2"], x, w))
    out = D.DeviceArray(len(c["m2"]))
    r = {}
    for mode, name in ((D.FP64, "fp64"), (D.MIXED, "mixed")):
        run = lambda: D.threshold_batch(d_m2, d_x, d_w, out, c["x_extent"], c["k"], mode, c["d"])
        ms_e, ms_g = eager_and_graph(D, run, args)
        ms = min(ms_e, ms_g)
        wt = FROZEN["C0_threshold"][name]
        r[name] = {"integrals_per_s": len(c["m2"]) / (ms * 1e-3), "us_per_step": ms * 1e3,
                   "us_eager": ms_e * 1e3, "us_graph": ms_g * 1e3, "launch": "hipGraph" if ms_g < ms_e else "eager",
                   "max_rel_err": rel_err(out.to_host(), ref),
                   "roofline_frac": wt["flop"] * len(c["m2"]) * 64 / (ms * 1e-3) / PEAK_FP64_VALU,
                   "ceiling": ceiling(wt)}
    # BATCH call sets in one launch (the same 256 integrals repeated): throughput without the fixed
    # per-launch latency that dominates a 256-integral launch
    d_m2b = D.DeviceArray.from_host(np.tile(c["m2"], BATCH))
    outb = D.DeviceArray(BATCH * len(c["m2"]))
    for mode, name in ((D.FP64, "fp64"), (D.MIXED, "mixed")):
        ms = min(eager_and_graph(D, lambda: D.threshold_batch(d_m2b, d_x, d_w, outb, c["x_extent"], c["k"], mode,
                                                              c["d"]), args))
        r[name][f"integrals_per_s_batched_x{BATCH}"] = BATCH * len(c["m2"]) / (ms * 1e-3)
    t_cpu, runs = cpu_timed(ref_fn, 2.0)
    r["cpu_integrals_per_s"] = len(c["m2"]) / t_cpu
    r["config"] = "256 m2 ~ U[-0.8,4], 64 GL nodes in q^2 over [0,20], d=4"
    return r


def finite_T(D, qp, args):
    import diffrg_amd.workloads as W
    from oracle import oracle as O
    c = W.finite_T()
    x, w = qp.get_points(c["nx"]), qp.get_weights(c["nx"])
    temps = sorted(c["rules"])
    starts = np.concatenate([[0], np.cumsum([len(c["rules"][t][0]) for t in temps])])
    off = {t: int(starts[i]) for i, t in enumerate(temps)}
    Dh = D.DeviceArray.from_host
    dev = {"T": Dh(c["T"]), "p0": Dh(c["p0"]), "pabs": Dh(c["pabs"]), "m1": Dh(c["m1"]), "m2": Dh(c["m2"]),
           "mx": Dh(np.concatenate([c["rules"][t][0] for t in temps])),
           "mw": Dh(np.concatenate([c["rules"][t][1] for t in temps])),
           "m_off": Dh(np.array([off[t] for t in c["T"]], np.int32)),
           "m_n": Dh(np.array([len(c["rules"][t][0]) for t in c["T"]], np.int32)), "x": Dh(x), "w": Dh(w)}
    mats = [c["rules"][t] for t in c["T"]]
    ref_fn = lambda: O.c2_finiteT(c["T"], c["p0"], c["pabs"], c["m1"], c["m2"], mats, x, w, c["q_max"], c["k"], 1.0,
                                  c["d"], c["radial_mode"])
    ref = ref_fn()
    out = D.DeviceArray(len(c["T"]))
    evals = int(sum(c["nx"] * (2 * len(c["rules"][t][0]) + 1) for t in c["T"]))
    r = {"evals_per_step": evals}
    for mode, name in ((D.FP64, "fp64"), (D.MIXED, "mixed")):
        run = lambda: D.ft_batch(dev, out, c["q_max"], c["k"], mode, c["d"], c["radial_mode"])
        best = ft_best_block(D, run)
        D.ft_set_block(best)
        ms_e, ms_g = eager_and_graph(D, run, args)
        ms = min(ms_e, ms_g)
        got = out.to_host()
        err = np.abs(got - ref) / np.abs(ref)
        wt = FROZEN["C2_finite_T"][name]
        pairs = int(sum(c["nx"] * len(c["rules"][t][0]) for t in c["T"]))
        r[name] = {"integrals_per_s": len(c["T"]) / (ms * 1e-3), "evals_per_s": evals / (ms * 1e-3),
                   "us_per_step": ms * 1e3, "us_eager": ms_e * 1e3, "us_graph": ms_g * 1e3,
                   "launch": "hipGraph" if ms_g < ms_e else "eager", "max_rel_err": float(err.max()), "block": best,
                   "roofline_frac": wt["flop"] * pairs / (ms * 1e-3) / PEAK_FP64_VALU, "ceiling": ceiling(wt),
                   "valu_issue_frac": valu_issue_frac(wt["valu"], pairs, ms * 1e-3),
                   "max_rel_err_per_T": {str(t): float(err[c["T"] == t].max()) for t in W.TEMPS}}
    # BATCH call sets in one launch (the same 512 externals repeated)
    devb = dict(dev)
    for key in ("T", "p0", "pabs", "m1", "m2"):
        devb[key] = Dh(np.tile(c[key], BATCH))
    devb["m_off"] = Dh(np.tile(np.array([off[t] for t in c["T"]], np.int32), BATCH))
    devb["m_n"] = Dh(np.tile(np.array([len(c["rules"][t][0]) for t in c["T"]], np.int32), BATCH))
    outb = D.DeviceArray(BATCH * len(c["T"]))
    for mode, name in ((D.FP64, "fp64"), (D.MIXED, "mixed")):
        run = lambda: D.ft_batch(devb, outb, c["q_max"], c["k"], mode, c["d"], c["radial_mode"])
        best = ft_best_block(D, run)
        D.ft_set_block(best)
        ms = min(eager_and_graph(D, run, args))
        r[name][f"integrals_per_s_batched_x{BATCH}"] = BATCH * len(c["T"]) / (ms * 1e-3)
        r[name][f"block_batched_x{BATCH}"] = best
        if name == "mixed":
            got = outb.to_host()
            r[name][f"max_rel_err_batched_x{BATCH}"] = float(np.max(np.abs(got - np.tile(ref, BATCH)) /
                                                                     np.abs(np.tile(ref, BATCH))))
    D.ft_set_block(0)
    r["mixed_over_plain"] = r["mixed"]["integrals_per_s"] / r["fp64"]["integrals_per_s"]
    r[f"mixed_over_plain_batched_x{BATCH}"] = (r["mixed"][f"integrals_per_s_batched_x{BATCH}"] /
                                              r["fp64"][f"integrals_per_s_batched_x{BATCH}"])
    t_cpu, _ = cpu_timed(ref_fn, 2.0)
    r["cpu_integrals_per_s"] = len(c["T"]) / t_cpu
    r["config"] = "512 externals, T in {0.01..0.3}, explicit Matsubara n in [-64,64], 48 GL in |q| over [0,6]"
    return r


def uneven(D, qp, args):
    import diffrg_amd.workloads as W
    from oracle import oracle as O
    c = W.uneven()
    n = len(c["p"])
    b = D.SEBatch(qp, c["p"], c["nq"], c["na"], c["tables"], *[[v] * n for v in c["coord"]],
                  x_extent=c["x_extent"], k=c["k"])
    xq = [(qp.get_points(int(v)), qp.get_weights(int(v))) for v in c["nq"]]
    xa = [(qp.get_points(int(v)), qp.get_weights(int(v))) for v in c["na"]]
    coords = [O.logcoord(len(t), *c["coord"]) for t in c["tables"]]
    ref_fn = lambda: O.c1_selfenergy(c["p"], c["tables"], coords, xq, xa, c["x_extent"], c["k"])
    t0 = time.perf_counter()
    ref = ref_fn()
    t_cpu = time.perf_counter() - t0
    out = D.DeviceArray(n)
    r = {"evals_per_step": b.evals}
    per_graph = 10
    steps = max(1, args.steps // (4 * per_graph))
    for mode, name in ((D.FP64, "fp64"), (D.MIXED, "mixed")):
        shape = b.autotune(mode, reps=2, cache_path=None)
        ms, _ = time_steps(D, lambda: b.graph_launch(out, mode, launches=per_graph), steps, 2)
        ms /= per_graph
        wt = FROZEN["C3_uneven"][name]
        r[name] = {"integrals_per_s": n / (ms * 1e-3), "evals_per_s": b.evals / (ms * 1e-3), "ms_per_batch": ms,
                   "max_rel_err": rel_err(out.to_host(), ref), "shape": shape,
                   "roofline_frac": wt["flop"] * b.evals / (ms * 1e-3) / PEAK_FP64_VALU, "ceiling": ceiling(wt),
                   "valu_issue_frac": sum(int(q) * int(a) ** 2 * se_ops(int(v), int(a), shape["block"],
                                                                          b.unroll(mode))[1]
                                          for q, a, v in zip(c["nq"], c["na"], b.variants(mode)))
                                      / 64 / (ms * 1e-3)
                                      / PEAK_VALU_WAVE_INSTR}
        log(f"uneven {name}: {ms:.3f} ms/batch, err {r[name]['max_rel_err']:.2e}, shape {shape}")
    r["mixed_over_plain"] = r["mixed"]["integrals_per_s"] / r["fp64"]["integrals_per_s"]
    r["cpu_integrals_per_s"] = n / t_cpu
    r["config"] = ("4096 two-angle integrals, radial order in {16,32,64,128}, angular in {8,16,32,64}, own 32-256 "
                   "point Z table and p; one persistent launch per batch, hipGraph of 10 batches")
    return r


KAT_LIB = os.path.join(ROOT, "tests", "native", "libdfrg_kat.so")


def template(D, qp, args):
    """C1 through the drop-in boundary: Integrator4D2AngGPU<NT, KERNEL> with the self-energy as a
    user KERNEL (LinearInterpolator1D dressing), its 64 requests made as request_data makes them and
    read back.  The C++ side is the template test library (tests/native/libdfrg_kat.so, a user of
    the drop-in headers).  Device time: HIP events on the request thread's stream around the batch's
    launch and the reads (the kernel reads the tuples from pinned host memory and stores the results
    there); host time: the whole call set (64 requests, launch, reads).  Also timed: the same 64
    integrals as 64 get() calls, one launch and one synchronisation each, the reference's per-call
    pattern."""
    import ctypes as C
    import diffrg_amd.workloads as W
    from oracle import oracle as O
    kat = C.CDLL(KAT_LIB)
    kat.kat_last_error.restype = C.c_char_p
    c = W.selfenergy()
    Z = np.ascontiguousarray(c["Z"])
    n = len(c["p"])
    evals = n * c["nq"] * c["na"] ** 2
    P = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    coord = O.logcoord(c["n_tab"], *c["coord"])
    xq = (qp.get_points(c["nq"]), qp.get_weights(c["nq"]))
    xa = (qp.get_points(c["na"]), qp.get_weights(c["na"]))
    r = {"evals_per_call_set": evals,
         "path": "Integrator4D2AngGPU<NT, SelfEnergyKernel> + request_data/update_data (physics/utils.hh)"}
    for fp32, name, tol in ((0, "fp64", 1e-12), (1, "fp32_eval_fp64_acc", 1e-5)):
        out, pg = np.empty(n), np.empty(n)
        td, th, ph = C.c_double(), C.c_double(), np.empty(5)
        rc = kat.kat_selfenergy_request_bench(c["nq"], c["na"], C.c_double(c["x_extent"]), C.c_double(c["k"]), n,
                                              P(Z), len(Z), C.c_double(c["coord"][0]), C.c_double(c["coord"][1]),
                                              C.c_double(c["coord"][2]), fp32, 20, P(out), P(pg), C.byref(td),
                                              C.byref(th), P(ph))
        if rc != 0:
            raise RuntimeError(kat.kat_last_error())
        ref = O.c1_selfenergy(pg, [c["Z"]] * n, [coord] * n, [xq] * n, [xa] * n, c["x_extent"], c["k"])
        err = rel_err(out, ref)
        w = FROZEN["C1_template"]["fp64" if not fp32 else "fp32"]
        t = td.value
        r[name] = {"integrals_per_s": n / t, "evals_per_s": evals / t, "us_device_per_call_set": t * 1e6,
                   "us_host_per_call_set": th.value * 1e6, "max_rel_err_vs_fp64_oracle": err,
                   "within_tolerance": err <= tol, "tolerance": tol,
                   "flop_per_eval": w["flop"], "valu_per_eval": w["valu"], "ceiling": ceiling(w),
                   "roofline_frac": w["flop"] * evals / t / PEAK_FP64_VALU,
                   "valu_issue_frac": valu_issue_frac(w["valu"], evals, t),
                   "us_phases": {"host_requests": ph[0] * 1e6, "host_launch": ph[1] * 1e6, "host_reads": ph[2] * 1e6,
                                 "device_launch_to_enqueued": ph[3] * 1e6, "device_enqueued_to_read": ph[4] * 1e6}}
        log(f"template {name}: {t * 1e6:.1f} us device, {th.value * 1e6:.1f} us host per call set, err {err:.2e}")
    # the same integrals one get() at a time (FP64)
    out, tg = np.empty(n), C.c_double()
    pv = np.ascontiguousarray(pg, dtype=np.float64)
    rc = kat.kat_selfenergy_get_bench(c["nq"], c["na"], C.c_double(c["x_extent"]), C.c_double(c["k"]), P(pv), n,
                                      P(Z), len(Z), C.c_double(c["coord"][0]), C.c_double(c["coord"][1]),
                                      C.c_double(c["coord"][2]), 5, P(out), C.byref(tg))
    if rc != 0:
        raise RuntimeError(kat.kat_last_error())
    ref = O.c1_selfenergy(pv, [c["Z"]] * n, [coord] * n, [xq] * n, [xa] * n, c["x_extent"], c["k"])
    r["fp64_get_per_integral"] = {"us_per_get": tg.value * 1e6, "integrals_per_s": 1.0 / tg.value,
                                  "max_rel_err_vs_fp64_oracle": rel_err(out, ref),
                                  "pattern": "64 get() calls, each one launch and one synchronisation"}
    log(f"template fp64 get(): {tg.value * 1e6:.1f} us per integral")
    return r


def _flow_sample(reqs, m, seed):
    """m flows spanning the grid sizes: the sorted sizes cut into m strata, one seeded pick in each."""
    rng = np.random.default_rng(seed)
    order = sorted(range(len(reqs)), key=lambda i: reqs[i]["n_grid"])
    edges = np.linspace(0, len(order), m + 1).astype(int)
    return sorted(order[int(rng.integers(edges[j], max(edges[j] + 1, edges[j + 1])))] for j in range(m))


def flow_step_roofline(tm, nx=64):
    """flow_step_kernel over the steady-state run, from the server's HIP events around each graph
    launch (25 step launches and the base-step node per graph, so the gaps between the launches
    inside a graph count as kernel time): frozen FLOP per node of both threshold integrals x 64
    nodes x 4 RK stages per grid-point step."""
    wt = FROZEN["C4_flow_node"]["fp64"]
    node_evals = tm["point_steps"] * 4 * nx
    achieved = wt["flop"] * node_evals / tm["device_s"] if tm["device_s"] > 0 else 0.0
    return {"graph_launches": tm["launches"], "device_s": tm["device_s"], "point_steps": tm["point_steps"],
            "node_evals": node_evals, "flop_per_node": wt["flop"], "achieved_tflops": achieved / 1e12,
            "roofline_frac": achieved / PEAK_FP64_VALU, "ceiling": ceiling(wt),
            "valu_issue_frac": valu_issue_frac(wt["valu"], node_evals, tm["device_s"]),
            "timing": "HIP events around each graph launch on the server's stream"}


def flows(D, qp, args):
    """Flow serving: saturation throughput (flows submitted at once), then a steady-state run of
    Poisson arrivals at 0.7x saturation, with served/offered rate and p50/p99 latency over the
    window that excludes the first and last 10% of arrivals, and final potentials of a seeded
    sample of flows spanning the grid sizes against the CPU reference."""
    import diffrg_amd.workloads as W
    from oracle import oracle as O
    budget = 64 << 20
    # saturation: n_sat flows submitted at t = 0
    n_sat = min(args.flows, 512)
    fs = D.FlowServer(qp, hbm_budget=budget, max_resident=128, steps_per_launch=25)
    reqs = W.flows(n_sat, rate=1.0)
    t0 = time.perf_counter()
    for q in reqs:
        fs.submit(q["n_grid"], q["T"], q["mu"], q["rho_max"], q["N"], q["m2_0"], q["lambda"])
    done = 0
    while done < n_sat:
        _, _, done = fs.progress()
    sat = n_sat / (time.perf_counter() - t0)
    fs.close()
    # steady state: Poisson arrivals at 0.7 x saturation
    n_flows = args.flows
    rate = 0.7 * sat
    reqs = W.flows(n_flows, rate=rate)
    fs = D.FlowServer(qp, hbm_budget=budget, max_resident=128, steps_per_launch=25)
    pending = list(range(n_flows))
    ids = [None] * n_flows
    t0 = time.perf_counter()
    finished = 0
    while finished < n_flows:
        now = time.perf_counter() - t0
        while pending and reqs[pending[0]]["arrival"] <= now:
            i = pending.pop(0)
            q = reqs[i]
            ids[i] = fs.submit(q["n_grid"], q["T"], q["mu"], q["rho_max"], q["N"], q["m2_0"], q["lambda"])
        r, qd, finished = fs.progress()
        if r == 0 and qd == 0 and pending:
            time.sleep(max(0.0, min(0.002, reqs[pending[0]]["arrival"] - (time.perf_counter() - t0))))
    elapsed = time.perf_counter() - t0
    tm = fs.timing()
    res = [fs.result(fid) for fid in ids]
    lat = np.array([l for _, l in res])
    arr = np.array([q["arrival"] for q in reqs])
    # window: arrivals between the 10% and 90% quantiles; a flow completes at arrival + latency
    lo, hi = np.quantile(arr, 0.1), np.quantile(arr, 0.9)
    in_win = (arr >= lo) & (arr <= hi)
    comp = arr + lat
    offered = in_win.sum() / (hi - lo)
    served = ((comp >= lo) & (comp <= hi)).sum() / (hi - lo)
    # CPU reference for a seeded sample spanning the grid sizes
    x, w = qp.get_points(64), qp.get_weights(64)
    m = max(4, int(math.ceil(args.flow_check_frac * n_flows))) if args.flow_check_frac > 0 else 6
    sample = _flow_sample(reqs, m, seed=0xD1FF0004)
    fps = [O.flow_params(reqs[i]["n_grid"], reqs[i]["rho_max"], reqs[i]["T"], reqs[i]["mu"], reqs[i]["N"],
                         reqs[i]["m2_0"], reqs[i]["lambda"]) for i in sample]
    t1 = time.perf_counter()
    if args.flow_check_frac > 0:
        urefs = O.flow_lpa_many(fps, x, w)  # one flow per thread
    else:
        urefs = [O.flow_lpa(fp, x, w) for fp in fps]  # each flow's right-hand side over the threads
    cpu_s = time.perf_counter() - t1
    diffs = [float(np.max(np.abs(res[i][0] - u))) for i, u in zip(sample, urefs)]
    fs.close()
    return {"flows": n_flows, "saturation_flows_per_s": sat, "saturation_flows": n_sat,
            "offered_flows_per_s": rate, "run_s": elapsed,
            "steady_state": {"window_s": [float(lo), float(hi)], "flows_in_window": int(in_win.sum()),
                             "offered_flows_per_s": float(offered), "served_flows_per_s": float(served),
                             "served_over_offered": float(served / offered),
                             "p50_latency_s": float(np.quantile(lat[in_win], 0.5)),
                             "p99_latency_s": float(np.quantile(lat[in_win], 0.99))},
            "p50_latency_s_all": float(np.quantile(lat, 0.5)), "p99_latency_s_all": float(np.quantile(lat, 0.99)),
            "step_kernel": flow_step_roofline(tm),
            "cpu_check": {"flows": len(sample), "fraction": len(sample) / n_flows,
                          "n_grid": [reqs[i]["n_grid"] for i in sample], "max_abs_u_diff": max(diffs),
                          "tolerance": 1e-8, "within_tolerance": max(diffs) <= 1e-8, "cpu_seconds": cpu_s,
                          "threads": O.num_threads(),
                          "parallel_over": "flows" if args.flow_check_frac > 0 else "grid points"},
            "config": "O(4) LPA flows, n_grid ~ U{64..1024}, T ~ U[0.01,0.3], mu ~ U[0,0.4], RK4 2000 steps "
                      "k=1 -> 1e-3, two 64-node threshold integrals per point and stage; HBM budget 64 MiB, "
                      "128 slots, 25 steps per graph launch"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; the multi-GPU sharding is described in the "
                                   "design notes and not implemented", "n_gpus": args.gpus}))
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    import diffrg_amd as D
    D.init(0)
    qp = D.QuadratureProvider()
    se = selfenergy(D, qp, args)
    n, evals = se["n"], se["evals"]
    fp, mx = se["res"]["fp64"], se["res"]["mixed"]
    ok = mx["err"] <= MIXED_TOL
    out = {
        "metric": "integrals/sec at rel.err<=1e-8 vs FP64 (mixed vs plain); p50/p99 flow latency",
        "value": n / (mx["ms"] * 1e-3) if ok else 0.0,
        "unit": "integrals/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": mx["ms"],
        "higher_is_better": True, "scaling": "strong", "vs_baseline": None,
        "dtype": "fp32-seed/fp64-acc", "data": "synthetic (SplitMix64 seed 0xD1FF0000+c)",
        "config": {"workload": "two-angle self-energy: 64 p on log grid [1e-3,40], nodes {64,32,32}, "
                               "128-point Z table (4,194,304 evals per step)",
                   "integrals_per_step": n, "evals_per_step": evals,
                   "launch": f"{mx['launch']} (eager launches or hipGraph of {GRAPH_SETS} call sets, the faster)"},
        "mixed": {"integrals_per_s": n / (mx["ms"] * 1e-3), "evals_per_s": evals / (mx["ms"] * 1e-3),
                  "max_rel_err_vs_fp64_oracle": mx["err"], "shape": mx["shape"], "launch": mx["launch"],
                  "us_eager": mx["ms_eager"] * 1e3, "us_graph": finite(mx["ms_graph"] * 1e3),
                  f"integrals_per_s_{N_STREAMS}_streams": finite(n / (mx["ms_streams"] * 1e-3)),
                  f"batched_x{BATCH}": mx.get(f"batched_x{BATCH}")},
        "plain_fp64": {"integrals_per_s": n / (fp["ms"] * 1e-3), "evals_per_s": evals / (fp["ms"] * 1e-3),
                       "ms_per_step": fp["ms"], "max_rel_err_vs_oracle": fp["err"], "shape": fp["shape"],
                       "launch": fp["launch"], "us_eager": fp["ms_eager"] * 1e3, "us_graph": finite(fp["ms_graph"] * 1e3),
                       f"integrals_per_s_{N_STREAMS}_streams": finite(n / (fp["ms_streams"] * 1e-3)),
                       f"batched_x{BATCH}": fp.get(f"batched_x{BATCH}")},
        "mixed_over_plain": fp["ms"] / mx["ms"],
        f"mixed_over_plain_batched_x{BATCH}": (mx[f"batched_x{BATCH}"]["integrals_per_s"] /
                                               fp[f"batched_x{BATCH}"]["integrals_per_s"]
                                               if f"batched_x{BATCH}" in mx else None),
    }
    na = se["na"]
    v1 = se["variant"][1]
    # FLOP per evaluation: the frozen weights; VALU per evaluation: the current
    # assembly of the loop that ran (for the issue fraction)
    w1, w0 = FROZEN["C1_selfenergy"]["mixed"], FROZEN["C1_selfenergy"]["fp64"]
    fl1, fl0 = w1["flop"], w0["flop"]
    va1 = se_ops(v1, na, mx["shape"]["block"], mx["unroll"])[1]
    va0 = se_ops(se["variant"][0], na, fp["shape"]["block"], fp["unroll"])[1]
    achieved = fl1 * evals / (mx["ms"] * 1e-3) / 1e12
    peak = PEAK_FP64_VALU / 1e12
    out["roofline"] = {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                       "frac": achieved / peak, "traffic": measured_traffic(), "kernel": f"se_kernel<MIXED, {mx['shape']['block']}>",
                       "integrand_form": {1: "general", 2: "narrow"}[v1],
                       "flop_per_eval": fl1, "flop_per_eval_source": "frozen flop-weights block",
                       "ceiling": ceiling(w1), "valu_per_eval": va1,
                       "loop": se_loop(na, mx["shape"]["block"], mx["unroll"]),
                       "evals_per_launch": evals,
                       "valu_issue_frac": valu_issue_frac(va1, evals, mx["ms"] * 1e-3),
                       "plain_fp64": {"flop_per_eval": fl0, "valu_per_eval": va0, "ceiling": ceiling(w0),
                                      "loop": se_loop(na, fp["shape"]["block"], fp["unroll"]),
                                      "achieved": fl0 * evals / (fp["ms"] * 1e-3) / 1e12,
                                      "frac": fl0 * evals / (fp["ms"] * 1e-3) / PEAK_FP64_VALU,
                                      "valu_issue_frac": valu_issue_frac(va0, evals, fp["ms"] * 1e-3)}}
    # the published peak confirmed on this device: independent FMA chains on every SIMD
    m64, m32 = D.valu_peak(True) / 1e12, D.valu_peak(False) / 1e12
    out["roofline"]["peak_measured"] = {"fp64_fma_tflops": m64, "fp32_fma_tflops": m32,
                                        "frac_of_measured_fp64": achieved / m64,
                                        "probe": "8 independent FMA chains per lane, 8 waves per SIMD, best of 3"}
    print(f"measured FMA peak: fp64 {m64:.1f} TFLOP/s, fp32 (not packed) {m32:.1f} TFLOP/s", file=sys.stderr)
    if not args.no_cpu_baseline:
        t_cpu, runs = cpu_timed(se["ref_fn"], args.cpu_seconds)
        from oracle import oracle as O
        out["cpu_baseline"] = {"value": n / t_cpu, "unit": "integrals/s", "cores": O.num_threads(), "kind": "port",
                               "sample": f"{runs} runs of the full call set ({n} integrals, {evals} evals), C "
                                         f"restatement of the reference TBB loop (OpenMP), median {t_cpu * 1e3:.1f} ms",
                               "evals_per_s": evals / t_cpu, "cpu_model": cpu_model()}
    if not args.only_headline:
        work = {}
        chosen = set(args.workloads.split(","))
        for name, fn in (("template", template), ("threshold", threshold), ("finite_T", finite_T),
                         ("uneven", uneven), ("flows", flows)):
            if name not in chosen:
                continue
            t0 = time.perf_counter()
            work[name] = fn(D, qp, args)
            log(f"{name} done in {time.perf_counter() - t0:.1f} s")
        out["workloads"] = work
    out["device"] = D.device_props()["name"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
