"""Summaries of rocprofv3 CSV output (kernel traces and counter collections).

    python -m diffrg_amd.prof_summary trace  <run>_kernel_trace.csv [--last N]
    python -m diffrg_amd.prof_summary pmc    <run>_counter_collection.csv [--last N]

Dispatches are grouped by kernel instance (template instantiation id, workgroup size, grid size
and LDS size), so the plain and mixed instantiations of one kernel and the launch shapes an
autotuner tried are reported separately.  --last N keeps the last N dispatches of each group
(the timed region of a benchmark that ends with N timed launches of one shape).
"""
from __future__ import annotations

import argparse
import collections
import csv
import statistics


def _key(r):
    grid = r.get("Grid_Size_X", r.get("Grid_Size"))
    wg = r.get("Workgroup_Size_X", r.get("Workgroup_Size"))
    return (r["Kernel_Name"][:48], r["Kernel_Id"], wg, grid, r["LDS_Block_Size"], r["VGPR_Count"])


def trace(path: str, last: int = 0):
    groups = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        groups[_key(r)].append((int(r["Correlation_Id"]), int(r["End_Timestamp"]) - int(r["Start_Timestamp"])))
    out = []
    for k, v in groups.items():
        v = [d for _, d in sorted(v)]
        if last:
            v = v[-last:]
        out.append((k, len(v), statistics.mean(v) / 1e3, statistics.median(v) / 1e3, min(v) / 1e3, sum(v) / 1e3))
    out.sort(key=lambda t: -t[5])
    print(f"{'kernel':48s} {'id':>4s} {'wg':>4s} {'grid':>8s} {'lds':>6s} {'vgpr':>4s} {'calls':>6s} "
          f"{'avg_us':>8s} {'med_us':>8s} {'min_us':>8s}")
    for k, n, avg, med, mn, _ in out:
        print(f"{k[0]:48s} {k[1]:>4s} {k[2]:>4s} {k[3]:>8s} {k[4]:>6s} {k[5]:>4s} {n:6d} {avg:8.2f} {med:8.2f} {mn:8.2f}")


def pmc(path: str, last: int = 0):
    groups = collections.defaultdict(lambda: collections.defaultdict(list))
    for r in csv.DictReader(open(path)):
        groups[_key(r)][r["Counter_Name"]].append((int(r["Dispatch_Id"]), float(r["Counter_Value"])))
    print(f"{'kernel':48s} {'id':>4s} {'wg':>4s} {'grid':>8s} {'lds':>6s} {'counter':24s} {'n':>5s} {'median':>14s}")
    for k, counters in groups.items():
        for name, v in sorted(counters.items()):
            v = [x for _, x in sorted(v)]
            if last:
                v = v[-last:]
            print(f"{k[0]:48s} {k[1]:>4s} {k[2]:>4s} {k[3]:>8s} {k[4]:>6s} {name:24s} {len(v):5d} "
                  f"{statistics.median(v):14.1f}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("kind", choices=("trace", "pmc"))
    ap.add_argument("csv")
    ap.add_argument("--last", type=int, default=0)
    a = ap.parse_args()
    (trace if a.kind == "trace" else pmc)(a.csv, a.last)


if __name__ == "__main__":
    main()
