"""Operation counts of the integrand loops, read from the gfx950 assembly of this library.

    python -m diffrg_amd.isa_count            # innermost loops of every kernel in csrc/integrals.hip
    python -m diffrg_amd.isa_count --weights  # FP-op weights of the FP64 math functions (ocml)

The FLOP-per-evaluation figures of the roofline come from here: the source is compiled for
gfx950 with the library's own flags to assembly (-S, device only, into a temporary folder), the
innermost loops of each kernel are found from their backward branches, and their instructions
are counted with FMA = 2 FLOP; add, sub, mul, min/max, floor/rint/fract, ldexp and the hardware
transcendentals (exp, log, rcp, rsq, sqrt) and med3 (a clamp) = 1; moves, compares, selects, conversions,
integer and LDS instructions = 0.  VALU counts every vector instruction, `trans` the
quarter-rate transcendental ones.
"""
from __future__ import annotations

import argparse
import collections
import os
import re
import subprocess
import tempfile

from . import build as B

ROOT = os.path.dirname(os.path.abspath(__file__))

_WEIGHTS_SRC = r"""
#include <hip/hip_runtime.h>
extern "C" __global__ void w_exp(double* a) { int i = threadIdx.x; a[i] = exp(a[i]); }
extern "C" __global__ void w_log1p(double* a) { int i = threadIdx.x; a[i] = log1p(a[i]); }
extern "C" __global__ void w_div(double* a) { int i = threadIdx.x; a[i] = a[i] / a[i + 64]; }
extern "C" __global__ void w_sqrt(double* a) { int i = threadIdx.x; a[i] = sqrt(a[i]); }
extern "C" __global__ void w_expm1(double* a) { int i = threadIdx.x; a[i] = expm1(a[i]); }
"""


def flop(op: str) -> int:
    if not op.startswith("v_") or not ("f64" in op or "f32" in op):
        return 0
    if any(t in op for t in ("cvt", "cmp", "cndmask", "mov", "class", "frexp")):
        return 0
    if any(t in op for t in ("fma", "fmac", "mad")):
        return 2
    if any(t in op for t in ("add", "sub", "mul", "exp", "log", "rcp", "rsq", "sqrt", "ldexp", "div_", "fract",
                             "rndne", "floor", "ceil", "trunc", "max", "min", "med3")):
        return 1
    return 0


def is_trans(op: str) -> bool:
    return any(op.startswith(t) for t in ("v_exp", "v_log", "v_rcp", "v_rsq", "v_sqrt"))


def _ops(lines):
    out = []
    for line in lines:
        s = line.strip()
        if s and not s.startswith((".", ";")) and not s.endswith(":"):
            out.append(s.split()[0])
    return out


def count(ops) -> dict:
    valu = [o for o in ops if o.startswith("v_")]
    return {"valu": len(valu), "flop": sum(flop(o) for o in ops), "trans": sum(is_trans(o) for o in valu),
            "f64": sum("f64" in o for o in valu), "lds": sum(o.startswith("ds_") for o in ops),
            "salu": sum(o.startswith("s_") for o in ops)}


def compile_asm(src: str, out_dir: str) -> str:
    out = os.path.join(out_dir, os.path.basename(src) + ".s")
    cmd = [B.HIPCC, *B.COMMON_FLAGS, "-S", "--cuda-device-only", "-o", out, src]
    subprocess.run(cmd, check=True, capture_output=True)
    return out


def inner_loops(asm: str, min_valu: int = 20):
    """(kernel symbol, loop counts) for every innermost loop with at least min_valu VALU instructions."""
    text = open(asm).read()
    starts = [(m.start(), m.group(1)) for m in re.finditer(r"^(_Z\w+|\w+_kernel\w*):", text, re.M)]
    res = []
    for k, (pos, name) in enumerate(starts):
        end = starts[k + 1][0] if k + 1 < len(starts) else len(text)
        lines = text[pos:end].split("\n")
        labels = {m.group(1): i for i, l in enumerate(lines) for m in [re.match(r"^(\.LBB\w+):", l)] if m}
        loops = []
        for i, l in enumerate(lines):
            m = re.search(r"s_(?:cbranch_\w+|branch)\s+(\.LBB\w+)", l)
            if m and m.group(1) in labels and labels[m.group(1)] < i:
                loops.append((labels[m.group(1)], i))
        for b, e in loops:
            if any(b2 > b and e2 <= e and (b2, e2) != (b, e) for b2, e2 in loops):
                continue
            c = count(_ops(lines[b:e + 1]))
            if c["valu"] >= min_valu:
                res.append((name, c))
    return res


def weights(out_dir: str) -> dict:
    src = os.path.join(out_dir, "weights.hip")
    with open(src, "w") as f:
        f.write(_WEIGHTS_SRC)
    text = open(compile_asm(src, out_dir)).read()
    return {m.group(1): count(_ops(m.group(2).split("\n")))
            for m in re.finditer(r"^w_(\w+):(.*?)s_endpgm", text, re.M | re.S)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--weights", action="store_true")
    ap.add_argument("--source", default=os.path.join(ROOT, "csrc", "integrals.hip"))
    a = ap.parse_args()
    with tempfile.TemporaryDirectory() as d:
        if a.weights:
            for name, c in weights(d).items():
                print(f"{name:8s} {c}")
            return
        for name, c in inner_loops(compile_asm(a.source, d)):
            print(f"{name[:60]:60s} {c}")


if __name__ == "__main__":
    main()
