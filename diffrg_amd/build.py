"""Build libdfrg.so (the gfx950 HIP runtime and kernels) in-tree.

    python -m diffrg_amd.build            # or diffrg_amd.build.build()

Every source under diffrg_amd/csrc/*.hip is compiled with hipcc for gfx950 and linked into
diffrg_amd/libdfrg.so.  Packed FP32 instructions are disabled for all device code.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
# stand-ins for the reference headers the drop-in does not replace (json, math, types, regulators,
# coordinates ...), for builds without the reference tree; never ahead of the reference's own
STANDALONE = os.path.join(INCLUDE, "standalone")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libdfrg.so")

ARCH = os.environ.get("DFRG_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

COMMON_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++20", "-fPIC", f"-I{INCLUDE}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-value", "-Wno-unused-variable",
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src).replace(".hip", ".o"))
    cmd = [HIPCC, *COMMON_FLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{msg}")
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources for {ARCH}")
    return LIB


KAT_SRC = os.path.join(ROOT, "tests", "native", "kat_kernels.hip")
KAT_LIB = os.path.join(ROOT, "tests", "native", "libdfrg_kat.so")


def build_kat(verbose: bool = True) -> str:
    """Compile the test library that instantiates the C++ integrator templates (tests/native)."""
    tmp = KAT_LIB + ".tmp"
    cmd = [HIPCC, *COMMON_FLAGS, f"-I{STANDALONE}", "-fopenmp", "-shared", KAT_SRC, "-o", tmp, f"-L{PKG}", "-ldfrg",
           "-Wl,-rpath,$ORIGIN/../../diffrg_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {KAT_SRC}:\n{msg}")
    os.replace(tmp, KAT_LIB)
    if verbose:
        print(f"built {KAT_LIB}")
    return KAT_LIB


if __name__ == "__main__":
    build()
    build_kat()
    sys.exit(0)
