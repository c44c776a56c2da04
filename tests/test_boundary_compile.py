"""The drop-in boundary compiles against the reference's own interfaces (hipcc -c for gfx950, no
GPU, nothing linked or run; tests/native/compile_boundary.hip):

- every integrator, the QuadratureProvider and request_data accept a parameter object with only the
  reference JSONValue getters (the JSON constructors are templates, no JSONValue type is shipped);
- the reference's generated kernels Examples/ONfiniteT/flows/V/V.kernel and
  Examples/YangMills/SP/flows/AA/ZA.kernel compile unchanged, by include path, in the integrators
  their wrappers use (skipped where the reference tree is not present);
- the drop-in tree ships no header at the reference's common/json.hh, math.hh, types.hh or
  complex_math.hh paths, so the reference's own versions serve the rest of the framework.
"""
import os
import shutil
import subprocess
import tempfile

import pytest

from diffrg_amd import build as B

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "compile_boundary.hip")
REF = "/root/reference"
REF_V = os.path.join(REF, "Examples", "ONfiniteT", "flows", "V")
REF_ZA = os.path.join(REF, "Examples", "YangMills", "SP", "flows", "AA")

needs_hipcc = pytest.mark.skipif(shutil.which(B.HIPCC) is None and not os.path.exists(B.HIPCC),
                                 reason="hipcc not found")


def _compile(extra):
    with tempfile.TemporaryDirectory() as d:
        cmd = [B.HIPCC, *B.COMMON_FLAGS, f"-I{B.STANDALONE}", "-fopenmp", *extra, "-c", SRC, "-o",
               os.path.join(d, "boundary.o")]
        r = subprocess.run(cmd, capture_output=True, text=True)
    msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    assert r.returncode == 0, msg[-4000:]


@needs_hipcc
def test_integrators_bind_reference_json_getters():
    _compile([])


@needs_hipcc
@pytest.mark.skipif(not (os.path.isdir(REF_V) and os.path.isdir(REF_ZA)), reason="reference tree not present")
def test_reference_generated_kernels_compile_unchanged():
    _compile(["-DDFRG_REFERENCE_KERNELS", f"-I{REF_V}", f"-I{REF_ZA}"])


def test_drop_in_ships_no_reference_common_headers():
    common = os.path.join(ROOT, "include", "DiFfRG", "common")
    for name in ("json.hh", "math.hh", "types.hh", "complex_math.hh", "utils.hh", "cuda_prefix.hh"):
        assert not os.path.exists(os.path.join(common, name)), name
    # the stand-ins live apart, for builds without the reference tree
    for name in ("json.hh", "math.hh", "types.hh"):
        assert os.path.exists(os.path.join(ROOT, "include", "standalone", "DiFfRG", "common", name))
