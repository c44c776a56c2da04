"""The per-evaluation operation counts that bench.py prices its roofline with agree with the
current gfx950 assembly of se_kernel (hipcc -S, device only; no GPU needed)."""
import os
import shutil
import tempfile

import pytest

import bench
from diffrg_amd import isa_count as I

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which(I.B.HIPCC) is None and not os.path.exists(I.B.HIPCC), reason="hipcc not found")
def test_se_kernel_counts_match_bench():
    with tempfile.TemporaryDirectory() as d:
        loops = I.inner_loops(I.compile_asm(os.path.join(ROOT, "diffrg_amd", "csrc", "integrals.hip"), d))
    for mode in (0, 1):
        for block in (64, 128, 256, 512, 1024):
            name = f"se_kernelILi{mode}ELi{block}E"
            # the integrand loops: the ones with transcendental instructions and more than one
            # table lookup; the paired loop reads the radial terms once for two evaluations.  The
            # mixed kernel holds both integrand forms; the narrow one has fewer FP64 instructions.
            body = sorted((c for n, c in loops if name in n and c["trans"] >= 8), key=lambda c: (c["f64"], c["lds"]))
            variants = (0,) if mode == 0 else (2, 1)
            assert len(body) == 2 * len(variants), (name, body)
            for j, c in enumerate(body):
                variant, loop = variants[j // 2], ("pairs", "single")[j % 2]
                na = 64 if loop == "pairs" else 8
                assert bench.se_loop(na, block) == loop
                assert bench.se_ops(variant, na, block) == (c["flop"] / 2, c["valu"] / 2), (name, variant, loop, c)


def test_se_loop_choice():
    # se_tile_sum: a lane owns two angular pairs when na^2 >= 2 * BLOCK
    assert bench.se_loop(32, 256) == "pairs" and bench.se_loop(32, 512) == "pairs"
    assert bench.se_loop(32, 1024) == "single" and bench.se_loop(16, 256) == "single"
    assert bench.se_loop(64, 1024) == "pairs"
