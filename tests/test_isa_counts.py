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
    kernels = [(0, b, 1) for b in (64, 128, 256, 512, 1024)] + \
              [(1, b, u) for b in (64, 128, 256, 512, 1024) for u in (1, 2)]
    for mode, block, unroll in kernels:
        name = f"se_kernelILi{mode}ELi{block}ELi{unroll}E"
        # the integrand loops: the ones with transcendental instructions (4 per FP64 evaluation, 5
        # per mixed one); the mixed kernel holds both integrand forms, told apart by their FP64
        # instructions per evaluation (narrow 20, general 21).  Per form: the paired loop reads
        # three radial terms per radial node (fewer LDS reads than the single loop per two
        # evaluations), and with unroll 2 a paired loop of four evaluations per step comes first.
        body = [c for n, c in loops if name in n and c["trans"] >= 8]
        per_trans = 4 if mode == 0 else 5
        forms = {}
        for c in body:
            evals = c["trans"] // per_trans
            form = 0 if mode == 0 else (2 if c["f64"] // evals == 20 else 1)
            forms.setdefault(form, []).append((evals, c["lds"], c))
        assert sorted(forms) == ([0] if mode == 0 else [1, 2]), (name, body)
        for form, ls in forms.items():
            ls.sort(key=lambda t: (-t[0], t[1]))
            want = ["pairs2", "pairs", "single"] if unroll == 2 else ["pairs", "single"]
            assert [t[0] for t in ls] == ([4, 2, 2] if unroll == 2 else [2, 2]), (name, ls)
            for kind, (evals, _, c) in zip(want, ls):
                per_eval = (c["flop"] / evals, c["valu"] / evals)
                if kind == "pairs" and unroll == 2:
                    continue  # the remainder step of an odd radial count, not priced
                na = 8 if kind == "single" else 64
                assert bench.se_loop(na, block, unroll) == kind
                assert bench.se_ops(form, na, block, unroll) == per_eval, (name, form, kind, c)


def test_se_loop_choice():
    # se_tile_sum: a lane owns two angular pairs when na^2 >= 2 * BLOCK
    assert bench.se_loop(32, 256) == "pairs" and bench.se_loop(32, 512) == "pairs"
    assert bench.se_loop(32, 1024) == "single" and bench.se_loop(16, 256) == "single"
    assert bench.se_loop(64, 1024) == "pairs"


def _kernel_body(asm_text, symbol_prefix):
    import re
    m = re.search(r"^(" + symbol_prefix + r"\w*):(.*?)s_endpgm", asm_text, re.M | re.S)
    assert m, symbol_prefix
    return [l.strip() for l in m.group(2).split("\n")]


@pytest.mark.skipif(shutil.which(I.B.HIPCC) is None and not os.path.exists(I.B.HIPCC), reason="hipcc not found")
def test_cross_block_publication_lowering():
    """The cross-block finish (dfrg::dev::publish / wait_stores / take_ticket / fetch) uses relaxed
    agent-scope atomics, not release/acquire; its correctness rests on how gfx950 lowers them:
    the partial's store and the finishing block's load carry sc1 (device scope: written through
    to, and read from, the level where the XCDs are coherent), the store is complete (vmcnt
    counts stores on gfx9) before the ticket atomic is issued, and the ticket is an atomic with a
    returned value (sc0).  This test fails if a compiler or flag change alters that lowering."""
    with tempfile.TemporaryDirectory() as d:
        text = open(I.compile_asm(os.path.join(ROOT, "diffrg_amd", "csrc", "integrals.hip"), d)).read()
    for sym in ("_ZN12_GLOBAL__N_19se_kernelILi1ELi512ELi2E", "_ZN12_GLOBAL__N_19se_kernelILi0ELi256ELi1E"):
        body = _kernel_body(text, sym)
        stores = [k for k, l in enumerate(body) if l.startswith("global_store_dwordx2") and l.endswith(" sc1")]
        assert stores, (sym, "no device-scope store of a partial")
        i = stores[0]
        j = next(k for k in range(i + 1, len(body)) if body[k].startswith("global_atomic_add"))
        assert "s_waitcnt vmcnt(0)" in body[i + 1:j], (sym, body[i:j + 1])
        assert body[j].endswith(" sc0"), body[j]  # returns the pre-increment ticket
        assert any(l.startswith("global_load_dwordx2") and l.endswith(" sc1") for l in body[j:]), sym
