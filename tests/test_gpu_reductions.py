"""dfrg::dev::wave_sum8 and wave_sum4 (include/dfrg/device.hh), cross-lane reductions of eight and
four vectors (the flow server uses wave_sum4), on one wave through tests/native/libdfrg_kat.so.

Pinned by exact arithmetic: integer-valued inputs sum exactly, so every lane must hold exactly the
total of the vector its lane bits name.  With general values, a vector's total must not depend on
its position among the eight or on the other vectors (bitwise), which the flow server's
alone-or-in-the-mix determinism relies on.
"""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_dp = C.POINTER(C.c_double)


@pytest.fixture(scope="module")
def kat(gpu):
    lib = C.CDLL(os.path.join(ROOT, "tests", "native", "libdfrg_kat.so"))
    lib.kat_last_error.restype = C.c_char_p
    return lib


def run(kat, vecs):
    vecs = np.ascontiguousarray(vecs, dtype=np.float64)
    assert vecs.shape == (8, 64)
    out = np.empty(64)
    rc = kat.kat_wave_sum8(vecs.ctypes.data_as(_dp), out.ctypes.data_as(_dp))
    assert rc == 0, kat.kat_last_error()
    return out


def vector_of_lane(l):
    return 4 * ((l >> 3) & 1) + 2 * ((l >> 4) & 1) + ((l >> 5) & 1)


def sum8_lane(v):
    return 8 * (v >> 2) + 16 * ((v >> 1) & 1) + 32 * (v & 1)


def test_lane_layout_exact(kat):
    rng = np.random.default_rng(8)
    print("[case] wave_sum8 exact", flush=True)
    vecs = rng.integers(-1000, 1000, (8, 64)).astype(np.float64)
    out = run(kat, vecs)
    want = vecs.sum(axis=1)
    for l in range(64):
        assert out[l] == want[vector_of_lane(l)], l
    for v in range(8):
        assert all(out[sum8_lane(v) + j] == want[v] for j in range(8))


def test_total_independent_of_position(kat):
    rng = np.random.default_rng(9)
    vecs = rng.standard_normal((8, 64)) * np.exp(rng.uniform(-20, 20, (8, 64)))
    print("[case] wave_sum8 positions", flush=True)
    base = run(kat, vecs)
    totals = {v: base[sum8_lane(v)] for v in range(8)}
    for rep in range(6):
        perm = rng.permutation(8)
        others = rng.standard_normal((8, 64))
        mix = np.where(rng.random(8)[:, None] < 0.5, vecs[perm], others)
        keep = [i for i in range(8) if np.array_equal(mix[i], vecs[perm[i]])]
        out = run(kat, mix)
        for i in keep:
            assert out[sum8_lane(i)] == totals[perm[i]]  # bitwise
        for i in range(8):
            assert out[sum8_lane(i)] == pytest.approx(mix[i].sum(), rel=1e-12, abs=1e-300 + 1e-12 * np.abs(mix[i]).sum())


def run4(kat, vecs):
    vecs = np.ascontiguousarray(vecs, dtype=np.float64)
    assert vecs.shape == (4, 64)
    out = np.empty(64)
    rc = kat.kat_wave_sum4(vecs.ctypes.data_as(_dp), out.ctypes.data_as(_dp))
    assert rc == 0, kat.kat_last_error()
    return out


def test_sum4_lane_layout_exact(kat):
    rng = np.random.default_rng(10)
    print("[case] wave_sum4 exact", flush=True)
    vecs = rng.integers(-1000, 1000, (4, 64)).astype(np.float64)
    out = run4(kat, vecs)
    want = vecs.sum(axis=1)
    for l in range(64):
        assert out[l] == want[2 * ((l >> 4) & 1) + ((l >> 5) & 1)], l


def test_sum4_is_sum8_tree(kat):
    # each of the four vectors gets bitwise the total wave_sum8 gives it in the same position, for
    # values whose sums round
    rng = np.random.default_rng(11)
    print("[case] wave_sum4 against wave_sum8", flush=True)
    for rep in range(6):
        vecs = rng.standard_normal((8, 64)) * np.exp(rng.uniform(-20, 20, (8, 64)))
        out8 = run(kat, vecs)
        out4 = run4(kat, vecs[:4])
        for v in range(4):
            assert all(out4[sum8_lane(v) + j] == out8[sum8_lane(v)] for j in range(16))
