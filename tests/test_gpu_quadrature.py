"""GPU quadrature generation (dfrg_gl_generate, dfrg_matsubara_generate, dfrg_qp_*) against the
reference's GL tables, the oracle's restatement of the reference rules, and closed forms."""
import json
import math
import os

import mpmath as mp
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "gl_tables.json")))


def test_gl_device_vs_reference_tables(gpu):
    for n in GOLD["orders"]:
        print(f"[case] gl order {n}", flush=True)
        x, w = gpu.gl_generate(n)
        tx = np.array([float(mp.mpf(s)) for s in GOLD["tables"][str(n)]["x"]])
        tw = np.array([float(mp.mpf(s)) for s in GOLD["tables"][str(n)]["w"]])
        assert np.all(np.abs(x - tx) <= 4.5e-16 * tx), n
        assert np.all(np.abs(w - tw) <= 4.5e-16 * tw), n


@pytest.mark.parametrize("n", [5, 100, 200, 333, 1000])
def test_gl_device_vs_oracle_other_orders(gpu, n):
    print(f"[case] gl order {n} vs oracle", flush=True)
    x, w = gpu.gl_generate(n)
    ox, ow = O.gl_nodes(n)
    # the long-double oracle itself drifts to ~4e-15 at order 1000; the GPU rule is double-double
    tol = 1e-15 if n <= 200 else 1e-14
    assert np.max(np.abs(x - ox) / ox) < tol
    assert np.max(np.abs(w - ow) / ow) < 10 * tol


@pytest.mark.parametrize("T,E", [(0.01, 1.0), (0.05, 1.0), (0.1, 0.3), (0.3, 2.0), (1.0, 10.0), (0.02, 1.9)])
def test_matsubara_device_vs_oracle(gpu, T, E):
    print(f"[case] matsubara T={T} E={E}", flush=True)
    x, w, Tr = gpu.matsubara_generate(T, E, 2, 32, 128, 48, 0)
    ox, ow, oT = O.matsubara(T, E, 2, 32, 128, 48, 0)
    assert len(x) == len(ox) and Tr == oT
    # both within ~1e-11 of the exact rule (IMTQLX is the less accurate of the two)
    assert np.max(np.abs(x - ox) / ox) < 3e-11
    assert np.max(np.abs(w - ow) / ow) < 3e-11
    s = O.matsubara_sum(x, w, Tr, lambda v: 1.0 / (1.0 + v * v))
    assert s == pytest.approx(0.5 / math.tanh(0.5 / T), rel=1e-13)


def test_matsubara_device_vs_mpmath(gpu):
    print("[case] matsubara n=24 vs 40-digit eigen-decomposition", flush=True)
    n, T = 24, 0.1
    x, w, _ = gpu.matsubara_generate(T, 1.0, 2, n, n, 48, 1)
    mp.mp.dps = 40
    a = [2 * mp.pi ** 2 / (4 * j + 1) / (4 * j + 5) for j in range(n)]
    a[0] = mp.pi ** 2 / 15
    b = [mp.pi ** 4 / ((4 * (j + 1) - 1) * (4 * (j + 1) + 3)) / (4 * (j + 1) + 1) ** 2 for j in range(n)]
    A = mp.matrix(n, n)
    for i in range(n):
        A[i, i] = a[i]
        if i < n - 1:
            A[i, i + 1] = A[i + 1, i] = mp.sqrt(b[i])
    Ev, Q = mp.eigsy(A)
    idx = sorted(range(n), key=lambda i: Ev[i])
    xr = np.array([float(2 * mp.pi * T / mp.sqrt(Ev[i])) for i in idx])
    wr = np.array([float(T * (mp.pi ** 2 / 6) * Q[0, i] ** 2 / Ev[i]) for i in idx])
    assert np.max(np.abs(x - xr) / xr) < 1e-13
    assert np.max(np.abs(w - wr) / wr) < 1e-13


@pytest.mark.parametrize("E", [0.05, 1.0, 7.0])
def test_matsubara_vacuum_rule(gpu, E):
    print(f"[case] vacuum rule E={E}", flush=True)
    x, w, Tr = gpu.matsubara_generate(0.0, E, 2, 32, 128, 48, 0)
    ox, ow, oT = O.matsubara(0.0, E, 2, 32, 128, 48, 0)
    assert Tr == 0.0 and len(x) == len(ox) == 48
    assert np.max(np.abs(x - ox) / ox) < 1e-14
    assert np.max(np.abs(w - ow) / ow) < 1e-13
    s = O.matsubara_sum(x, w, 0.0, lambda v: 1.0 / (E * E + v * v))
    assert s == pytest.approx(1.0 / (2 * E), rel=3e-6)


def test_provider_cache_and_errors(gpu, qp):
    print("[case] provider cache", flush=True)
    p64 = qp.get_points(64)
    x, _ = gpu.gl_generate(64)
    assert np.array_equal(p64, x)
    assert qp.get_device_points(64) == qp.get_device_points(64)
    n1, T1, dx1, _ = qp.get_device_matsubara(0.1, 1.0)
    n2, T2, dx2, _ = qp.get_device_matsubara(0.1 * (1 + 5e-7), 1.0 * (1 - 5e-7))  # within the 1e-6 key match
    assert dx1 == dx2 and n1 == n2
    n3, _, dx3, _ = qp.get_device_matsubara(0.1 * (1 + 5e-6), 1.0)
    assert dx3 != dx1
    mx = qp.get_matsubara_points(0.1, 1.0)
    ox, _, _ = O.matsubara_storage_rule(0.1, 1.0)
    assert len(mx) == len(ox)
    with pytest.raises(gpu.DfrgInvalidArgument):
        gpu.QuadratureProvider(vacuum_quad_size=0)
    with pytest.raises(gpu.DfrgInvalidArgument):
        qp.get_points(0)
