"""CPU checks that pin the oracle's quadrature rules (the GPU parity tests compare against it).

Sources of truth: the reference's 30-digit GLQuadrature<N> tables (fixture), the reference's
own known answers (test_quadrature.cc, test_diagonalization.cc, test_matsubara.cc), closed forms
and 50-digit mpmath eigen-decompositions.
"""
import json
import math
import os

import mpmath as mp
import numpy as np
import pytest

from oracle import oracle as O

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "gl_tables.json")))


@pytest.mark.parametrize("n", GOLD["orders"])
def test_gl_matches_reference_tables(n):
    # the 30-digit tables rounded to double; the oracle must agree to <= 1 ulp
    x, w = O.gl_nodes(n)
    tx = np.array([float(mp.mpf(s)) for s in GOLD["tables"][str(n)]["x"]])
    tw = np.array([float(mp.mpf(s)) for s in GOLD["tables"][str(n)]["w"]])
    assert np.all(np.abs(x - tx) <= 2.3e-16 * np.maximum(tx, 1e-300) + 1e-300)
    assert np.all(np.abs(w - tw) <= 2.3e-16 * tw)
    assert np.all(np.diff(x) > 0)
    assert abs(np.sum(w) - 1.0) < 1e-15


@pytest.mark.parametrize("n", [4, 8, 16, 32, 64, 128])
def test_gl_from_jacobi_matrix_and_imtqlx(n):
    # test_quadrature.cc:10-33: make_quadrature on the Legendre Jacobi matrix vs the tables
    b = np.array([i * i / (4.0 * i * i - 1.0) for i in range(1, n + 1)])
    x, w = O.make_quadrature(np.zeros(n), b, 2.0)
    tx = np.array([float(mp.mpf(s)) for s in GOLD["tables"][str(n)]["x"]])
    tw = np.array([float(mp.mpf(s)) for s in GOLD["tables"][str(n)]["w"]])
    assert np.max(np.abs(x - (2 * tx - 1))) < 1e-14
    assert np.max(np.abs(w - 2 * tw) / (2 * tw)) < 5e-12  # QL weight accuracy ~1e-12 at n=128


def test_imtqlx_2x2_and_3x3_closed_forms():
    rng = np.random.default_rng(11)
    for _ in range(20):
        a, b, c = rng.uniform(1, 10, 3)
        d, z = O.imtqlx([a, b], [c, 0.0], [1.0, 0.0])
        disc = math.sqrt(a * a - 2 * a * b + b * b + 4 * c * c)
        assert d[0] == pytest.approx((a + b - disc) / 2, rel=1e-14)
        assert d[1] == pytest.approx((a + b + disc) / 2, rel=1e-14)
        # first eigenvector components (test_diagonalization.cc:26-40)
        v1 = -0.5 * (-a + b + disc) / (c * math.sqrt(1 + (-a + b + disc) ** 2 / (4 * c * c)))
        v2 = -0.5 * (-a + b - disc) / (c * math.sqrt(1 + (-a + b - disc) ** 2 / (4 * c * c)))
        assert abs(z[0]) == pytest.approx(abs(v1), rel=1e-12)
        assert abs(z[1]) == pytest.approx(abs(v2), rel=1e-12)
    for _ in range(10):
        a = rng.uniform(1, 10, 3)
        e = np.sqrt(rng.uniform(1, 10, 2))
        d, z = O.imtqlx(a, [e[0], e[1], 0.0], [1.0, 0.0, 0.0])
        M = np.diag(a) + np.diag(e, 1) + np.diag(e, -1)
        ev, V = np.linalg.eigh(M)
        assert np.allclose(d, ev, rtol=1e-13)
        assert np.allclose(np.abs(z), np.abs(V[0]), rtol=1e-11)


@pytest.mark.parametrize("T", [0.013, 0.05, 0.2, 0.61, 0.97])
def test_matsubara_coth_closed_form(T):
    # test_matsubara.cc:11-23: T sum 1/(1+(2 pi n T)^2) = coth(1/2T)/2 at 1e-13
    x, w, Tr = O.matsubara(T, 1.0)
    s = O.matsubara_sum(x, w, Tr, lambda v: 1.0 / (1.0 + v * v))
    assert s == pytest.approx(0.5 / math.tanh(0.5 / T), rel=1e-13)


def test_matsubara_general_closed_form():
    # test_matsubara.cc:24-43
    rng = np.random.default_rng(5)
    n_done = 0
    while n_done < 25:
        T, a, b = rng.uniform(0.001, 0.1), rng.uniform(0.1, 1.0), rng.uniform(0.01, 0.1)
        if a / T > 100.0:
            continue  # the T = 0 rule applies there (matsubara.cu:14); it is checked at 3e-6 below
        n_done += 1
        x, w, Tr = O.matsubara(T, a)
        s = O.matsubara_sum(x, w, Tr, lambda v: 1.0 / (a * a + b * v + v * v))
        r = math.sqrt(4 * a * a - b * b)
        ref = -(math.sinh(r / (2 * T)) / (r * (math.cos(b / (2 * T)) - math.cosh(r / (2 * T)))))
        assert s == pytest.approx(ref, rel=1e-13)


def test_matsubara_vacuum_limit():
    # test_matsubara.cc:44-59: T = 0, sum of 1/(a^2+x^2) -> 1/(2a) to 3e-6
    rng = np.random.default_rng(3)
    for a in rng.uniform(0.01, 10.0, 20):
        x, w, Tr = O.matsubara(0.0, a)
        assert Tr == 0.0 and len(x) == 48
        s = O.matsubara_sum(x, w, Tr, lambda v: 1.0 / (a * a + v * v))
        assert s == pytest.approx(1.0 / (2 * a), rel=3e-6)


def test_matsubara_sizes():
    # predict_size (matsubara.cu:12-21) and the storage clamps (quadrature_provider.cu:52)
    assert O.matsubara_predict_size(0.0, 1.0) == -48
    assert O.matsubara_predict_size(0.001, 1.0) == -48  # E/T > 100: vacuum rule
    n = O.matsubara_predict_size(0.1, 1.0, 2, 48, 0)
    assert n == int(math.ceil((5 + int(math.sqrt(400.0 / (math.pi ** 2 * 0.1)))) / 2) * 2)
    x, w, _ = O.matsubara_storage_rule(1.0, 10.0)  # predicted 26, clamped up to min 32
    assert len(x) == 32
    x, w, _ = O.matsubara_storage_rule(0.02, 1.9)  # 5 + int(sqrt(760/(pi^2 0.02))) = 67 -> even 68
    assert len(x) == 68
    x, w, _ = O.matsubara_storage_rule(0.05, 1.0)  # 5 + int(sqrt(400/(pi^2 0.05))) = 33 -> even 34
    assert len(x) == 34
    x, w, _ = O.matsubara_storage_rule(0.01, 0.1)
    assert len(x) == 32
    x, w, _ = O.matsubara_storage_rule(0.01, 0.999)
    assert len(x) == 68


def _monien_mp(n, T):
    mp.mp.dps = 40
    a = [2 * mp.pi ** 2 / (4 * j + 1) / (4 * j + 5) for j in range(n)]
    a[0] = mp.pi ** 2 / 15
    b = [mp.pi ** 4 / ((4 * (j + 1) - 1) * (4 * (j + 1) + 3)) / (4 * (j + 1) + 1) ** 2 for j in range(n)]
    A = mp.matrix(n, n)
    for i in range(n):
        A[i, i] = a[i]
        if i < n - 1:
            A[i, i + 1] = A[i + 1, i] = mp.sqrt(b[i])
    E, Q = mp.eigsy(A)
    idx = sorted(range(n), key=lambda i: E[i])
    x = np.array([float(2 * mp.pi * T / mp.sqrt(E[i])) for i in idx])
    w = np.array([float(T * (mp.pi ** 2 / 6) * Q[0, i] ** 2 / E[i]) for i in idx])
    return x, w


def test_matsubara_nodes_vs_mpmath():
    x, w, _ = O.matsubara(0.1, 1.0, 2, 24, 24, 48, 1)
    xr, wr = _monien_mp(24, 0.1)
    assert np.max(np.abs(x - xr) / xr) < 1e-12
    assert np.max(np.abs(w - wr) / wr) < 1e-12
