"""The restated regulator and thermal functions (oracle/shapes.py) pinned without the formulas
themselves: limits and special cases that tie the regulator families together, the continuum
identities of the thermal functions, and high-precision (mpmath) values and derivatives."""
import math

import mpmath as mp
import numpy as np
import pytest

from oracle import shapes as S

K2 = np.array([1.0, 0.25, 2.0, 1.0, 0.04])
Q2 = np.array([0.3, 2.0, 0.05, 0.999, 0.01])


def test_regulators_at_zero_momentum_equal_k2():
    # every family with R_B(k^2, q^2 -> 0) = k^2 (the mass-like IR regularisation)
    for kind, par in S.KINDS:
        got = S.RB(kind, par, K2, 1e-12 * K2)
        assert np.allclose(got, K2, rtol=1e-9), (kind, par)


def test_polynomial_exp_order_1_is_exponential_b1():
    assert np.allclose(S.RB("polynomial_exp", 1, K2, Q2), S.RB("exponential", 1.0, K2, Q2), rtol=1e-15)


def test_polynomial_exp_large_order_is_litim_below_k():
    # sum_c x^c/c -> -log(1 - x) for x < 1, so k^2 exp(-f) -> k^2 (1 - x) = k^2 - q^2
    k2 = np.array([1.0, 2.0, 0.5])
    q2 = np.array([0.3, 0.4, 0.1])
    assert np.allclose(S.RB("polynomial_exp", 60, k2, q2), k2 - q2, rtol=1e-12)


def test_smoothed_litim_small_alpha_is_litim():
    k2 = np.array([1.0, 1.0, 2.0, 2.0])
    q2 = np.array([0.5, 1.5, 0.2, 3.0])
    assert np.allclose(S.RB("smoothed_litim", 1e-4, k2, q2), S.RB("litim", None, k2, q2), rtol=0, atol=1e-12)


def test_bosonic_b1_closed_form_and_limit():
    # b = 1: R_B = q^2/(e^x - 1) -> k^2 (1 - x/2 + ...) for x -> 0
    x = np.array([1e-4, 0.5, 2.0])
    got = S.RB("bosonic", 1.0, 1.0, x)
    assert np.allclose(got, [float(mp.mpf(v) / mp.expm1(v)) for v in x], rtol=1e-14)
    assert got[0] == pytest.approx(1 - 0.5e-4, rel=1e-8)


def test_exponential_mpmath_values():
    for b in (1.0, 2.0):
        for k2, q2 in zip(K2, Q2):
            want = float(mp.mpf(k2) * mp.exp(-(mp.mpf(q2) / k2) ** int(b)))
            assert S.RB("exponential", b, k2, q2) == pytest.approx(want, rel=1e-14)


def test_complex_step_derivatives_match_mpmath():
    # the derivative operators the GPU tests use, on the exponential family against mpmath
    f = lambda k2, q2: S.RB("exponential", 2.0, k2, q2)
    for k2, q2 in zip(K2, Q2):
        g = lambda kk: mp.mpf(kk) ** 2 * mp.exp(-(mp.mpf(q2) / mp.mpf(kk) ** 2) ** 2)
        want_t = float(mp.sqrt(k2) * mp.diff(g, mp.sqrt(k2)))  # k dR/dk
        assert S.t_derivative(f, k2, q2) == pytest.approx(want_t, rel=1e-13)
        h = lambda qq: mp.mpf(k2) * mp.exp(-(mp.mpf(qq) / k2) ** 2)
        assert S.q2_derivative(f, k2, q2) == pytest.approx(float(mp.diff(h, q2)), rel=1e-13)


def test_thermal_functions_identities():
    x = np.array([0.1, 1.0, 3.0, -0.7])
    T = np.array([0.05, 0.3, 1.0, 0.2])
    # coth(x/2T) = 1 + 2 n_B(x, T) and tanh(x/2T) = 1 - 2 n_F(x, T)
    assert np.allclose(S.coth_s(x, T), 1 + 2 * S.n_b(x, T), rtol=1e-13)
    assert np.allclose(S.tanh_s(x, T), 1 - 2 * S.n_f(x, T), rtol=1e-13)
    assert np.allclose(S.coth_s(x, T) ** 2 - S.csch_s(x, T) ** 2, 1.0, rtol=1e-12)
    assert np.allclose(S.tanh_s(x, T) ** 2 + S.sech_s(x, T) ** 2, 1.0, rtol=1e-12)
    for xi, Ti in zip(x, T):
        assert S.n_b(xi, Ti) == pytest.approx(float(1 / mp.expm1(mp.mpf(xi) / Ti)), rel=1e-14)
