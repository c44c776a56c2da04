"""Regulator shape functions and thermal functions of DiFfRG, restated for the tests.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the reference values of the regulator and
thermodynamics checks; the product never imports it.

Only the defining functions are restated -- the regulator R_B(k^2, q^2) of each family and the
thermal functions coth, tanh, sech, csch, n_B, n_F of e/T -- written for numpy arrays that may be
complex, so that every derived quantity (R_B-dot = k dR_B/dk, dR_B/dq^2, R_F = sqrt(R_B + q^2) - q,
its t- and q^2-derivatives, and the e-derivatives of the thermal functions) is checked against a
complex-step or high-precision numerical derivative of these, not against a second copy of the
formulas.  The defining functions themselves are pinned in tests/test_oracle_shapes.py by limits
and special cases (Litim as the small-alpha limit of the smoothed Litim and the large-order limit of
the polynomial-exponential, order 1 = exponential with b = 1, R_B(k^2, 0) = k^2, ...).

Reference definitions (DiFfRG/include/DiFfRG/physics/regulators.hh):
  Litim              :24-51    R_B = (k^2 - q^2) theta(k^2 - q^2)
  Bosonic<b>         :72-123   R_B = q^2 x^(b-1) / (exp(x^b) - 1),      x = q^2/k^2
  Exponential<b>     :144-191  R_B = k^2 exp(-x^b)
  SmoothedLitim<a>   :212-242  R_B = (k^2 - q^2) / (1 + exp((x - 1)/a))
  PolynomialExp<n>   :625-695  R_B = k^2 exp(-sum_{c=1..n} x^c / c)
Thermal functions (thermodynamics.hh:28-135): coth/tanh/sech/csch of e/(2T), n_B = 1/(exp(e/T) - 1),
n_F = 1/(exp(e/T) + 1); the FiniteT forms take the T/x -> 0 limit sign(x) (coth, tanh) and 0 where
cosh/sinh overflow (sech, csch).
"""
from __future__ import annotations

import numpy as np

# (name, parameter) of the regulator kinds the device tests instantiate, in kat_regulator_eval order
KINDS = [("litim", None), ("bosonic", 2.0), ("bosonic", 1.0), ("exponential", 1.0), ("exponential", 2.0),
         ("smoothed_litim", 2e-3), ("smoothed_litim", 1e-4), ("polynomial_exp", 8), ("polynomial_exp", 1),
         ("polynomial_exp", 60)]
# which of (RB, RBdot, dq2RB, RF, RFdot, dq2RF) each kind defines
DEFINES = {"litim": (1, 1, 0, 1, 1, 0), "bosonic": (1, 1, 1, 1, 1, 1), "exponential": (1, 1, 1, 1, 1, 1),
           "smoothed_litim": (1, 1, 0, 1, 1, 0), "polynomial_exp": (1, 1, 1, 1, 1, 1)}


def RB(kind, par, k2, q2):
    """R_B(k^2, q^2) of a regulator family (numpy, real or complex arguments)."""
    k2 = np.asarray(k2)
    q2 = np.asarray(q2)
    x = q2 / k2
    if kind == "litim":
        return np.where(np.real(k2) > np.real(q2), k2 - q2, 0.0 * q2)
    if kind == "bosonic":
        return q2 * x ** (par - 1.0) / np.expm1(x ** par)
    if kind == "exponential":
        return k2 * np.exp(-(x ** par))
    if kind == "smoothed_litim":
        # (k^2 - q^2) / (1 + e^z), the logistic factor evaluated without overflow for large z
        z = (x - 1.0) / par
        zr = np.real(z)
        with np.errstate(over="ignore", invalid="ignore"):
            e_neg = np.exp(-np.where(zr > 0, z, 0.0))
            e_pos = np.exp(np.where(zr > 0, 0.0, z))
        return (k2 - q2) * np.where(zr > 0, e_neg / (1.0 + e_neg), 1.0 / (1.0 + e_pos))
    if kind == "polynomial_exp":
        f = sum(x ** c / c for c in range(1, par + 1))
        return k2 * np.exp(-f)
    raise ValueError(kind)


def RF(kind, par, k2, q2):
    return np.sqrt(RB(kind, par, k2, q2) + q2) - np.sqrt(q2)


def t_derivative(f, k2, q2, h=1e-20):
    """d f/dt at fixed q^2, t = ln k (= 2 k^2 df/dk^2), by a complex step in k^2."""
    return 2.0 * k2 * np.imag(f(k2 + 1j * h * k2, q2)) / (h * k2)


def q2_derivative(f, k2, q2, h=1e-20):
    """d f/dq^2 at fixed k^2, by a complex step in q^2."""
    return np.imag(f(k2, q2 + 1j * h * q2)) / (h * q2)


# thermal functions of (x, T)
def coth_s(x, T):
    return 1.0 / np.tanh(x / (2.0 * T))


def tanh_s(x, T):
    return np.tanh(x / (2.0 * T))


def sech_s(x, T):
    return 1.0 / np.cosh(x / (2.0 * T))


def csch_s(x, T):
    return 1.0 / np.sinh(x / (2.0 * T))


def n_b(x, T):
    return 1.0 / np.expm1(x / T)


def n_f(x, T):
    return 1.0 / (np.exp(x / T) + 1.0)
