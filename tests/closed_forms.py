"""Closed forms shared by the CPU oracle tests and the GPU parity tests."""
import math

import numpy as np
from scipy import special


def S_d(d):
    return 2 * math.pi ** (d / 2) / special.gamma(d / 2)


def poly_int(c, lo, hi, extra_power=0):
    """int_lo^hi q^extra_power * sum c_i q^i dq"""
    return sum(ci * (hi ** (i + extra_power + 1) - lo ** (i + extra_power + 1)) / (i + extra_power + 1)
               for i, ci in enumerate(c))


def tx0_closed(kind, d, q_ext, k, x0_ext, ns, T, c, a, p, z):
    """Finite-T x0 forms at T > 0: the spatial integral times T sum_{|n|<ns} f(2 pi n T) plus
    (2pi)^-1 int_{2 pi ns T <= |q0| <= k x0_ext} f(q0) dq0."""
    if kind == 2:
        spatial = poly_int(c, 0, q_ext, 2) * poly_int(a, -1, 1) * poly_int(p, 0, 2 * math.pi)
    elif kind == 1:
        spatial = S_d(d - 1) / 2 * poly_int(c, 0, q_ext, d - 2) * poly_int(a, -1, 1)
    else:
        spatial = S_d(d - 1) * poly_int(c, 0, q_ext, d - 2)
    f = np.polynomial.polynomial.polyval
    Q, q0s = k * x0_ext, 2 * math.pi * ns * T
    fs = T * sum(f(2 * math.pi * n * T, z) for n in range(-(ns - 1), ns))
    fi = (poly_int(z, q0s, Q) + poly_int(z, -Q, -q0s)) / (2 * math.pi)
    return spatial / (2 * math.pi) ** (d - 1) * (fs + fi)
