"""CPU checks that pin the oracle's integrators, integrands and LPA flow.

Measures are pinned by the reference tests' known answers (polynomial and volume integrals,
integrator_*_cpu.cc, integrator_finiteTq0_cpu.cc); integrands by independent continuum
integrals (scipy adaptive quadrature in |q|, not in x = q^2/k^2); the Matsubara closed form of
the flow's threshold integral by an explicit frequency sum; the RK4 stepper by its convergence
order.
"""
import math

import numpy as np
import pytest
from scipy import integrate, special

from closed_forms import tx0_closed
from oracle import oracle as O


def V_d(d, r):
    return math.pi ** (d / 2) / special.gamma(d / 2 + 1) * r ** d


def S_d(d):
    return 2 * math.pi ** (d / 2) / special.gamma(d / 2)


def is_close(a, b, eps):
    """DiFfRG's is_close (math.hh:156-170): absolute OR relative agreement within eps."""
    diff = abs(a - b)
    return diff <= eps or diff <= max(abs(a), abs(b)) * eps


def poly_int(c, lo, hi, extra_power=0):
    """int_lo^hi q^extra_power * sum c_i q^i dq"""
    return sum(ci * (hi ** (i + extra_power + 1) - lo ** (i + extra_power + 1)) / (i + extra_power + 1)
               for i, ci in enumerate(c))


# ----------------------------------------------------------------------------- reference KATs

@pytest.mark.parametrize("d", [1, 2, 3, 4])
def test_radial_known_answers(d):
    rng = np.random.default_rng(100 + d)
    x, w = O.gl_nodes(64)
    for _ in range(4):
        x_ext, k = rng.uniform(1, 2), rng.uniform(0.1, 1)
        q_ext = math.sqrt(x_ext * k * k)
        vol = O.kat_radial(d, x, w, x_ext, k, [1, 0, 0, 0, 0, 0], 0.0)
        assert is_close(vol, V_d(d, q_ext) / (2 * math.pi) ** d, 1e-2 if d == 1 else 1e-6)
        c = rng.uniform(-1, 1, 6)
        if d == 1:
            c[0] = 0.0
        const = rng.uniform(-1, 1)
        ref = const + S_d(d) * poly_int(c, 0, q_ext, d - 1) / (2 * math.pi) ** d
        assert is_close(O.kat_radial(d, x, w, x_ext, k, c, const), ref, 1e-6)
        # even powers: polynomial in x = q^2/k^2 for even d, so GL is exact to rounding
        if d in (2, 4):
            ce = np.array([c[0], 0, c[2], 0, c[4], 0])
            ref = S_d(d) * poly_int(ce, 0, q_ext, d - 1) / (2 * math.pi) ** d
            assert O.kat_radial(d, x, w, x_ext, k, ce, 0.0) == pytest.approx(ref, rel=1e-13)


def test_4d2ang_known_answers():
    rng = np.random.default_rng(7)
    xq, wq = O.gl_nodes(64)
    xa, wa = O.gl_nodes(24)
    for _ in range(4):
        x_ext, k = rng.uniform(1, 2), rng.uniform(0.1, 1)
        q_ext = math.sqrt(x_ext * k * k)
        vol = O.kat_4d2ang(xq, wq, xa, wa, x_ext, k, [1, 0, 0, 0, 0, 0], [1, 0, 0, 0], [1, 0, 0, 0], 0.0)
        assert is_close(vol, V_d(4, q_ext) / (2 * math.pi) ** 4, 1e-6)
        # the sqrt(1 - cos1^2) end points limit the 24-node cos1 rule to ~4e-5 relative
        assert vol == pytest.approx(V_d(4, q_ext) / (2 * math.pi) ** 4, rel=1e-4)
        c, a, b = rng.uniform(-1, 1, 6), rng.uniform(-1, 1, 4), rng.uniform(-1, 1, 4)
        const = rng.uniform(-1, 1)
        ref = const + poly_int(c, 0, q_ext, 3) * ((4 * a[0] + a[2]) * math.pi / 8) * poly_int(b, -1, 1) * 2 * math.pi \
            / (2 * math.pi) ** 4
        assert is_close(O.kat_4d2ang(xq, wq, xa, wa, x_ext, k, c, a, b, const), ref, 1e-4)


def test_4d_known_answers():
    rng = np.random.default_rng(8)
    xq, wq = O.gl_nodes(64)
    xa, wa = O.gl_nodes(24)
    for _ in range(2):
        x_ext, k = rng.uniform(1, 2), rng.uniform(0.1, 1)
        q_ext = math.sqrt(x_ext * k * k)
        one = [1, 0, 0, 0]
        vol = O.kat_4d(xq, wq, xa, wa, x_ext, k, [1, 0, 0, 0, 0, 0], one, one, one, 0.0)
        assert is_close(vol, V_d(4, q_ext) / (2 * math.pi) ** 4, 1e-6)
        assert vol == pytest.approx(V_d(4, q_ext) / (2 * math.pi) ** 4, rel=1e-4)
        c, a, b, p = rng.uniform(-1, 1, 6), rng.uniform(-1, 1, 4), rng.uniform(-1, 1, 4), rng.uniform(-1, 1, 4)
        const = rng.uniform(-1, 1)
        ref = const + poly_int(c, 0, q_ext, 3) * ((4 * a[0] + a[2]) * math.pi / 8) * poly_int(b, -1, 1) \
            * poly_int(p, 0, 2 * math.pi) / (2 * math.pi) ** 4
        assert is_close(O.kat_4d(xq, wq, xa, wa, x_ext, k, c, a, b, p, const), ref, 1e-4)


@pytest.mark.parametrize("d", [2, 3, 4])
def test_finiteT_known_answers(d):
    # integrator_finiteTq0_cpu.cc:36-62: kernel 1/(T^2 + val^2 q0^2)
    rng = np.random.default_rng(20 + d)
    x, w = O.gl_nodes(64)
    for _ in range(6):
        x_ext, T, k, val = rng.uniform(1, 2), rng.uniform(0.01, 1), rng.uniform(0.05, 1), rng.uniform(0.8, 1.2)
        E = k  # the integrator tracks E = k (integrator_finiteTq0_cpu.hh:79-82)
        mx, mw, Tr = O.matsubara_storage_rule(T, E)
        q_ext = math.sqrt(x_ext * k * k)
        ref = V_d(d - 1, q_ext) / (2 * math.pi) ** (d - 1) / T / math.tanh(0.5 / val) / 2 / val
        got = O.kat_finiteT(d, x, w, x_ext, k, mx, mw, T, [1, 0, 0, 0, 0, 0], [T * T, 0, val * val, 0], 0.0)
        assert np.isfinite(got)
        assert is_close(got, ref, 1e-2 if d == 2 else 1e-6)


@pytest.mark.parametrize("d", [2, 3, 4])
def test_angle_known_answers(d):
    # integrator_angle_cpu.cc: volume V_d/(2pi)^d and S_d/2 int poly q^{d-1} dq int cos_poly / (2pi)^d
    rng = np.random.default_rng(30 + d)
    xq, wq = O.gl_nodes(64)
    xa, wa = O.gl_nodes(12)
    for _ in range(4):
        x_ext, k = rng.uniform(1, 2), rng.uniform(0.1, 1)
        q_ext = math.sqrt(x_ext * k * k)
        vol = O.kat_angle(d, xq, wq, xa, wa, x_ext, k, [1, 0, 0, 0, 0, 0], [1, 0, 0, 0], 0.0)
        assert is_close(vol, V_d(d, q_ext) / (2 * math.pi) ** d, 1e-6)
        c, a = rng.uniform(-1, 1, 6), rng.uniform(-1, 1, 4)
        const = rng.uniform(-1, 1)
        ref = const + S_d(d) / 2 * poly_int(c, 0, q_ext, d - 1) / (2 * math.pi) ** d * poly_int(a, -1, 1)
        assert is_close(O.kat_angle(d, xq, wq, xa, wa, x_ext, k, c, a, const), ref, 1e-6)
        if d in (2, 4):  # polynomial in x and cos: exact to rounding
            ce = np.array([c[0], 0, c[2], 0, c[4], 0])
            ref = S_d(d) / 2 * poly_int(ce, 0, q_ext, d - 1) / (2 * math.pi) ** d * poly_int(a, -1, 1)
            assert O.kat_angle(d, xq, wq, xa, wa, x_ext, k, ce, a, 0.0) == pytest.approx(ref, rel=1e-12)


@pytest.mark.parametrize("D", [1, 2, 3])
def test_cartesian_known_answers(D):
    # integrator_{1,2,3}D_cartesian_cpu.cc: volume prod(hi - lo)/(2pi)^D and products of cubics over
    # random boxes (Gauss-Legendre with 32 nodes integrates cubics exactly, so to rounding); uneven
    # node counts and ranges per axis catch a swapped axis, range or weight
    rng = np.random.default_rng(60 + D)
    ns = [32, 20, 12][:D]
    rules = [O.gl_nodes(n) for n in ns]
    for _ in range(4):
        lo, hi = rng.uniform(-2, -1, D), rng.uniform(1, 2, D)
        one = np.zeros(4 * D)
        one[0::4] = 1.0
        vol = O.kat_cartesian(rules, lo, hi, one, 0.0)
        assert vol == pytest.approx(np.prod(hi - lo) / (2 * math.pi) ** D, rel=1e-13)
        c = rng.uniform(-1, 1, 4 * D)
        const = rng.uniform(-1, 1)
        ref = const + np.prod([poly_int(c[4 * i:4 * i + 4], lo[i], hi[i]) for i in range(D)]) / (2 * math.pi) ** D
        got = O.kat_cartesian(rules, lo, hi, c, const)
        assert is_close(got, ref, 1e-6)
        assert got == pytest.approx(ref, rel=1e-12, abs=1e-14)


FREQ_FORMS = [(3, 0, 0), (2, 1, 0), (2, 1, 1), (3, 1, 0), (3, 1, 1)]


def freq_closed(dim, q0_axis, px0, q_ext, ext, k, c, a, p, z, const):
    """integrator_3D_cpu.cc, integrator_{2,3}D{pq0,px0}_cpu.cc: the reference tests' closed forms."""
    Q = k * ext if px0 else ext
    if dim == 2:
        val = S_d(2) / 2 * poly_int(c, 0, q_ext, 1) * poly_int(a, -1, 1) * poly_int(z, -Q, Q) / (2 * math.pi) ** 3
    else:
        val = poly_int(c, 0, q_ext, 2) * poly_int(a, -1, 1) * poly_int(p, 0, 2 * math.pi)
        val /= (2 * math.pi) ** (4 if q0_axis else 3)
        if q0_axis:
            val *= poly_int(z, -Q, Q)
    return const + val


@pytest.mark.parametrize("form", FREQ_FORMS, ids=["3D", "2Dpq0", "2Dpx0", "3Dpq0", "3Dpx0"])
def test_freq_known_answers(form):
    # volumes V_{d}(q_ext) (times 2 Q with a frequency axis) / (2pi)^D and products of polynomials;
    # exact to rounding when the q polynomial makes the radial integrand polynomial in x = q^2/k^2
    # (odd powers for the q-measures of 3 spatial dimensions, even for 2); the frequency axis has
    # its own order, so a swapped table or weight fails
    dim, q0_axis, px0 = form
    rng = np.random.default_rng(80 + 10 * dim + 2 * q0_axis + px0)
    xq, wq = O.gl_nodes(64)
    xa, wa = O.gl_nodes(12)
    xb, wb = O.gl_nodes(10)
    for _ in range(3):
        x_ext, ext, k = rng.uniform(1, 2), rng.uniform(1, 2), rng.uniform(0.1, 1)
        q_ext = math.sqrt(x_ext) * k
        one6, one4 = np.array([1.0, 0, 0, 0, 0, 0]), np.array([1.0, 0, 0, 0])
        vol = O.kat_freq(dim, q0_axis, px0, xq, wq, xa, wa, xb, wb, x_ext, ext, k, one6, one4, one4, one4, 0.0)
        ref = freq_closed(dim, q0_axis, px0, q_ext, ext, k, one6, one4, one4, one4, 0.0)
        assert is_close(vol, ref, 1e-6)
        c, a, p, z = rng.uniform(-1, 1, 6), rng.uniform(-1, 1, 4), rng.uniform(-1, 1, 4), rng.uniform(-1, 1, 4)
        const = rng.uniform(-1, 1)
        got = O.kat_freq(dim, q0_axis, px0, xq, wq, xa, wa, xb, wb, x_ext, ext, k, c, a, p, z, const)
        assert is_close(got, freq_closed(dim, q0_axis, px0, q_ext, ext, k, c, a, p, z, const), 1e-6)
        ce = c.copy()
        ce[slice(0, None, 2) if dim == 3 else slice(1, None, 2)] = 0.0
        got = O.kat_freq(dim, q0_axis, px0, xq, wq, xa, wa, xb, wb, x_ext, ext, k, ce, a, p, z, 0.0)
        assert got == pytest.approx(freq_closed(dim, q0_axis, px0, q_ext, ext, k, ce, a, p, z, 0.0), rel=1e-12)


TX0_FORMS = [(0, 2), (0, 3), (0, 4), (1, 3), (1, 4), (2, 4)]


@pytest.mark.parametrize("form", TX0_FORMS, ids=["1D-d2", "1D-d3", "1D-d4", "angle-d3", "angle-d4", "4D"])
def test_finiteTx0_known_answers(form):
    # integrator_{,angle_,4D_}finiteTx0_cpu.cc: at T > 0 the Matsubara sum is exact and the log-mapped
    # frequency integral of a polynomial converges to rounding, so with radial powers that make the
    # q measure polynomial in x = q^2/k^2 the closed form holds to ~1e-13; at T = 0 the reference
    # test's volume formula at its tolerance (the map starts at 1e-3 and its weight omits the
    # shift, a 1e-5-level effect there)
    kind, d = form
    rng = np.random.default_rng(200 + 10 * kind + d)
    xq, wq = O.gl_nodes(64)
    xa, wa = O.gl_nodes(8)
    ns, nint = 16, 64
    x0, w0 = O.gl_nodes(nint)
    for T in (0.3, 0.05):
        k, x_ext = rng.uniform(0.3, 1), rng.uniform(1, 2)
        x0_ext = ns * 10 * 2 * math.pi * T / k * rng.uniform(1, 2) + 20 / k
        c, a, p, z = (rng.uniform(-1, 1, m) for m in (6, 4, 4, 4))
        c[(0 if (d - 3) % 2 else 1)::2] = 0.0  # powers m with d - 3 + m even
        const = rng.uniform(-1, 1)
        got = O.kat_finiteTx0(kind, d, xq, wq, xa, wa, x0, w0, x_ext, x0_ext, ns, T, k, c, a, p, z, const)
        ref = const + tx0_closed(kind, d, math.sqrt(x_ext) * k, k, x0_ext, ns, T, c, a, p, z)
        assert got == pytest.approx(ref, rel=1e-12, abs=1e-13)
    # T = 0: integrator_finiteTx0_gpu.cu's volume integral, x0_extent = 1000 / k
    if kind == 0 and d > 2:
        k, x_ext = 0.5, 1.5
        x0_ext = 1000 / k * 1.3
        x0, w0 = O.gl_nodes(nint + ns)
        one6, one4 = np.array([1.0, 0, 0, 0, 0, 0]), np.array([1.0, 0, 0, 0])
        got = O.kat_finiteTx0(0, d, xq, wq, xa, wa, x0, w0, x_ext, x0_ext, ns, 0.0, k, one6, one4, one4, one4, 0.0)
        ref = V_d(d - 1, math.sqrt(x_ext) * k) / (2 * math.pi) ** (d - 1) * (k * x0_ext * 2 / (2 * math.pi))
        assert is_close(got, ref, 5e-5)


def _sum_closed(T, val):
    # T sum_n 1/(T^2 + val^2 (2 pi n T)^2) = coth(1/(2 val)) / (2 val T)
    return 1 / T / math.tanh(0.5 / val) / 2 / val


def test_4d_finiteT_known_answers():
    # integrator_4D_finiteTq0_cpu.cc:42-71 (volume x Matsubara sum, tolerance 5e-5), extended to
    # polynomials in q, cos and phi, whose spatial integral factorises from the sum
    rng = np.random.default_rng(44)
    xq, wq = O.gl_nodes(64)
    xa, wa = O.gl_nodes(16)
    for _ in range(4):
        x_ext, T, k, val = rng.uniform(1, 2), rng.uniform(0.01, 1), rng.uniform(0.05, 1), rng.uniform(0.8, 1.2)
        mx, mw, _ = O.matsubara_storage_rule(T, k)
        q_ext = math.sqrt(x_ext * k * k)
        den = [T * T, 0, val * val, 0, 0]
        one = [1, 0, 0, 0]
        vol = O.kat_4d_finiteT(xq, wq, xa, wa, x_ext, k, mx, mw, T, [1, 0, 0, 0, 0, 0], one, one, den, 0.0)
        assert is_close(vol, V_d(3, q_ext) / (2 * math.pi) ** 3 * _sum_closed(T, val), 5e-5)
        c, a, p = rng.uniform(-1, 1, 6), rng.uniform(-1, 1, 4), rng.uniform(-1, 1, 4)
        const = rng.uniform(-1, 1)
        ref = const + poly_int(c, 0, q_ext, 2) * poly_int(a, -1, 1) * poly_int(p, 0, 2 * math.pi) \
            / (2 * math.pi) ** 3 * _sum_closed(T, val)
        got = O.kat_4d_finiteT(xq, wq, xa, wa, x_ext, k, mx, mw, T, c, a, p, den, const)
        assert is_close(got, ref, 5e-5)


@pytest.mark.parametrize("d", [3, 4])
def test_angle_finiteT_known_answers(d):
    # integrator_angle_finiteTq0_gpu.cu:36-56 (volume x sum), extended to poly(q) poly(cos)
    rng = np.random.default_rng(50 + d)
    xq, wq = O.gl_nodes(64)
    xa, wa = O.gl_nodes(16)
    for _ in range(4):
        x_ext, T, k, val = rng.uniform(1, 2), rng.uniform(0.01, 1), rng.uniform(0.05, 1), rng.uniform(0.8, 1.2)
        mx, mw, _ = O.matsubara_storage_rule(T, k)
        q_ext = math.sqrt(x_ext * k * k)
        den = [T * T, 0, val * val, 0]
        vol = O.kat_angle_finiteT(d, xq, wq, xa, wa, x_ext, k, mx, mw, T, [1, 0, 0, 0, 0, 0], [1, 0, 0, 0], den, 0.0)
        assert is_close(vol, V_d(d - 1, q_ext) / (2 * math.pi) ** (d - 1) * _sum_closed(T, val), 5e-5)
        c, a = rng.uniform(-1, 1, 6), rng.uniform(-1, 1, 4)
        const = rng.uniform(-1, 1)
        ref = const + S_d(d - 1) / 2 * poly_int(c, 0, q_ext, d - 2) * poly_int(a, -1, 1) \
            / (2 * math.pi) ** (d - 1) * _sum_closed(T, val)
        got = O.kat_angle_finiteT(d, xq, wq, xa, wa, x_ext, k, mx, mw, T, c, a, den, const)
        assert is_close(got, ref, 5e-5)


# ----------------------------------------------------------------------------- coordinates, lerp

def test_log_coordinates_and_lerp():
    c = O.logcoord(128, 1e-3, 40.0, 6.0)
    assert O.logcoord_forward(c, 0) == pytest.approx(1e-3, rel=1e-15)
    assert O.logcoord_forward(c, 127) == pytest.approx(40.0, rel=1e-14)
    for i in [0, 1, 17, 63, 126, 127]:
        assert O.logcoord_backward(c, O.logcoord_forward(c, i)) == pytest.approx(i, abs=1e-10)
    # data linear in the index is reproduced exactly (linear_interpolation_1D_cpu.cc:28-38)
    data = np.arange(128, dtype=float) * 0.5 + 3.0
    for y in np.geomspace(2e-3, 39.0, 37):
        assert O.lerp_log(data, c, y) == pytest.approx(0.5 * O.logcoord_backward(c, y) + 3.0, rel=1e-13)
    assert O.lerp_log(data, c, 0.0) == 3.0          # clamped below
    assert O.lerp_log(data, c, 100.0) == data[-1]   # clamped above
    assert O.lerp_lin(np.arange(11.0), 0.0, 1.0, 0.35) == pytest.approx(3.5, rel=1e-14)


def test_regulators_closed_forms():
    for k2, q2 in [(1.0, 0.3), (0.25, 2.0), (2.0, 0.05)]:
        for b in (1.0, 2.0):
            xb = (q2 / k2) ** b
            assert O.exp_RB(b, k2, q2) == pytest.approx(k2 * math.exp(-xb), rel=1e-15)
            assert O.exp_RBdot(b, k2, q2) == pytest.approx(2 * k2 * math.exp(-xb) * (1 + b * xb), rel=1e-15)
            # RBdot = k dR/dk at fixed q (numerical derivative in k)
            h = 1e-6
            kk = math.sqrt(k2)
            dk = (O.exp_RB(b, (kk * (1 + h)) ** 2, q2) - O.exp_RB(b, (kk * (1 - h)) ** 2, q2)) / (2 * h)
            assert O.exp_RBdot(b, k2, q2) == pytest.approx(dk, rel=1e-8)
            dq = (O.exp_RB(b, k2, q2 * (1 + h)) - O.exp_RB(b, k2, q2 * (1 - h))) / (2 * h * q2)
            assert O.exp_dq2RB(b, k2, q2) == pytest.approx(dq, rel=1e-7)
    assert O.litim_RB(1.0, 0.3) == pytest.approx(0.7)
    assert O.litim_RB(1.0, 1.3) == 0.0
    assert O.litim_RBdot(1.0, 0.3) == 2.0 and O.litim_RBdot(1.0, 2.0) == 0.0


# ----------------------------------------------------------------------------- built-in integrands

def _R(q2):
    return math.exp(-q2)


def _Rdot(q2):
    return 2 * math.exp(-q2) * (1 + q2)


def test_threshold_integrand_vs_continuum():
    x, w = O.gl_nodes(64)
    m2 = np.array([-0.8, -0.3, 0.0, 0.7, 4.0])
    got = O.c0_threshold(m2, x, w, 20.0, 1.0, 1.0)
    for i, m in enumerate(m2):
        f = lambda q: S_d(4) / (2 * math.pi) ** 4 * q ** 3 * _Rdot(q * q) / (q * q + _R(q * q) + m) ** 2
        ref, _ = integrate.quad(f, 0, math.sqrt(20.0), epsabs=0, epsrel=1e-13, limit=200)
        assert got[i] == pytest.approx(ref, rel=1e-11)


def test_selfenergy_integrand_vs_continuum():
    # constant dressing Z = 1.25 so that G is analytic; the cos1 rule is limited by the
    # sqrt(1-cos1^2) end points, so compare at 1e-5 with 64 angular nodes
    xq, wq = O.gl_nodes(64)
    xa, wa = O.gl_nodes(64)
    Zc = 1.25
    tab = np.full(16, Zc)
    coord = O.logcoord(16, 1e-3, 40.0, 6.0)
    for p in (0.05, 1.3, 7.0):
        got = O.c1_selfenergy([p], [tab], [coord], [(xq, wq)], [(xa, wa)], 20.0)[0]
        G = lambda y2: 1.0 / (Zc * y2 + _R(y2))

        def inner(q):
            g = lambda c: math.sqrt(1 - c * c) * G(p * p + q * q - 2 * p * q * c)
            val, _ = integrate.quad(g, -1, 1, epsabs=0, epsrel=1e-12)
            return q ** 3 * _Rdot(q * q) * G(q * q) ** 2 * val

        outer, _ = integrate.quad(inner, 0, math.sqrt(20.0), epsabs=0, epsrel=1e-11, limit=200)
        ref = 2 * math.pi * 2.0 * outer / (2 * math.pi) ** 4  # phi -> 2 pi, cos2 -> 2
        assert got == pytest.approx(ref, rel=2e-5)


def test_finiteT_integrand_vs_continuum():
    x, w = O.gl_nodes(48)
    T, p0, pa, m1, m2 = 0.1, 2 * math.pi * 0.1 * 2, 0.7, -0.3, 1.1
    n_max = 64
    mx, mw = 2 * math.pi * T * np.arange(1, n_max + 1), np.full(n_max, T)
    got = O.c2_finiteT([T], [p0], [pa], [m1], [m2], [(mx, mw)], x, w, 6.0, 1.0, 1.0, 4, 1)[0]
    G = lambda wn, s, m: 1.0 / (wn * wn + s + _R(s) + m)
    ref = 0.0
    for n in range(-n_max, n_max + 1):
        wn = 2 * math.pi * n * T
        f = lambda q: 4 * math.pi * q * q / (2 * math.pi) ** 3 * _Rdot(q * q) * G(wn, q * q, m1) ** 2 \
            * G(wn + p0, q * q + pa * pa, m2)
        v, _ = integrate.quad(f, 0, 6.0, epsabs=0, epsrel=1e-13, limit=200)
        ref += T * v
    assert got == pytest.approx(ref, rel=1e-10)


# ----------------------------------------------------------------------------- LPA flow

def test_bose_sum_closed_form_vs_explicit_sum():
    # I_T uses S2(E) = T sum_n (w_n^2 + E^2)^-2 in closed form.  A one-node rule isolates S2
    # at that node; compare with the explicit frequency sum.
    n = np.arange(-400000, 400001)
    for T in (0.01, 0.07, 0.3):
        wn2 = (2 * math.pi * T * n) ** 2
        for m2 in (-0.5, 0.2, 3.0):
            for x0 in (0.01, 0.3, 0.9):
                x_ext, k = 20.0, 1.0
                q2 = x0 * x_ext
                q = math.sqrt(q2)
                pref = 4 * math.pi * (q / 2 * k * k) / (2 * math.pi) ** 3 * x_ext * _Rdot(q2)
                S2 = O.lpa_threshold(m2, T, k, [x0], [1.0], x_ext) / pref
                E2 = q2 + _R(q2) + m2
                ref = T * np.sum(1.0 / (wn2 + E2) ** 2)
                assert S2 == pytest.approx(ref, rel=1e-12)
    # and the 64-node radial rule against the continuum (the x = q^2 map of the d = 3 element
    # leaves a sqrt(x) end point, so Gauss-Legendre converges algebraically: ~1e-5 here)
    x, w = O.gl_nodes(64)
    for T, m2 in ((0.05, 0.2), (0.3, -0.4)):
        got = O.lpa_threshold(m2, T, 1.0, x, w, 20.0)

        def f(q):
            E = math.sqrt(q * q + _R(q * q) + m2)
            S2 = 1 / (math.tanh(E / (2 * T)) * 4 * E ** 3) + 1 / (8 * T * E * E * math.sinh(E / (2 * T)) ** 2)
            return 4 * math.pi * q * q / (2 * math.pi) ** 3 * _Rdot(q * q) * S2

        ref, _ = integrate.quad(f, 0, math.sqrt(20.0), epsabs=0, epsrel=1e-12, limit=100)
        assert got == pytest.approx(ref, rel=1e-4)


def test_flow_rhs_for_linear_potential():
    x, w = O.gl_nodes(64)
    fp = O.flow_params(33, 4.0, 0.1, 0.0, 4.0, 0.2, 1.0)
    rho = np.linspace(0, 4.0, 33)
    u = 0.2 + 0.5 * rho
    k = 0.6
    got = O.flow_rhs(fp, u, k, x, w)
    for i in (0, 5, 32):
        Ipi = O.lpa_threshold(u[i], 0.1, k, x, w)
        Isg = O.lpa_threshold(u[i] + 2 * rho[i] * 0.5, 0.1, k, x, w)
        assert got[i] == pytest.approx(0.5 * (3.0 * 0.5 * Ipi + 1.5 * Isg), rel=1e-12)


def test_flow_constant_potential_is_fixed_point():
    x, w = O.gl_nodes(16)
    fp = O.flow_params(16, 4.0, 0.2, 0.1, 4.0, 0.3, 0.0)
    u = O.flow_lpa(fp, x, w, k_end=0.1, n_steps=20)
    assert np.all(u == 0.3)


def test_flow_rk4_fourth_order():
    x, w = O.gl_nodes(32)
    fp = O.flow_params(24, 4.0, 0.1, 0.0, 4.0, 0.2, 1.0)
    u1 = O.flow_lpa(fp, x, w, k_end=0.5, n_steps=5)
    u2 = O.flow_lpa(fp, x, w, k_end=0.5, n_steps=10)
    u4 = O.flow_lpa(fp, x, w, k_end=0.5, n_steps=20)
    e1, e2 = np.max(np.abs(u1 - u4)), np.max(np.abs(u2 - u4))
    ratio = (e1 - e2) / max(e2, 1e-300)
    assert 8.0 < e1 / e2 < 40.0, (e1, e2)
