"""The DiFfRG-compatible C++ integrator templates (include/DiFfRG/...) with user KERNEL structs,
compiled by hipcc into tests/native/libdfrg_kat.so and called through its C wrappers.

Each GPU template is compared with the oracle on the same Gauss-Legendre nodes (relative error
<= 1e-12) and with the reference tests' closed forms using the reference's is_close semantics;
the host (TBB-named) templates are checked the same way.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest
from scipy import special

import diffrg_amd.workloads as W
from closed_forms import tx0_closed
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KAT = os.path.join(ROOT, "tests", "native", "libdfrg_kat.so")
TOL = 1e-12
_dp = C.POINTER(C.c_double)


@pytest.fixture(scope="module")
def kat(gpu):
    lib = C.CDLL(KAT)
    lib.kat_last_error.restype = C.c_char_p
    return lib


def arr(v):
    return np.ascontiguousarray(v, dtype=np.float64)


def P(a):
    return a.ctypes.data_as(_dp)


def call(kat, name, *args):
    rc = getattr(kat, name)(*args)
    assert rc == 0, (name, rc, kat.kat_last_error())


def is_close(a, b, eps):
    diff = abs(a - b)
    return diff <= eps or diff <= max(abs(a), abs(b)) * eps


def V_d(d, r):
    return math.pi ** (d / 2) / special.gamma(d / 2 + 1) * r ** d


def S_d(d):
    return 2 * math.pi ** (d / 2) / special.gamma(d / 2)


def poly_int(c, lo, hi, p=0):
    return sum(ci * (hi ** (i + p + 1) - lo ** (i + p + 1)) / (i + p + 1) for i, ci in enumerate(c))


D = C.c_double


@pytest.mark.parametrize("d", [1, 2, 3, 4])
def test_radial_template(kat, qp, d):
    rng = np.random.default_rng(40 + d)
    x, w = qp.get_points(64), qp.get_weights(64)
    for device in (1, 0):
        x_ext, k = rng.uniform(1, 2), rng.uniform(0.1, 1)
        c = arr(rng.uniform(-1, 1, 6))
        const = rng.uniform(-1, 1)
        print(f"[case] radial d={d} device={device}", flush=True)
        out = D()
        call(kat, "kat_radial", device, d, 64, D(x_ext), D(k), P(c), D(const), C.byref(out))
        ref = O.kat_radial(d, x, w, x_ext, k, c, const)
        assert out.value == pytest.approx(ref, rel=TOL, abs=TOL * abs(const))
        q_ext = math.sqrt(x_ext) * k
        closed = const + S_d(d) * poly_int(c, 0, q_ext, d - 1) / (2 * math.pi) ** d
        if d > 1:
            assert is_close(out.value, closed, 1e-6)


def test_radial_complex_and_requests(kat, qp):
    c = arr([0.3, -0.2, 0.5, 0.1, -0.4, 0.2])
    re, im = D(), D()
    print("[case] radial complex", flush=True)
    call(kat, "kat_radial_complex", 64, D(1.5), D(0.7), P(c), D(0.25), C.byref(re), C.byref(im))
    x, w = qp.get_points(64), qp.get_weights(64)
    ref = O.kat_radial(4, x, w, 1.5, 0.7, c, 0.0)
    assert re.value == pytest.approx(0.25 + ref, rel=TOL)
    assert im.value == pytest.approx(2 * ref, rel=TOL)
    ks = arr(np.linspace(0.2, 1.0, 9))
    out = np.empty(9)
    print("[case] radial futures", flush=True)
    call(kat, "kat_radial_requests", 64, D(1.5), P(ks), 9, P(c), P(out))
    for i, k in enumerate(ks):
        assert out[i] == pytest.approx(O.kat_radial(4, x, w, 1.5, k, c, 0.0), rel=TOL)


@pytest.mark.parametrize("n", [20000, 65536 + 100])
def test_many_requests_in_one_batch(kat, qp, n):
    # n futures read only after all are requested: one batch, its tuple and result buffers grown
    # well past any earlier batch of the integrator; beyond 65,536 requests the full batch is
    # launched at once and the rest go to the next batch
    ks = arr(np.linspace(0.05, 2.0, n))
    c = arr([0.3, -0.2, 0.5, 0.1, -0.4, 0.2])
    out = np.empty(n)
    print(f"[case] {n} radial requests in one batch", flush=True)
    call(kat, "kat_radial_requests", 64, D(1.5), P(ks), n, P(c), P(out))
    x, w = qp.get_points(64), qp.get_weights(64)
    ref = np.array([O.kat_radial(4, x, w, 1.5, k, c, 0.0) for k in ks])
    assert np.max(np.abs(out - ref)) <= TOL * np.max(np.abs(ref))  # ref may cross zero in k


def test_radial_template_float(kat, qp):
    # NT = ctype = float: FP32 nodes and arithmetic, against the FP64 oracle on the same order
    x, w = qp.get_points(64), qp.get_weights(64)
    c = arr([0.3, -0.2, 0.5, 0.1, -0.4, 0.2])
    g, r = D(), D()
    print("[case] radial float", flush=True)
    call(kat, "kat_radial_float", 64, C.c_float(1.5), C.c_float(0.7), P(c), C.c_float(0.25), C.byref(g), C.byref(r))
    ref = 0.25 + O.kat_radial(4, x, w, 1.5, 0.7, c, 0.0)
    assert g.value == pytest.approx(ref, rel=2e-6)
    assert r.value == g.value


def test_radial_template_complex_float(kat, qp):
    # NT = complex<float> (reference NT list, types.hh:130-167): both components on FP32 nodes
    c = arr([0.3, -0.2, 0.5, 0.1, -0.4, 0.2])
    x, w = qp.get_points(64), qp.get_weights(64)
    out = np.empty(4)
    print("[case] radial complex<float>", flush=True)
    call(kat, "kat_radial_complex_float", 64, C.c_float(1.5), C.c_float(0.7), P(c), C.c_float(0.25), P(out))
    ref = O.kat_radial(4, x, w, 1.5, 0.7, c, 0.0)
    assert out[0] == pytest.approx(0.25 + ref, rel=2e-6) and out[1] == pytest.approx(2 * ref, rel=2e-6)
    assert out[2] == out[0] and out[3] == out[1]


def test_radial_template_dual(kat, qp):
    # NT = a dual number (the reference's autodiff::real, types.hh:130-167): both channels are
    # integrated; each must match the oracle's radial integral of its polynomial, on the GPU and
    # TBB templates, for get() and a batched request()
    rng = np.random.default_rng(5)
    x, w = qp.get_points(64), qp.get_weights(64)
    f, g = arr(rng.uniform(-1, 1, 6)), arr(rng.uniform(-1, 1, 6))
    for tbb in (0, 1):
        out = np.empty(4)
        print(f"[case] radial dual tbb={tbb}", flush=True)
        call(kat, "kat_radial_dual", 64, D(1.5), D(0.7), P(f), P(g), tbb, P(out))
        want_f = f[0] + O.kat_radial(4, x, w, 1.5, 0.7, f, 0.0)
        want_g = g[0] + O.kat_radial(4, x, w, 1.5, 0.7, g, 0.0)
        assert out[0] == pytest.approx(want_f, rel=TOL) and out[1] == pytest.approx(want_g, rel=TOL)
        assert out[2] == out[0] and out[3] == out[1]


def test_concurrent_callers(kat):
    # 8 host threads share one radial and one two-angle integrator, mixing get() and request()
    # (the reference's get_data calls get() from TBB threads): every result bitwise equals the same
    # call made alone
    rng = np.random.default_rng(12)
    n = 64
    ks = arr(rng.uniform(0.2, 1.0, n))
    c, a, b = arr(rng.uniform(-1, 1, 6)), arr(rng.uniform(-1, 1, 4)), arr(rng.uniform(-1, 1, 4))
    conc, seq = np.empty(2 * n), np.empty(2 * n)
    print("[case] 8 threads x 64 calls on shared integrators", flush=True)
    call(kat, "kat_concurrent", 8, P(ks), n, P(c), P(a), P(b), P(conc), P(seq))
    assert np.array_equal(conc, seq)


def test_two_angle_template(kat, qp):
    rng = np.random.default_rng(7)
    xq, wq = qp.get_points(64), qp.get_weights(64)
    xa, wa = qp.get_points(24), qp.get_weights(24)
    for device in (1, 0):
        x_ext, k = rng.uniform(1, 2), rng.uniform(0.1, 1)
        c, a, b = arr(rng.uniform(-1, 1, 6)), arr(rng.uniform(-1, 1, 4)), arr(rng.uniform(-1, 1, 4))
        const = rng.uniform(-1, 1)
        print(f"[case] 2-angle device={device}", flush=True)
        out = D()
        call(kat, "kat_2ang", device, 64, 24, 24, D(x_ext), D(k), P(c), P(a), P(b), D(const), C.byref(out))
        ref = O.kat_4d2ang(xq, wq, xa, wa, x_ext, k, c, a, b, const)
        assert out.value == pytest.approx(ref, rel=TOL, abs=TOL * abs(const))
        q_ext = math.sqrt(x_ext) * k
        closed = const + poly_int(c, 0, q_ext, 3) * ((4 * a[0] + a[2]) * math.pi / 8) * poly_int(b, -1, 1) \
            * 2 * math.pi / (2 * math.pi) ** 4
        assert is_close(out.value, closed, 1e-4)


def test_three_angle_template(kat, qp):
    rng = np.random.default_rng(8)
    xq, wq = qp.get_points(32), qp.get_weights(32)
    xa, wa = qp.get_points(12), qp.get_weights(12)
    for device in (1, 0):
        x_ext, k = rng.uniform(1, 2), rng.uniform(0.1, 1)
        c, a, b, p = (arr(rng.uniform(-1, 1, n)) for n in (6, 4, 4, 4))
        const = rng.uniform(-1, 1)
        print(f"[case] 3-angle device={device}", flush=True)
        out = D()
        call(kat, "kat_3ang", device, 32, 12, D(x_ext), D(k), P(c), P(a), P(b), P(p), D(const), C.byref(out))
        ref = O.kat_4d(xq, wq, xa, wa, x_ext, k, c, a, b, p, const)
        assert out.value == pytest.approx(ref, rel=TOL, abs=TOL * abs(const))


@pytest.mark.parametrize("nx,na", [(37, 13), (5, 3), (129, 7), (300, 11), (33, 16), (3, 41)])
def test_angle_templates_odd_sizes(kat, qp, nx, na):
    # Orders whose node counts do not divide the thread stride, so every digit of a thread's
    # node walk carries (the stride's lower digits are nonzero); a few sizes give more nodes than
    # one pass of the grid, so threads take several strides.  Unit plans: rows of fewer than 256
    # inner nodes are grouped (flat units), 256 or more make a unit per radial node (na = 16:
    # exactly 256 for the 2-angle form), and more than 65,536 split each row (na = 41: 68,921
    # inner nodes of the 3-angle form, two units per radial node).
    rng = np.random.default_rng(nx * 100 + na)
    xq, wq = qp.get_points(nx), qp.get_weights(nx)
    xa, wa = qp.get_points(na), qp.get_weights(na)
    x_ext, k = rng.uniform(1, 2), rng.uniform(0.1, 1)
    c, a, b, p = (arr(rng.uniform(-1, 1, n)) for n in (6, 4, 4, 4))
    const = rng.uniform(-1, 1)
    print(f"[case] 2- and 3-angle templates nx={nx} na={na}", flush=True)
    out = D()
    call(kat, "kat_2ang", 1, nx, na, na, D(x_ext), D(k), P(c), P(a), P(b), D(const), C.byref(out))
    ref = O.kat_4d2ang(xq, wq, xa, wa, x_ext, k, c, a, b, const)
    assert out.value == pytest.approx(ref, rel=TOL, abs=TOL * abs(const))
    call(kat, "kat_3ang", 1, nx, na, D(x_ext), D(k), P(c), P(a), P(b), P(p), D(const), C.byref(out))
    ref = O.kat_4d(xq, wq, xa, wa, x_ext, k, c, a, b, p, const)
    assert out.value == pytest.approx(ref, rel=TOL, abs=TOL * abs(const))


@pytest.mark.parametrize("d", [2, 3, 4])
def test_finiteT_template(kat, qp, d):
    rng = np.random.default_rng(60 + d)
    x, w = qp.get_points(64), qp.get_weights(64)
    for device in (1, 0):
        x_ext, T, k, val = rng.uniform(1, 2), rng.uniform(0.01, 1), rng.uniform(0.05, 1), rng.uniform(0.8, 1.2)
        c = arr([1, 0, 0, 0, 0, 0])
        den = arr([T * T, 0, val * val, 0])
        print(f"[case] finite-T d={d} device={device} T={T:.3f} k={k:.3f}", flush=True)
        out, nm = D(), C.c_int()
        call(kat, "kat_finiteT", device, d, 64, D(x_ext), D(T), D(k), P(c), P(den), D(0.0), C.byref(out), C.byref(nm))
        # E tracks k (2.5% hysteresis) from the default 10 T; the rule is the provider's at (T, E = k)
        E = k if abs(k - 10 * T) / max(k, 10 * T) > 2.5e-2 else 10 * T
        mx, mw, _ = O.matsubara_storage_rule(T, E)
        ref = O.kat_finiteT(d, x, w, x_ext, k, mx, mw, T, c, den, 0.0)
        if device:
            assert nm.value == len(mx)
        assert out.value == pytest.approx(ref, rel=1e-11)  # Monien rules agree to ~1e-11 (QL vs bisection)
        q_ext = math.sqrt(x_ext) * k
        closed = V_d(d - 1, q_ext) / (2 * math.pi) ** (d - 1) / T / math.tanh(0.5 / val) / 2 / val
        assert is_close(out.value, closed, 1e-2 if d == 2 else 1e-6)


def test_finiteT_energy_tracking(kat):
    # IntegratorFiniteTq0GPU::get: E follows k when it moves by more than 2.5% (integrator_finiteTq0_gpu.hh:278-281)
    T = 0.05
    ks = arr([1.0, 0.99, 0.97, 0.9, 0.89, 0.5, 0.1])
    nm = (C.c_int * len(ks))()
    E = np.empty(len(ks))
    print("[case] finite-T E tracking", flush=True)
    call(kat, "kat_finiteT_track", 32, D(1.5), D(T), P(ks), len(ks), nm, P(E))
    expect, e = [], 10 * T
    for k in ks:
        if abs(k - e) / max(k, e) > 2.5e-2:
            e = k
        expect.append(e)
    assert np.allclose(E, expect)
    for i, k in enumerate(ks):
        assert nm[i] == len(O.matsubara_storage_rule(T, E[i])[0])


def test_builtin_integrands_as_user_kernels(kat, qp):
    # threshold
    m2 = arr(W.threshold(32)["m2"])
    out = np.empty(32)
    print("[case] threshold KERNEL", flush=True)
    call(kat, "kat_threshold", 64, D(20.0), D(1.0), P(m2), 32, P(out))
    x, w = qp.get_points(64), qp.get_weights(64)
    assert np.max(np.abs(out - O.c0_threshold(m2, x, w)) / O.c0_threshold(m2, x, w)) <= TOL
    # self-energy with LinearInterpolator1D<double, LogarithmicCoordinates1D<double>>
    c = W.selfenergy(n_p=8, nq=32, na=16)
    p, Z = arr(c["p"]), arr(c["Z"])
    out = np.empty(8)
    print("[case] self-energy KERNEL", flush=True)
    call(kat, "kat_selfenergy", 32, 16, D(20.0), D(1.0), P(p), 8, P(Z), len(Z), D(1e-3), D(40.0), D(6.0), P(out))
    coord = O.logcoord(len(Z), 1e-3, 40.0, 6.0)
    ref = O.c1_selfenergy(p, [Z] * 8, [coord] * 8, [(qp.get_points(32), qp.get_weights(32))] * 8,
                          [(qp.get_points(16), qp.get_weights(16))] * 8)
    assert np.max(np.abs(out - ref) / ref) <= TOL
    # finite-T two-propagator, reference x-mapping and Monien rule at E = k = 1
    f = W.finite_T(n=10)
    out = np.empty(10)
    print("[case] finite-T KERNEL", flush=True)
    call(kat, "kat_finiteT_twoprop", 48, D(20.0), D(1.0), P(arr(f["T"])), P(arr(f["p0"])), P(arr(f["pabs"])),
         P(arr(f["m1"])), P(arr(f["m2"])), 10, P(out))
    mats = [O.matsubara_storage_rule(t, 1.0)[:2] for t in f["T"]]
    ref = O.c2_finiteT(f["T"], f["p0"], f["pabs"], f["m1"], f["m2"], mats, qp.get_points(48), qp.get_weights(48),
                       20.0, 1.0, 1.0, 4, 0)
    assert np.max(np.abs(out - ref) / ref) <= 1e-11


@pytest.mark.parametrize("d", [2, 3, 4])
def test_angle_template(kat, qp, d):
    # IntegratorAngleGPU/TBB<d> (integrator_angle_gpu.cu): oracle parity and the polynomial closed form
    rng = np.random.default_rng(70 + d)
    xq, wq = qp.get_points(64), qp.get_weights(64)
    xa, wa = qp.get_points(12), qp.get_weights(12)
    for device in (1, 0):
        x_ext, k = rng.uniform(1, 2), rng.uniform(0.1, 1)
        c, a = arr(rng.uniform(-1, 1, 6)), arr(rng.uniform(-1, 1, 4))
        const = rng.uniform(-1, 1)
        print(f"[case] angle d={d} device={device}", flush=True)
        out = D()
        call(kat, "kat_angle", device, d, 64, 12, D(x_ext), D(k), P(c), P(a), D(const), C.byref(out))
        ref = O.kat_angle(d, xq, wq, xa, wa, x_ext, k, c, a, const)
        assert out.value == pytest.approx(ref, rel=TOL, abs=TOL * abs(const))
        q_ext = math.sqrt(x_ext) * k
        closed = const + S_d(d) / 2 * poly_int(c, 0, q_ext, d - 1) / (2 * math.pi) ** d * poly_int(a, -1, 1)
        assert is_close(out.value, closed, 1e-6)


_dint = C.POINTER(C.c_int)
CART_SHAPES = [(32,), (7,), (300,), (32, 32), (20, 12), (16, 300), (32, 32, 32), (24, 20, 12)]


def _cart_closed(c, lo, hi, const):
    nd = len(lo)
    return const + np.prod([poly_int(c[4 * i:4 * i + 4], lo[i], hi[i]) for i in range(nd)]) / (2 * math.pi) ** nd


def _cart(kat, device, req, n, lo, hi, k, const, c):
    out = D()
    nn = np.ascontiguousarray(n, dtype=np.int32)
    call(kat, "kat_cartesian", device, req, len(n), nn.ctypes.data_as(_dint), P(lo), P(hi), D(k), D(const), P(c),
         C.byref(out))
    return out.value


@pytest.mark.parametrize("n", CART_SHAPES, ids=lambda n: "x".join(map(str, n)))
def test_cartesian_template(kat, qp, n):
    # Integrator{1,2,3}DCartesianGPU/TBB (integrator_{1,2,3}D_cartesian_gpu.cu): oracle parity, the
    # closed form of the cubic products, and request() bitwise get().  The shapes cover the flat
    # units (several when an axis has more than 256 nodes) and the row units (qy qz planes of 256
    # and more nodes), with node counts and ranges differing per axis.
    nd = len(n)
    rng = np.random.default_rng(90 + sum(n))
    rules = [(qp.get_points(m), qp.get_weights(m)) for m in n]
    for device in (1, 0):
        lo, hi = arr(rng.uniform(-2, -1, nd)), arr(rng.uniform(1, 2, nd))
        c = arr(rng.uniform(-1, 1, 4 * nd))
        const, k = rng.uniform(-1, 1), rng.uniform(0, 1)
        print(f"[case] cartesian n={n} device={device}", flush=True)
        got = _cart(kat, device, 0, n, lo, hi, k, const, c)
        ref = O.kat_cartesian(rules, lo, hi, c, const)
        assert got == pytest.approx(ref, rel=TOL, abs=TOL * abs(const))
        assert is_close(got, _cart_closed(c, lo, hi, const), 1e-6)
        one = np.zeros(4 * nd)
        one[0::4] = 1.0
        vol = _cart(kat, device, 0, n, lo, hi, k, 0.0, arr(one))
        assert is_close(vol, np.prod(hi - lo) / (2 * math.pi) ** nd, 1e-6)
        if device:
            assert _cart(kat, device, 1, n, lo, hi, k, const, c) == got  # request: bitwise get()


@pytest.mark.parametrize("nd", [1, 2, 3])
def test_cartesian_range_mutators(kat, qp, nd):
    # set_q?_min / set_q?_max (integrator_2D_cartesian_gpu.hh:144-171): a request keeps the range it
    # was made with, later calls and requests use the new one, copies made before keep theirs, and
    # set_q?_max before set_q?_min gives the same box
    rng = np.random.default_rng(120 + nd)
    n = [32, 20, 12][:nd]
    nn = np.ascontiguousarray(n, dtype=np.int32)
    lo, hi = arr(rng.uniform(-2, -1, nd)), arr(rng.uniform(1, 2, nd))
    lo2, hi2 = arr(rng.uniform(-3, -2, nd)), arr(rng.uniform(0.5, 1, nd))
    c = arr(rng.uniform(-1, 1, 4 * nd))
    k = 0.5
    out = np.zeros(6)
    print(f"[case] cartesian ranges nd={nd}", flush=True)
    call(kat, "kat_cartesian_ranges", nd, nn.ctypes.data_as(_dint), P(lo), P(hi), P(lo2), P(hi2), D(k), P(c),
         P(out))
    first = _cart(kat, 1, 0, n, lo, hi, k, 0.0, c)
    assert out[0] == first and out[4] == first
    assert out[1] == out[2] == out[3]
    assert out[5] == pytest.approx(out[3], rel=1e-12)
    rules = [(qp.get_points(m), qp.get_weights(m)) for m in n]
    assert out[3] == pytest.approx(O.kat_cartesian(rules, lo2, hi2, c, 0.0), rel=TOL, abs=1e-14)
    assert out[3] != pytest.approx(first, rel=1e-3)


@pytest.mark.parametrize("nt", [0, 1, 2], ids=["double", "complex", "dual"])
def test_cartesian_number_types(kat, qp, nt):
    # integrator_1D_cartesian_gpu.cu: double, complex<double> and a dual number, GPU and host
    rng = np.random.default_rng(130 + nt)
    x, w = qp.get_points(32), qp.get_weights(32)
    lo, hi = rng.uniform(-2, -1), rng.uniform(1, 2)
    c = arr(rng.uniform(-1, 1, 4))
    const = rng.uniform(-1, 1)
    integral = O.kat_cartesian([(x, w)], [lo], [hi], c, 0.0)
    second = [0.0, 0.5, 0.25][nt] * integral
    for device in (1, 0):
        out = np.zeros(2)
        print(f"[case] cartesian nt={nt} device={device}", flush=True)
        call(kat, "kat_cartesian1_nt", device, nt, 32, D(lo), D(hi), D(0.5), D(const), P(c), P(out))
        assert out[0] == pytest.approx(const + integral, rel=TOL, abs=TOL)
        assert out[1] == pytest.approx(second, rel=TOL, abs=TOL)
        assert is_close(out[0], _cart_closed(c, [lo], [hi], const), 1e-6)


FREQ = {"3D": (0, 3, 0, 0), "2Dpq0": (1, 2, 1, 0), "2Dpx0": (2, 2, 1, 1), "3Dpq0": (3, 3, 1, 0), "3Dpx0": (4, 3, 1, 1)}


@pytest.mark.parametrize("name", list(FREQ))
@pytest.mark.parametrize("sizes", [(64, 12, 10), (20, 8, 40)], ids=["64x12x10", "20x8x40"])
def test_freq_template(kat, qp, name, sizes):
    # Integrator3DGPU/TBB and Integrator{2,3}D{pq0,px0}GPU/TBB (integrator_3D_gpu.cu,
    # integrator_{2,3}D{pq0,px0}_gpu.cu): oracle parity, the reference tests' closed forms, and
    # request() bitwise get(); the frequency axis has its own order (10 or 40 nodes)
    form, dim, q0_axis, px0 = FREQ[name]
    nx, na, nb = sizes
    rng = np.random.default_rng(140 + form + nx)
    xq, wq = qp.get_points(nx), qp.get_weights(nx)
    xa, wa = qp.get_points(na), qp.get_weights(na)
    xb, wb = qp.get_points(nb), qp.get_weights(nb)
    for device in (1, 0):
        x_ext, ext, k = rng.uniform(1, 2), rng.uniform(1, 2), rng.uniform(0.1, 1)
        c, a, p, z = arr(rng.uniform(-1, 1, 6)), arr(rng.uniform(-1, 1, 4)), arr(rng.uniform(-1, 1, 4)), arr(
            rng.uniform(-1, 1, 4))
        const = rng.uniform(-1, 1)
        print(f"[case] {name} sizes={sizes} device={device}", flush=True)
        out = D()
        call(kat, "kat_freq", device, form, 0, nx, na, na, nb, D(x_ext), D(ext), D(k), P(c), P(a), P(p), P(z),
             D(const), C.byref(out))
        ref = O.kat_freq(dim, q0_axis, px0, xq, wq, xa, wa, xb, wb, x_ext, ext, k, c, a, p, z, const)
        assert out.value == pytest.approx(ref, rel=TOL, abs=TOL * abs(const))
        q_ext = math.sqrt(x_ext) * k
        Q = k * ext if px0 else ext
        if dim == 2:
            closed = S_d(2) / 2 * poly_int(c, 0, q_ext, 1) * poly_int(a, -1, 1) * poly_int(z, -Q, Q) / (2 * math.pi) ** 3
        else:
            closed = poly_int(c, 0, q_ext, 2) * poly_int(a, -1, 1) * poly_int(p, 0, 2 * math.pi)
            closed /= (2 * math.pi) ** (4 if q0_axis else 3)
            if q0_axis:
                closed *= poly_int(z, -Q, Q)
        assert is_close(out.value, const + closed, 1e-6 if nx >= 64 else 1e-4)
        if device:
            req = D()
            call(kat, "kat_freq", device, form, 1, nx, na, na, nb, D(x_ext), D(ext), D(k), P(c), P(a), P(p), P(z),
                 D(const), C.byref(req))
            assert req.value == out.value


@pytest.mark.parametrize("name", ["3D", "3Dpq0", "3Dpx0"])
def test_freq_template_unequal_angles(kat, name):
    # integrator_3D_gpu.hh:64-65, integrator_3Dpq0_gpu.hh:69-70: cos and phi orders must be equal
    form = FREQ[name][0]
    z4 = arr(np.zeros(4))
    for device in (1, 0):
        out = D()
        rc = kat.kat_freq(device, form, 0, 16, 8, 6, 8, D(1.0), D(1.0), D(0.5), P(arr(np.zeros(6))), P(z4), P(z4),
                          P(z4), D(0.0), C.byref(out))
        assert rc == -2
        assert kat.kat_last_error() == b"Grid sizes must be currently equal for all angular dimensions!"


TX0 = {"1D-d2": (0, 2), "1D-d3": (0, 3), "1D-d4": (0, 4), "angle-d3": (1, 3), "angle-d4": (1, 4), "4D": (2, 4)}


@pytest.mark.parametrize("name", list(TX0))
def test_finiteTx0_template(kat, qp, name):
    # Integrator{,Angle,4D}FiniteTx0GPU/TBB (integrator_{,angle_,4D_}finiteTx0_gpu.cu): oracle parity at
    # T > 0 and T = 0 (where the sum is off and the integral takes the summands' nodes too), the
    # T > 0 closed form, request() bitwise get(), and set_T(T) bitwise a new integrator at T
    kind, d = TX0[name]
    rng = np.random.default_rng(300 + 10 * kind + d)
    nx, na, nint, ns = 48, 8, 40, 12
    xq, wq = qp.get_points(nx), qp.get_weights(nx)
    xa, wa = qp.get_points(na), qp.get_weights(na)
    for T in (0.25, 0.0):
        x0, w0 = (qp.get_points(nint), qp.get_weights(nint)) if T > 0 else (qp.get_points(nint + ns),
                                                                             qp.get_weights(nint + ns))
        for device in (1, 0):
            k, x_ext = rng.uniform(0.3, 1), rng.uniform(1, 2)
            x0_ext = ns * 10 * 2 * math.pi * T / k * rng.uniform(1, 2) + 20 / k
            c, a, p, z = (arr(rng.uniform(-1, 1, m)) for m in (6, 4, 4, 4))
            c[(0 if (d - 3) % 2 else 1)::2] = 0.0
            const = rng.uniform(-1, 1)
            print(f"[case] finiteTx0 {name} T={T} device={device}", flush=True)

            def run(req, T_first):
                out = D()
                call(kat, "kat_finiteTx0", device, kind, d, req, nx, na, nint, ns, D(x_ext), D(x0_ext), D(T_first),
                     D(T), D(k), P(c), P(a), P(p), P(z), D(const), C.byref(out))
                return out.value

            got = run(0, T)
            ref = O.kat_finiteTx0(kind, d, xq, wq, xa, wa, x0, w0, x_ext, x0_ext, ns, T, k, c, a, p, z, const)
            assert got == pytest.approx(ref, rel=TOL, abs=TOL * abs(const))
            if T > 0:
                closed = const + tx0_closed(kind, d, math.sqrt(x_ext) * k, k, x0_ext, ns, T, c, a, p, z)
                assert got == pytest.approx(closed, rel=1e-10)
            if device:
                assert run(1, T) == got
                assert run(0, 0.7) == got


@pytest.mark.parametrize("d", [1, 2, 3, 4])
def test_constant_integrator(kat, d):
    # IntegratorConstant<d> (integrator_constant.cc): constant + x0 S_d / (2pi)^d, get and request
    rng = np.random.default_rng(400 + d)
    for _ in range(3):
        k, const, x0 = rng.uniform(0, 1), rng.uniform(-1, 1), rng.uniform(-1, 1)
        out = np.zeros(2)
        call(kat, "kat_constant", d, D(k), D(const), D(x0), P(out))
        assert out[0] == pytest.approx(const + S_d(d) * x0 / (2 * math.pi) ** d, rel=1e-15, abs=1e-15)
        assert out[1] == out[0]


def _sum_closed(T, val):
    return 1 / T / math.tanh(0.5 / val) / 2 / val


def _tracked_E(T, k):
    return k if abs(k - 10 * T) / max(k, 10 * T) > 2.5e-2 else 10 * T


def test_4d_finiteT_template(kat, qp):
    # Integrator4DFiniteTq0GPU/TBB (integrator_4D_finiteTq0_gpu.cu:41-72): oracle parity, and the
    # volume x Matsubara-sum closed form extended to polynomials in q, cos, phi (tolerance 5e-5)
    rng = np.random.default_rng(81)
    xq, wq = qp.get_points(64), qp.get_weights(64)
    xa, wa = qp.get_points(16), qp.get_weights(16)
    for device in (1, 0):
        x_ext, T, k, val = rng.uniform(1, 2), rng.uniform(0.01, 1), rng.uniform(0.05, 1), rng.uniform(0.8, 1.2)
        c, a, p = arr(rng.uniform(-1, 1, 6)), arr(rng.uniform(-1, 1, 4)), arr(rng.uniform(-1, 1, 4))
        den = arr([T * T, 0, val * val, 0, 0])
        const = rng.uniform(-1, 1)
        print(f"[case] 4D finite-T device={device} T={T:.3f} k={k:.3f}", flush=True)
        out, nm = D(), C.c_int()
        call(kat, "kat_4d_finiteT", device, 64, 16, D(x_ext), D(T), D(k), P(c), P(a), P(p), P(den), D(const),
             C.byref(out), C.byref(nm))
        mx, mw, _ = O.matsubara_storage_rule(T, _tracked_E(T, k))
        if device:
            assert nm.value == len(mx)
        ref = O.kat_4d_finiteT(xq, wq, xa, wa, x_ext, k, mx, mw, T, c, a, p, den, const)
        assert out.value == pytest.approx(ref, rel=1e-11, abs=1e-11 * abs(const))
        q_ext = math.sqrt(x_ext) * k
        closed = const + poly_int(c, 0, q_ext, 2) * poly_int(a, -1, 1) * poly_int(p, 0, 2 * math.pi) \
            / (2 * math.pi) ** 3 * _sum_closed(T, val)
        assert is_close(out.value, closed, 5e-5)


@pytest.mark.parametrize("d", [3, 4])
def test_angle_finiteT_template(kat, qp, d):
    # IntegratorAngleFiniteTq0GPU/TBB<d> (integrator_angle_finiteTq0_gpu.cu:36-56)
    rng = np.random.default_rng(90 + d)
    xq, wq = qp.get_points(64), qp.get_weights(64)
    xa, wa = qp.get_points(16), qp.get_weights(16)
    for device in (1, 0):
        x_ext, T, k, val = rng.uniform(1, 2), rng.uniform(0.01, 1), rng.uniform(0.05, 1), rng.uniform(0.8, 1.2)
        c, a = arr(rng.uniform(-1, 1, 6)), arr(rng.uniform(-1, 1, 4))
        den = arr([T * T, 0, val * val, 0])
        const = rng.uniform(-1, 1)
        print(f"[case] angle finite-T d={d} device={device} T={T:.3f} k={k:.3f}", flush=True)
        out, nm = D(), C.c_int()
        call(kat, "kat_angle_finiteT", device, d, 64, 16, D(x_ext), D(T), D(k), P(c), P(a), P(den), D(const),
             C.byref(out), C.byref(nm))
        mx, mw, _ = O.matsubara_storage_rule(T, _tracked_E(T, k))
        if device:
            assert nm.value == len(mx)
        ref = O.kat_angle_finiteT(d, xq, wq, xa, wa, x_ext, k, mx, mw, T, c, a, den, const)
        assert out.value == pytest.approx(ref, rel=1e-11, abs=1e-11 * abs(const))
        q_ext = math.sqrt(x_ext) * k
        closed = const + S_d(d - 1) / 2 * poly_int(c, 0, q_ext, d - 2) * poly_int(a, -1, 1) \
            / (2 * math.pi) ** (d - 1) * _sum_closed(T, val)
        assert is_close(out.value, closed, 5e-5)


def test_errors(kat):
    print("[case] unequal angular orders", flush=True)
    assert kat.kat_unequal_angles() == -2
    assert b"Grid sizes must be currently equal" in kat.kat_last_error()


# ------------------------------------------------------------------------------ batched requests

def _se_case(n_p=16, nq=32, na=16):
    c = W.selfenergy(n_p=n_p, nq=nq, na=na)
    return c, arr(c["p"]), arr(c["Z"])


def test_batched_requests_match_get_bitwise(kat, qp):
    c, p, Z = _se_case()
    req, get = np.empty(len(p)), np.empty(len(p))
    args = (32, 16, D(20.0), D(1.0), P(p), len(p), P(Z), len(Z), D(1e-3), D(40.0), D(6.0))
    print("[case] self-energy: 16 batched requests vs 16 get()", flush=True)
    call(kat, "kat_selfenergy", *args, P(req))
    call(kat, "kat_selfenergy_get", *args, P(get))
    assert np.array_equal(req, get)  # same block partition and reduction order
    coord = O.logcoord(len(Z), 1e-3, 40.0, 6.0)
    ref = O.c1_selfenergy(p, [Z] * len(p), [coord] * len(p), [(qp.get_points(32), qp.get_weights(32))] * len(p),
                          [(qp.get_points(16), qp.get_weights(16))] * len(p))
    assert np.max(np.abs(req - ref) / ref) <= TOL


def test_request_data_update_data_get_data(kat, qp):
    c, _, Z = _se_case()
    n_p = 24
    req, get, pg = np.empty(n_p), np.empty(n_p), np.empty(n_p)
    print("[case] request_data / update_data / get_data over a 24-point log grid", flush=True)
    call(kat, "kat_request_data", 32, 16, D(20.0), D(1.0), n_p, D(1e-2), D(10.0), D(3.0), P(Z), len(Z), D(1e-3),
         D(40.0), D(6.0), P(req), P(get), P(pg))
    assert np.allclose(pg, W.log_coord_forward(n_p, 1e-2, 10.0, 3.0), rtol=1e-15, atol=0)
    assert np.array_equal(req, get)
    coord = O.logcoord(len(Z), 1e-3, 40.0, 6.0)
    ref = O.c1_selfenergy(pg, [Z] * n_p, [coord] * n_p, [(qp.get_points(32), qp.get_weights(32))] * n_p,
                          [(qp.get_points(16), qp.get_weights(16))] * n_p)
    assert np.max(np.abs(req - ref) / ref) <= TOL


def test_finiteT_requests_follow_E(kat):
    # k moves by more than the 2.5% hysteresis several times: the rule (and the batch) changes
    ks = arr([1.0, 0.99, 0.9, 0.89, 0.7, 0.5, 0.49, 0.3, 0.1, 0.05])
    c = arr([1.0, 0.3, -0.2, 0.1, 0.0, 0.05])
    den = arr([0.04, 0.0, 1.1, 0.0])
    req, get = np.empty(len(ks)), np.empty(len(ks))
    print("[case] finite-T requests with E tracking", flush=True)
    call(kat, "kat_finiteT_requests", 32, D(2.0), D(0.2), P(ks), len(ks), P(c), P(den), P(req), P(get))
    print(f"  max rel diff {np.max(np.abs(req - get) / np.abs(get)):.3e}", flush=True)
    assert np.all(np.isfinite(req))
    assert np.array_equal(req, get)


@pytest.mark.parametrize("which", [0, 1, 2], ids=["angle", "4D_finiteT", "angle_finiteT"])
def test_variant_requests_match_get_bitwise(kat, which):
    # k crosses the 2.5% E hysteresis several times: the finite-T forms change rule and batch
    ks = arr([1.0, 0.99, 0.9, 0.89, 0.7, 0.5, 0.49, 0.3, 0.1, 0.05])
    req, get = np.empty(len(ks)), np.empty(len(ks))
    print(f"[case] requests vs get, variant {which}", flush=True)
    call(kat, "kat_variant_requests", which, 32, 8, D(1.5), D(0.2), P(ks), len(ks), P(req), P(get))
    assert np.all(np.isfinite(req)) and np.all(req != 0)
    assert np.array_equal(req, get)


@pytest.mark.parametrize("which", [0, 1, 2])
def test_launch_at_request_matches_get_bitwise(kat, which):
    # set_launch_at_request(true): every request is launched at the call as a batch of one, as the
    # reference's request() does; the values stay bitwise those of get() and of batched requests
    ks = arr([1.0, 0.9, 0.7, 0.5, 0.3])
    batched, get = np.empty(len(ks)), np.empty(len(ks))
    call(kat, "kat_variant_requests", which, 32, 8, D(1.5), D(0.2), P(ks), len(ks), P(batched), P(get))
    assert kat.kat_set_launch_at_request(1) == 1
    try:
        eager, get2 = np.empty(len(ks)), np.empty(len(ks))
        print(f"[case] launch at request, variant {which}", flush=True)
        call(kat, "kat_variant_requests", which, 32, 8, D(1.5), D(0.2), P(ks), len(ks), P(eager), P(get2))
    finally:
        assert kat.kat_set_launch_at_request(0) == 0
    assert np.array_equal(eager, get) and np.array_equal(get2, get) and np.array_equal(batched, get)


def test_request_batching_is_one_launch_per_batch(kat):
    c, p, Z = _se_case(n_p=64, nq=64, na=32)
    tg, tr = D(), D()
    print("[case] 64 self-energy requests: get() loop vs one batch", flush=True)
    call(kat, "kat_request_timing", 64, 32, D(20.0), D(1.0), P(p), len(p), P(Z), len(Z), D(1e-3), D(40.0), D(6.0),
         5, C.byref(tg), C.byref(tr))
    print(f"  64 x get(): {tg.value * 1e3:.3f} ms, 64 requests in one batch: {tr.value * 1e3:.3f} ms "
          f"({tg.value / tr.value:.1f}x)", flush=True)
    assert tr.value < tg.value


def test_get_after_large_batch_matches_oracle(kat, qp):
    """get() after a batch of 256 two-angle requests at {64, 32, 32} with the same integrator: the
    batch's partial slots and tickets must not leak into the scratch get() uses."""
    c = W.selfenergy(n_p=4)
    p, Z = arr(c["p"]), arr(c["Z"])
    n_batch = 256
    out_get, out_batch = np.empty(4), np.empty(n_batch)
    print("[case] 256 requests {64,32,32}, then 4 get()", flush=True)
    call(kat, "kat_selfenergy_after_batch", 64, 32, D(20.0), D(1.0), P(p), 4, P(Z), len(Z), D(1e-3), D(40.0),
         D(6.0), n_batch, P(out_get), P(out_batch))
    coord = O.logcoord(len(Z), 1e-3, 40.0, 6.0)
    xq, xa = (qp.get_points(64), qp.get_weights(64)), (qp.get_points(32), qp.get_weights(32))
    ref = O.c1_selfenergy(p, [Z] * 4, [coord] * 4, [xq] * 4, [xa] * 4)
    assert np.max(np.abs(out_get - ref) / ref) <= TOL
    pb = arr([p[i % 4] * (1.0 + 1e-3 * i) for i in range(n_batch)])
    refb = O.c1_selfenergy(pb[:8], [Z] * 8, [coord] * 8, [xq] * 8, [xa] * 8)
    assert np.max(np.abs(out_batch[:8] - refb) / refb) <= TOL


def test_request_data_launches_before_the_reads(kat):
    """request_data launches its batch when its loop ends: with host work between request_data and
    the reads longer than the batch, the reads find the results ready."""
    c = W.selfenergy()
    Z = arr(c["Z"])
    tf, tr = D(), D()
    print("[case] request_data, 5 ms of host work, then the reads", flush=True)
    call(kat, "kat_request_overlap", 64, 32, D(20.0), D(1.0), 64, P(Z), len(Z), D(1e-3), D(40.0), D(6.0), 5000, 5,
         C.byref(tf), C.byref(tr))
    print(f"  request_data + reads: {tf.value * 1e6:.1f} us; reads after host work: {tr.value * 1e6:.1f} us",
          flush=True)
    assert tf.value < 5e-3  # the batch is shorter than the host work
    assert tr.value < 0.5 * tf.value


@pytest.mark.parametrize("fp32", [0, 1])
def test_selfenergy_request_data_full_size(kat, qp, fp32):
    """The 64-momentum self-energy call set at {64, 32, 32} (4,194,304 evaluations) through
    Integrator4D2AngGPU and request_data, against the oracle: FP64 at 1e-12; the FP32 form
    (NT = float, LinearInterpolator1D<float>, FP64 accumulation) at 1e-5, its measured error is
    printed."""
    c = W.selfenergy()
    Z = arr(c["Z"])
    out, pg, td, th = np.empty(64), np.empty(64), D(), D()
    print(f"[case] self-energy call set through request_data, fp32={fp32}", flush=True)
    call(kat, "kat_selfenergy_request_bench", 64, 32, D(20.0), D(1.0), 64, P(Z), len(Z), D(1e-3), D(40.0), D(6.0),
         fp32, 3, P(out), P(pg), C.byref(td), C.byref(th), None)
    assert np.allclose(pg, c["p"], rtol=1e-15, atol=0)
    coord = O.logcoord(len(Z), 1e-3, 40.0, 6.0)
    xq, xa = (qp.get_points(64), qp.get_weights(64)), (qp.get_points(32), qp.get_weights(32))
    ref = O.c1_selfenergy(pg, [Z] * 64, [coord] * 64, [xq] * 64, [xa] * 64)
    err = np.max(np.abs(out - ref) / np.abs(ref))
    print(f"  device {td.value * 1e6:.1f} us, host {th.value * 1e6:.1f} us per call set; max rel err {err:.2e}",
          flush=True)
    assert err <= (1e-5 if fp32 else TOL)
