"""Explicit Runge-Kutta stepping of variables systems (include/dfrg/explicit_rk.hh, the reference's
TimeStepperRK::run_vars, timestepping/rk.cc:57-134) and the x_extent choice (include/dfrg/x_extent.hh,
FlowEquations::update_x_extent, flow_equations.cc:54-105), through tests/native/libdfrg_kat.so.

Pinned by the mathematics: each method's convergence order on a linear system with the matrix
exponential as exact solution (a wrong tableau entry breaks the order), the step sequence of the
end-time rule, the device ensemble against the host stepper, and the x_extent stopping criterion
checked with the closed-form radial integral of a Gaussian.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest
from scipy.linalg import expm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KAT = os.path.join(ROOT, "tests", "native", "libdfrg_kat.so")
_dp = C.POINTER(C.c_double)
D = C.c_double
ORDER = {1: 1, 2: 3, 3: 3, 4: 4}


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(KAT):
        pytest.skip("tests/native/libdfrg_kat.so not built")
    L = C.CDLL(KAT)
    L.kat_last_error.restype = C.c_char_p
    return L


def P(a):
    return a.ctypes.data_as(_dp)


def host_run(lib, method, A, y0, t0, t1, dt, output_dt, n_snap=0):
    A = np.ascontiguousarray(A, dtype=np.float64).reshape(-1, 4)
    y = np.ascontiguousarray(y0, dtype=np.float64).reshape(-1, 2).copy()
    snap = np.zeros((max(n_snap, 1), 2))
    nw = C.c_int()
    rc = lib.kat_rk_host(method, P(A), P(y), len(y), D(t0), D(t1), D(dt), D(output_dt), P(snap), n_snap,
                         C.byref(nw))
    return rc, y, snap[:nw.value]


A0 = np.array([[-1.0, 2.0], [-2.0, -1.0]])
Y0 = np.array([1.0, 0.5])


@pytest.mark.parametrize("method", [1, 2, 3, 4])
def test_host_convergence_order(lib, method):
    exact = expm(A0 * 1.0) @ Y0
    errs = []
    for dt in (0.1, 0.05, 0.025):
        rc, y, _ = host_run(lib, method, A0.ravel(), Y0, 0.0, 1.0, dt, dt)
        assert rc == 0, lib.kat_last_error()
        errs.append(np.linalg.norm(y[0] - exact))
    p = ORDER[method]
    for e0, e1 in zip(errs, errs[1:]):
        assert 0.75 * 2 ** p <= e0 / e1 <= 1.3 * 2 ** p, (method, errs)


def test_end_time_rule_and_outputs(lib):
    # deal.II DiscreteTime: the last step lands on t_stop, stretched when the remainder is < 5% of dt
    times = []
    for dt, want in ((0.33, [0.0, 0.33, 0.66, 1.0]), (0.3, [0.0, 0.3, 0.6, 0.9, 1.0]), (0.25, [0, .25, .5, .75, 1])):
        rc, _, snap = host_run(lib, 1, np.zeros(4), [1.0, 0.0], 0.0, 1.0, dt, dt, n_snap=16)
        assert rc == 0
        # y' = 0 for A = 0; snapshots are written at t_start and after every step
        assert len(snap) == len(want)
        n = C.c_int()
        assert lib.kat_rk_snapshots(D(0.0), D(1.0), D(dt), D(dt), C.byref(n)) == 0 and n.value == len(want)
        times.append(len(snap))
    # y' = 1 (as y1' = 1 via A acting on y0 = (1, 0)): y1(t) = t exactly, so snapshots give the times
    A = np.array([0.0, 0.0, 1.0, 0.0])
    rc, _, snap = host_run(lib, 4, A, [1.0, 0.0], 0.0, 1.0, 0.33, 0.33, n_snap=16)
    assert np.allclose(snap[:, 1], [0.0, 0.33, 0.66, 1.0], rtol=0, atol=1e-15)
    # output every second step
    rc, _, snap = host_run(lib, 4, A, [1.0, 0.0], 0.0, 1.0, 0.1, 0.2, n_snap=16)
    assert np.allclose(snap[:, 1], [0.0, 0.2, 0.4, 0.6, 0.8, 1.0], rtol=0, atol=1e-14)


def test_host_errors(lib):
    rc, _, _ = host_run(lib, 5, A0.ravel(), Y0, 0.0, 1.0, 0.1, 0.1)
    assert rc == -2 and b"method not implemented" in lib.kat_last_error()
    rc, _, _ = host_run(lib, 4, A0.ravel(), Y0, 0.0, 1.0, 0.0, 0.1)
    assert rc == -1
    rc, _, _ = host_run(lib, 4, A0.ravel(), Y0, 0.0, 1.0, 1e-9, 0.1)
    assert rc == -1 and b"dt too small" in lib.kat_last_error()
    rc, _, _ = host_run(lib, 4, np.array([60.0, 0, 0, 60.0]), Y0, 0.0, 20.0, 0.01, 0.1)
    assert rc == -2 and b"NaNs in solution" in lib.kat_last_error()


# ------------------------------------------------------------------------------------ device

def device_run(lib, method, A, y0, t0, t1, dt, output_dt, nx=0, c=0.0, n_snap=0):
    A = np.ascontiguousarray(A, dtype=np.float64)
    y = np.ascontiguousarray(y0, dtype=np.float64).copy()
    n = len(y) // 2
    snap = np.zeros(max(n_snap, 1) * n * 2)
    status = np.zeros(n, dtype=np.int32)
    rc = lib.kat_rk_device(method, P(A), P(y), n, D(t0), D(t1), D(dt), D(output_dt), nx, D(c), P(snap), n_snap,
                           status.ctypes.data_as(C.POINTER(C.c_int)))
    assert rc == 0, lib.kat_last_error()
    return y.reshape(n, 2), snap[:n_snap * n * 2].reshape(n_snap, n, 2), status


@pytest.mark.gpu
@pytest.mark.parametrize("method", [1, 2, 3, 4])
def test_device_ensemble_matches_host(lib, gpu, method):
    rng = np.random.default_rng(300 + method)
    n = 1000
    A = rng.uniform(-2, 1, (n, 4))
    y0 = rng.uniform(-1, 1, (n, 2))
    print(f"[case] RK ensemble method {method}, {n} systems", flush=True)
    t0, t1, dt, odt = 0.0, 1.3, 0.02, 0.1
    n_snap = C.c_int()
    assert lib.kat_rk_snapshots(D(t0), D(t1), D(dt), D(odt), C.byref(n_snap)) == 0
    yd, snap, status = device_run(lib, method, A.ravel(), y0.ravel(), t0, t1, dt, odt, n_snap=n_snap.value)
    rc, yh, snap_h = host_run(lib, method, A.ravel(), y0.ravel(), t0, t1, dt, odt, n_snap=64)
    assert rc == 0
    assert np.all(status == 0)
    scale = np.maximum(1.0, np.abs(yh))
    assert np.max(np.abs(yd - yh) / scale) <= 1e-13
    assert len(snap_h) == n_snap.value
    assert np.max(np.abs(snap[:, 0, :] - snap_h) / np.maximum(1.0, np.abs(snap_h))) <= 1e-13
    exact = np.array([expm(A[i].reshape(2, 2) * t1) @ y0[i] for i in range(0, n, 97)])
    tol = {1: 0.2, 2: 1e-4, 3: 1e-4, 4: 1e-6}[method]
    assert np.max(np.abs(yd[::97] - exact) / np.maximum(1, np.abs(exact))) <= tol


@pytest.mark.gpu
def test_device_ensemble_with_loop_integral_is_fourth_order(lib, gpu):
    # y0' gets a GL threshold integral of y0 (nonlinear RHS evaluated in-thread)
    rng = np.random.default_rng(9)
    n = 256
    A = rng.uniform(-1, 0.5, (n, 4)).ravel()
    y0 = rng.uniform(0.2, 1, (n, 2)).ravel()
    print("[case] RK4 ensemble with a loop integral", flush=True)
    ys = [device_run(lib, 4, A, y0, 0.0, 1.0, dt, 1.0, nx=64, c=-3.0)[0] for dt in (0.1, 0.05, 0.025, 0.0125)]
    d = [np.max(np.abs(a - b)) for a, b in zip(ys, ys[1:])]
    for e0, e1 in zip(d, d[1:]):
        assert 12 <= e0 / e1 <= 20, d


@pytest.mark.gpu
def test_device_ensemble_flags_blow_up(lib, gpu):
    A = np.array([60.0, 0, 0, 60.0, -1.0, 0, 0, -1.0])
    y, _, status = device_run(lib, 4, A, np.array([1.0, 1.0, 1.0, 1.0]), 0.0, 20.0, 0.01, 20.0)
    assert list(status) == [1, 0]
    assert np.allclose(y[1], np.exp(-20.0), rtol=1e-8)


# ------------------------------------------------------------------------------------ x_extent

def gauss_I(a, X):
    # int d^4q/(2pi)^4 exp(-a q^2) over q^2 <= X = S_4/(2pi)^4 * 0.5 * (1 - (1 + aX) e^{-aX}) / a^2
    return 2 * math.pi ** 2 / (2 * math.pi) ** 4 * 0.5 * (1 - (1 + a * X) * math.exp(-a * X)) / a ** 2


@pytest.mark.gpu
@pytest.mark.parametrize("a,tol", [(0.5, 1e-3), (0.1, 1e-4), (2.0, 1e-6)])
def test_x_extent_criterion(lib, gpu, a, tol):
    x = D()
    print(f"[case] x_extent exp(-{a} q^2) tol {tol}", flush=True)
    assert lib.kat_x_extent(0, D(a), 16, D(tol), C.byref(x)) == 0, lib.kat_last_error()
    X = x.value
    m = round(math.log(X) / math.log(1.15))
    assert X == pytest.approx(1.15 ** m, rel=1e-12)

    def ok(X):
        I1, I2, I3 = gauss_I(a, X), gauss_I(a, 2 * X), gauss_I(a, 10 * X)
        return abs((I2 - I1) / I1) <= tol and abs((I2 - I3) / I2) <= tol

    assert ok(X)
    if m > 0:
        assert not ok(X / 1.15)


@pytest.mark.gpu
def test_x_extent_special_cases(lib, gpu):
    x = D()
    assert lib.kat_x_extent(1, D(0.0), 16, D(1e-3), C.byref(x)) == 0 and x.value == 1.0
    # 1/(1+q^2)^3: I(X) is proportional to X^2/(1+X)^2, so |I(2X) - I(X)|/I(X) ~ 1/X and the
    # criterion at 1e-3 needs X ~ 1000 (GL on [0, X] resolves the peak near 0 only roughly)
    assert lib.kat_x_extent(2, D(0.0), 16, D(1e-3), C.byref(x)) == 0
    X = x.value
    assert 700 < X < 1700
