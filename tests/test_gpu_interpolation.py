"""The 2D/3D, FP32 ("Tex") and stacked interpolators (include/DiFfRG/physics/interpolation/) inside
kernels and on the host, through tests/native/libdfrg_kat.so.

Pinned by what the mathematics fixes: multilinear interpolation reproduces any table that is
multilinear in the indices exactly (a wrong corner, weight or clamp fails this), at clamped
indices for points outside the table; plus the reference tests' own data and expected values
(tests/physics/interpolation/*_gpu.cu) and the device == host agreement they check.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KAT = os.path.join(ROOT, "tests", "native", "libdfrg_kat.so")
_dp = C.POINTER(C.c_double)


@pytest.fixture(scope="module")
def kat(gpu):
    lib = C.CDLL(KAT)
    lib.kat_last_error.restype = C.c_char_p
    return lib


def P(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(_dp)


def run(kat, kind, cp, data, pts):
    cp, data, pts = (np.ascontiguousarray(v, dtype=np.float64) for v in (cp, data, pts))
    n = len(pts)
    dev, host = np.empty(n), np.empty(n)
    print(f"[case] interpolator kind {kind}, {len(data)} table values, {n} points", flush=True)
    rc = kat.kat_interp(kind, P(cp), P(data), len(data), P(pts), n, P(dev), P(host))
    assert rc == 0, kat.kat_last_error()
    return dev, host


def lin_axis(rng, dtype):
    n = int(rng.integers(10, 100))
    start = dtype(rng.uniform(1e-6, 1e-1))
    stop = dtype(rng.uniform(1, 100) + start)
    return n, start, stop


def lin_backward(x, n, start, stop, dtype):
    a = dtype((dtype(stop) - dtype(start)) / dtype(n - 1))
    return (dtype(x) - dtype(start)) / a


def points(rng, axes, m, dtype):
    # mostly inside, some beyond either end of every axis (clamping)
    cols = []
    for n, start, stop in axes:
        u = rng.uniform(-0.1, 1.1, m)
        cols.append((start + u * (stop - start)).astype(dtype).astype(np.float64))
    return np.stack(cols, axis=1)


def multilinear(coef, idx):
    # coef over the subsets of axes: sum_S c_S prod_{a in S} idx_a
    out = np.zeros(len(idx))
    d = idx.shape[1]
    for mask in range(1 << d):
        term = np.full(len(idx), coef[mask])
        for a in range(d):
            if mask >> a & 1:
                term = term * idx[:, a]
        out += term
    return out


@pytest.mark.parametrize("kind,dim,dtype,tol", [(0, 2, np.float64, 1e-12), (1, 3, np.float64, 1e-12),
                                                (3, 2, np.float32, 2e-5), (4, 3, np.float32, 2e-5),
                                                (7, 2, np.float32, 2e-5)])
def test_multilinear_tables_are_reproduced(kat, kind, dim, dtype, tol):
    rng = np.random.default_rng(200 + kind)
    for _ in range(3):
        axes = [lin_axis(rng, dtype) for _ in range(dim)]
        shape = [a[0] for a in axes]
        grids = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij")
        grid_idx = np.stack([g.ravel() for g in grids], axis=1)
        coef = rng.uniform(-1, 1, 1 << dim)
        data = multilinear(coef, grid_idx)
        pts = points(rng, axes, 300, dtype)
        dev, host = run(kat, kind, [v for a in axes for v in a], data, pts)
        idx = np.stack([np.clip(lin_backward(pts[:, a].astype(dtype), *axes[a], dtype).astype(np.float64), 0,
                                shape[a] - 1) for a in range(dim)], axis=1)
        want = multilinear(coef, idx)
        scale = np.sum(np.abs(coef)) * np.prod(shape)
        assert np.max(np.abs(dev - want)) <= tol * scale
        assert np.max(np.abs(host - want)) <= tol * scale
        assert np.max(np.abs(dev - host)) <= tol * scale


def test_reference_2D_3D_cases(kat):
    # linear_interpolation_2D_gpu.cu:24-56 (data = j -> clamped j index), 3D :36-73 (data = k)
    rng = np.random.default_rng(7)
    for kind, dim in ((7, 2), (3, 2), (4, 3)):
        axes = [lin_axis(rng, np.float32) for _ in range(dim)]
        shape = [a[0] for a in axes]
        grids = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij")
        data = grids[-1].ravel()
        pts = np.stack([(a[1] + rng.uniform(0, 1, 50) * (a[2] - a[1])).astype(np.float32).astype(np.float64)
                        for a in axes], axis=1)
        dev, host = run(kat, kind, [v for a in axes for v in a], data, pts)
        last = np.clip(lin_backward(pts[:, -1].astype(np.float32), *axes[-1], np.float32), 0, shape[-1])
        for res in (dev, host):
            assert np.all(np.abs(res - last) <= 1e-6 * np.maximum(1, np.abs(last)) + 1e-5)


def log_backward(y, n, start, stop, bias):
    f = np.float32
    a, b, c = f(bias), f((f(stop) - f(start)) / f(math.expm1(f(bias)))), f(start)
    with np.errstate(invalid="ignore"):  # below the table start: NaN, clamped by the caller
        return np.log1p((f(y) - c) / b) * f(n - 1) / a


def test_tex_1D_log_coordinates(kat):
    # tex_linear_interpolation_1D_gpu.cu:19-57: data = j on log coordinates
    rng = np.random.default_rng(11)
    for _ in range(4):
        n = int(rng.integers(10, 100))
        start = rng.uniform(1e-6, 1e-1)
        stop = rng.uniform(1, 100) + start
        bias = rng.uniform(1, 10)
        cp = [n, np.float32(start), np.float32(stop), np.float32(bias)]
        data = np.arange(n, dtype=np.float64) * 0.75 - 3.0
        pts = (start + rng.uniform(-0.05, 1.05, 200) * (stop - start)).astype(np.float32).astype(np.float64)
        dev, host = run(kat, 2, cp, data, pts)
        # far below the table start log1p has no real value: the NaN index clamps to the last point
        # (fmax(0, fmin(NaN, n - 1)), the reference's device clamp)
        idx = np.clip(np.nan_to_num(log_backward(pts, n, *cp[1:]).astype(np.float64), nan=n - 1), 0, n - 1)
        want = idx * 0.75 - 3.0
        assert np.max(np.abs(dev - want)) <= 2e-5 * n
        assert np.max(np.abs(dev - host)) <= 2e-5 * n


@pytest.mark.parametrize("kind,fermionic", [(5, False), (6, True)])
def test_tex_1D_stack(kat, kind, fermionic):
    # tex_linear_interpolation_1D_stack_gpu.cu:16-55: data = i p_size + j; frequencies on the stack
    rng = np.random.default_rng(20 + kind)
    for _ in range(4):
        m_start = int(rng.integers(-50, 5))
        m_size = int(rng.integers(2, 64))
        T = float(np.float32(rng.uniform(1e-3, 1e-1)))
        n = int(rng.integers(10, 100))
        start, stop, bias = (float(np.float32(v)) for v in (rng.uniform(1e-6, 1e-1), rng.uniform(1, 100),
                                                            rng.uniform(1, 10)))
        stop += start
        cp = [m_start, m_start + m_size, T, n, start, stop, bias]
        data = np.arange(m_size * n, dtype=np.float64)
        m_idx = rng.integers(0, m_size, 100)
        shift = 0.5 if fermionic else 0.0
        m = (m_idx + m_start + shift) * 2 * math.pi * T
        p = start + rng.uniform(-0.05, 1.05, 100) * (stop - start)
        pts = np.stack([m, p], axis=1).astype(np.float32).astype(np.float64)
        dev, host = run(kat, kind, cp, data, pts)
        pidx = np.clip(np.nan_to_num(log_backward(pts[:, 1], n, start, stop, bias).astype(np.float64), nan=n - 1),
                       0, n - 1)
        want = m_idx * n + pidx
        assert np.max(np.abs(dev - want) / np.maximum(1, want)) <= 2e-6
        assert np.max(np.abs(dev - host) / np.maximum(1, want)) <= 2e-6
        # beyond the last frequency: last table, |p| continued by the frequency excess
        mo = (m_start + m_size + 2 + shift) * 2 * math.pi * T
        po = 0.5 * (start + stop)
        dev, _ = run(kat, kind, cp, data, np.array([[np.float32(mo), np.float32(po)]], dtype=np.float64))
        dm = np.float32(mo) - np.float32((m_start + m_size - 1 + shift) * 2 * math.pi * T)
        pe = np.sqrt(np.float32(po) ** 2 + dm ** 2)
        want = (m_size - 1) * n + np.clip(log_backward(pe, n, start, stop, bias), 0, n - 1)
        assert abs(dev[0] - want) <= 2e-6 * want


def run_dual(kat, kind, cp, val, der, pts):
    cp, val, der, pts = (np.ascontiguousarray(v, dtype=np.float64) for v in (cp, val, der, pts))
    n = len(pts)
    dev, host = np.empty(2 * n), np.empty(2 * n)
    print(f"[case] dual-number interpolator kind {kind}, {len(val)} table values, {n} points", flush=True)
    rc = kat.kat_interp_dual(kind, P(cp), P(val), P(der), len(val), P(pts), n, P(dev), P(host))
    assert rc == 0, kat.kat_last_error()
    return dev.reshape(n, 2), host.reshape(n, 2)


def test_dual_number_channels(kat):
    # Dual-number data (the reference's autodiff::real, tex_linear_interpolation_1D.hh:58-72,
    # 86-106): each channel is the FP32 table interpolator on that channel's data, so both must be
    # bitwise the plain interpolator's results on the value data and on the derivative data.
    rng = np.random.default_rng(31)
    n1 = 40
    log_cp = [n1, np.float32(1e-3), np.float32(30.0), np.float32(4.0)]
    a2 = [lin_axis(rng, np.float32) for _ in range(2)]
    a3 = [(int(rng.integers(4, 12)),) + lin_axis(rng, np.float32)[1:] for _ in range(3)]
    T, n_s = 0.05, 24
    st_cp = [-3, 9, T, n_s, 1e-3, 20.0, 3.0]
    cases = [  # (dual kind, plain kind, cp, table size, points)
        (0, 2, log_cp, n1, (1e-3 + rng.uniform(-0.05, 1.05, 150) * 30).astype(np.float32).astype(np.float64)),
        (1, 3, [v for ax in a2 for v in ax], a2[0][0] * a2[1][0], points(rng, a2, 150, np.float32)),
        (2, 4, [v for ax in a3 for v in ax], a3[0][0] * a3[1][0] * a3[2][0], points(rng, a3, 150, np.float32)),
        (3, 5, st_cp, 12 * n_s, np.stack([(rng.integers(-5, 11, 150) * 2 * math.pi * T),
                                          1e-3 + rng.uniform(-0.05, 1.05, 150) * 20], axis=1)
         .astype(np.float32).astype(np.float64)),
    ]
    for kd, kp, cp, size, pts in cases:
        val = rng.uniform(-2, 2, size)
        der = rng.uniform(-5, 5, size)
        dev, host = run_dual(kat, kd, cp, val, der, pts)
        v_dev, v_host = run(kat, kp, cp, val, pts)
        d_dev, d_host = run(kat, kp, cp, der, pts)
        assert np.array_equal(dev[:, 0], v_dev) and np.array_equal(dev[:, 1], d_dev), kd
        assert np.array_equal(host[:, 0], v_host) and np.array_equal(host[:, 1], d_host), kd
    # update() with plain values replaces the value table and keeps the derivative table
    kd, kp, cp, size, pts = cases[0]
    val, der = rng.uniform(-2, 2, size), rng.uniform(-5, 5, size)
    dev, host = run_dual(kat, 5, cp, val, der, pts)
    assert np.array_equal(dev[:, 0], run(kat, kp, cp, 2 * val, pts)[0])
    assert np.array_equal(dev[:, 1], run(kat, kp, cp, der, pts)[0])


def log_index64(x, n, start, stop, bias):
    """The reference's FP64 index log1p((x - c)/b) (n-1)/a clamped to [0, n-1], a NaN index -> n-1
    (fmax(0, fmin(NaN, n-1)), the reference's device clamp), in IEEE arithmetic as the C++ does."""
    f = np.float64
    with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
        b = (f(stop) - f(start)) / f(np.expm1(f(bias)))
        idx = np.log1p((x - f(start)) / b) * (f(n) - 1.0) / f(bias)
    return np.clip(np.where(np.isnan(idx), n - 1.0, idx), 0, n - 1)


@pytest.mark.parametrize("n,bias,start,stop", [(128, 6.0, 1e-3, 40.0), (256, 1.0, 1e-3, 40.0), (32, 6.0, 1e-3, 40.0),
                                               (2000, 3.0, 1e-3, 40.0), (10, 6.0, 1e-3, 40.0), (2, 0.2, 1e-3, 40.0),
                                               (64, -3.0, 1e-3, 40.0), (8, 60.0, 1e-3, 40.0), (5, 300.0, 0.0, 1.0),
                                               (3, 5.0, 2.0, 1.0), (20, 4.0, 1.0, 1.0), (1, 5.0, 0.0, 1.0),
                                               (20, 0.0, 0.0, 1.0)])
def test_fp64_log_coordinates_on_the_device(kat, n, bias, start, stop):
    """LinearInterpolator1D<double> on FP64 log coordinates: on the device the index comes from a
    table of sub-cell starts (an FP32 estimate of the sub-cell, an FP64 series in it), with no
    branch on the grid.  A table linear in the index pins the index itself; the grids include
    sub-cells (bias/(n-1) > 1/4), negative bias, biases whose 1 + u leaves the FP32 range, a
    reversed axis, start == stop, one point and bias 0; the points include grid nodes, both sides of
    cell boundaries at 1e-7 and 1e-13, both ends, below the start (where log1p has no real value and
    the NaN index takes the last point), +-1e300, +-inf and NaN."""
    rng = np.random.default_rng(n + int(abs(bias) * 10))
    cp = [n, start, stop, bias]
    f = np.float64
    with np.errstate(all="ignore"):
        b = (f(stop) - f(start)) / f(np.expm1(f(bias)))
        nodes = b * np.expm1(bias * np.arange(n) / max(n - 1.0, 1.0)) + start
    nodes = nodes[np.isfinite(nodes)]
    near = np.concatenate([nodes * (1 + e) for e in (1e-7, -1e-7, 1e-13, -1e-13)])
    lo, hi = min(start, stop), max(start, stop)
    below = [start - s * b for s in (0.3, 0.5, 0.99, 1.0, 2.0)] if np.isfinite(b) else []
    pts = np.concatenate([nodes, near, rng.uniform(lo, hi, 400), np.exp(rng.uniform(np.log(lo + 1e-3), np.log(hi + 1e-3), 400)),
                          below, [0.0, start, stop, 1.5 * stop, 1e300, -1e300, np.inf, -np.inf, np.nan]])
    # the index within the reference's own FP64 rounding: the input (x - c)/b carries an error of
    # a few ulp, which (n-1)/a log1p turns into ~1e-16 (n-1)/|a| cells (times 1/(1 + u) <= e^|a| for
    # a < 0, where the top cells crowd towards u = -1); the log's own rounding ~1e-16 n
    amp = math.exp(-bias) if bias < 0 else 1.0
    tol_idx = 8 * 2.0 ** -52 * ((max(n, 2) - 1.0) / max(abs(bias), 1e-300) * amp + n) if bias != 0 else 0.0
    lin = np.arange(n, dtype=np.float64) * 0.75 - 3.0
    dev, host = run(kat, 8, cp, lin, pts)
    want = log_index64(pts, n, start, stop, bias) * 0.75 - 3.0
    assert np.all(np.isfinite(dev)) and np.all(np.isfinite(host))
    assert np.max(np.abs(dev - want)) <= 0.75 * tol_idx + 4 * 2.0 ** -52 * 3.0 * n, np.max(np.abs(dev - want))
    assert np.max(np.abs(host - want)) <= 0.75 * tol_idx + 4 * 2.0 ** -52 * 3.0 * n
    # a smooth table: device against the host's reference formula, at the table's cell slopes
    smooth = 1.0 + 0.5 * np.tanh(np.sin(np.arange(n) * 0.37) * 2.0)
    dev, host = run(kat, 8, cp, smooth, pts)
    slope = np.max(np.abs(np.diff(smooth))) if n > 1 else 0.0
    assert np.max(np.abs(dev - host)) <= slope * tol_idx + 4 * 2.0 ** -52


def test_fp64_log_coordinates_refused_bias(kat):
    # beyond |bias| = 700 exp(bias) overflows; the device index refuses such grids at construction
    cp = np.array([16, 0.0, 1.0, 800.0])
    pts = np.array([0.5])
    dev, host = np.empty(1), np.empty(1)
    rc = kat.kat_interp(8, P(cp), P(np.zeros(16)), 16, P(pts), 1, P(dev), P(host))
    assert rc != 0 and b"bias" in kat.kat_last_error()


def test_dual_number_fp64_interpolator(kat):
    # LinearInterpolator1D over dual data (reference linear_interpolation_1D.cu:76-91 instantiates
    # it for autodiff::real): FP64 lerp of both channels.  Value data linear in the index is
    # reproduced exactly; the derivative channel is checked against the lerp formula.
    n, start, stop, bias = 50, 1e-3, 40.0, 5.0
    rng = np.random.default_rng(32)
    j = np.arange(n, dtype=np.float64)
    val = 0.75 * j - 3.0
    der = np.sin(j) + j ** 2 / n
    pts = start + rng.uniform(-0.02, 1.05, 300) * (stop - start)
    dev, host = run_dual(kat, 4, [n, start, stop, bias], val, der, pts)
    b = (stop - start) / math.expm1(bias)
    with np.errstate(invalid="ignore"):
        idx = np.log1p((pts - start) / b) * (n - 1) / bias
    idx = np.clip(np.nan_to_num(idx, nan=n - 1), 0, n - 1)
    fl = np.floor(idx)
    want_d = der[fl.astype(int)] * (1 - idx + fl) + der[np.ceil(idx).astype(int)] * (idx - fl)
    assert np.max(np.abs(dev[:, 0] - (0.75 * idx - 3.0))) <= 1e-12 * n
    assert np.max(np.abs(dev[:, 1] - want_d)) <= 1e-12 * n
    assert np.max(np.abs(dev - host)) <= 1e-12 * n


def test_interpolator_errors(kat):
    print("[case] interpolator errors", flush=True)
    assert kat.kat_interp_errors(0) == -2 and b"non-owner" in kat.kat_last_error()
    assert kat.kat_interp_errors(1) == -1 and b"start must be smaller than stop" in kat.kat_last_error()
    assert kat.kat_interp_errors(2) == -2 and b"at least 2 points" in kat.kat_last_error()
    assert kat.kat_interp_dual_index() == -2 and b"dual-number data" in kat.kat_last_error()


# ------------------------------------------------------------------ C-ABI tables (dfrg_table_*)

def test_c_abi_table_matches_oracle(gpu):
    from oracle import oracle as O
    rng = np.random.default_rng(5)
    n, start, stop, bias = 128, 1e-3, 40.0, 6.0
    vals = np.cumsum(rng.uniform(0.0, 0.1, n)) + 1.0
    t = gpu.Table(vals, start, stop, bias)
    coord = O.logcoord(n, start, stop, bias)
    y = np.concatenate([rng.uniform(0, 45, 500), [start, stop, 0.0, 50.0, -1.0]])
    print("[case] dfrg_table_eval on 505 points", flush=True)
    d_y, d_out = gpu.DeviceArray.from_host(y), gpu.DeviceArray(len(y))
    t.eval(d_y, d_out)
    got = d_out.to_host()
    want = np.array([O.lerp_log(vals, coord, v) for v in y])
    assert np.allclose(got, want, rtol=1e-13, atol=0)
    # update() replaces the values the view points to
    t.update(2.0 * vals)
    t.eval(d_y, d_out)
    assert np.allclose(d_out.to_host(), 2.0 * want, rtol=1e-13, atol=0)
    assert t.view() != 0
    t.close()


def test_c_abi_table_errors(gpu):
    L = gpu.lib()
    h = C.c_void_p()
    v = np.ones(4)
    pv = v.ctypes.data_as(_dp)
    assert L.dfrg_table_create(C.byref(h), 1, 0.0, 1.0, 1.0, pv) == gpu.ERR_INVALID
    assert L.dfrg_table_create(C.byref(h), 4, 1.0, 1.0, 1.0, pv) == gpu.ERR_INVALID
    assert L.dfrg_table_create(C.byref(h), 4, 0.0, 1.0, 0.0, pv) == gpu.ERR_INVALID
    assert b"bias" in L.dfrg_last_error()
