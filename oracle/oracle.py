"""ctypes view of liboracle.so, the CPU restatement of DiFfRG's integrator path.

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` for
its reference values, and by the cpu_baseline leg of ``bench.py``.  The product
(``diffrg_amd``) never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = C.POINTER(C.c_double)


class LogCoord(C.Structure):
    _fields_ = [("start", C.c_double), ("stop", C.c_double), ("bias", C.c_double), ("n", C.c_uint)]


class SEIntegral(C.Structure):
    _fields_ = [("xq", _dp), ("wq", _dp), ("nq", C.c_int), ("xa", _dp), ("wa", _dp), ("na", C.c_int),
                ("tab", _dp), ("coord", LogCoord), ("p", C.c_double)]


class FTIntegral(C.Structure):
    _fields_ = [("T", C.c_double), ("p0", C.c_double), ("pabs", C.c_double), ("m1", C.c_double),
                ("m2", C.c_double), ("mx", _dp), ("mw", _dp), ("nm", C.c_int)]


class FlowParams(C.Structure):
    _fields_ = [("n_grid", C.c_int), ("rho_max", C.c_double), ("T", C.c_double), ("mu", C.c_double),
                ("N", C.c_double), ("m2_0", C.c_double), ("lambda_", C.c_double)]


def build() -> None:
    """Compile liboracle.so with the Makefile beside this file."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        L = _lib
        L.orc_splitmix64.restype = C.c_uint64
        L.orc_splitmix64.argtypes = [C.POINTER(C.c_uint64)]
        L.orc_uniform01.argtypes = [C.c_uint64, _dp, C.c_size_t]
        L.orc_gl_nodes.argtypes = [C.c_int, _dp, _dp]
        L.orc_imtqlx.argtypes = [C.c_int, _dp, _dp, _dp]
        L.orc_make_quadrature.argtypes = [C.c_int, _dp, _dp, C.c_double, _dp, _dp]
        L.orc_matsubara_predict_size.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int, C.c_int]
        L.orc_matsubara.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                    _dp, _dp, C.c_int, _dp]
        for name in ("orc_exp_RB", "orc_exp_RBdot", "orc_exp_dq2RB"):
            getattr(L, name).restype = C.c_double
            getattr(L, name).argtypes = [C.c_double, C.c_double, C.c_double]
        for name in ("orc_litim_RB", "orc_litim_RBdot"):
            getattr(L, name).restype = C.c_double
            getattr(L, name).argtypes = [C.c_double, C.c_double]
        L.orc_logcoord_forward.restype = C.c_double
        L.orc_logcoord_forward.argtypes = [C.POINTER(LogCoord), C.c_uint]
        L.orc_logcoord_backward.restype = C.c_double
        L.orc_logcoord_backward.argtypes = [C.POINTER(LogCoord), C.c_double]
        L.orc_lerp_log.restype = C.c_double
        L.orc_lerp_log.argtypes = [_dp, C.POINTER(LogCoord), C.c_double]
        L.orc_lerp_lin.restype = C.c_double
        L.orc_lerp_lin.argtypes = [_dp, C.c_uint, C.c_double, C.c_double, C.c_double]
        L.orc_c0_threshold.argtypes = [C.c_int, _dp, C.c_int, _dp, _dp, C.c_int, C.c_double, C.c_double,
                                       C.c_double, _dp]
        L.orc_c1_selfenergy.argtypes = [C.POINTER(SEIntegral), C.c_int, C.c_double, C.c_double, C.c_double, _dp]
        L.orc_c2_finiteT.argtypes = [C.c_int, C.c_int, C.POINTER(FTIntegral), C.c_int, _dp, _dp, C.c_int,
                                     C.c_double, C.c_double, C.c_double, _dp]
        L.orc_kat_radial.restype = C.c_double
        L.orc_kat_radial.argtypes = [C.c_int, _dp, _dp, C.c_int, C.c_double, C.c_double, _dp, C.c_double]
        L.orc_kat_4d2ang.restype = C.c_double
        L.orc_kat_4d2ang.argtypes = [_dp, _dp, C.c_int, _dp, _dp, C.c_int, C.c_double, C.c_double, _dp, _dp,
                                     _dp, C.c_double]
        L.orc_kat_4d.restype = C.c_double
        L.orc_kat_4d.argtypes = [_dp, _dp, C.c_int, _dp, _dp, C.c_int, C.c_double, C.c_double, _dp, _dp, _dp,
                                 _dp, C.c_double]
        L.orc_kat_finiteT.restype = C.c_double
        L.orc_kat_finiteT.argtypes = [C.c_int, _dp, _dp, C.c_int, C.c_double, C.c_double, _dp, _dp, C.c_int,
                                      C.c_double, _dp, _dp, C.c_double]
        L.orc_kat_angle.restype = C.c_double
        L.orc_kat_angle.argtypes = [C.c_int, _dp, _dp, C.c_int, _dp, _dp, C.c_int, C.c_double, C.c_double, _dp, _dp,
                                    C.c_double]
        L.orc_kat_cartesian.restype = C.c_double
        L.orc_kat_cartesian.argtypes = [C.c_int, C.POINTER(_dp), C.POINTER(_dp), C.POINTER(C.c_int), _dp, _dp, _dp,
                                        C.c_double]
        L.orc_kat_freq.restype = C.c_double
        L.orc_kat_freq.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_int, _dp, _dp, C.c_int, _dp, _dp, C.c_int,
                                   C.c_double, C.c_double, C.c_double, _dp, _dp, _dp, _dp, C.c_double]
        L.orc_kat_finiteTx0.restype = C.c_double
        L.orc_kat_finiteTx0.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_int, _dp, _dp, C.c_int, _dp, _dp, C.c_int,
                                        C.c_double, C.c_double, C.c_int, C.c_double, C.c_double, _dp, _dp, _dp, _dp,
                                        C.c_double]
        L.orc_kat_4d_finiteT.restype = C.c_double
        L.orc_kat_4d_finiteT.argtypes = [_dp, _dp, C.c_int, _dp, _dp, C.c_int, C.c_double, C.c_double, _dp, _dp,
                                         C.c_int, C.c_double, _dp, _dp, _dp, _dp, C.c_double]
        L.orc_kat_angle_finiteT.restype = C.c_double
        L.orc_kat_angle_finiteT.argtypes = [C.c_int, _dp, _dp, C.c_int, _dp, _dp, C.c_int, C.c_double, C.c_double,
                                            _dp, _dp, C.c_int, C.c_double, _dp, _dp, _dp, C.c_double]
        L.orc_lpa_threshold.restype = C.c_double
        L.orc_lpa_threshold.argtypes = [C.c_double, C.c_double, C.c_double, _dp, _dp, C.c_int, C.c_double,
                                        C.c_double]
        L.orc_flow_lpa.argtypes = [C.POINTER(FlowParams), _dp, _dp, C.c_int, C.c_double, C.c_double,
                                   C.c_double, C.c_int, _dp]
        L.orc_flow_rhs.argtypes = [C.POINTER(FlowParams), _dp, C.c_double, _dp, _dp, C.c_int, C.c_double,
                                   C.c_double, _dp]
        L.orc_flow_lpa_many.argtypes = [C.POINTER(FlowParams), C.c_int, _dp, _dp, C.c_int, C.c_double, C.c_double,
                                        C.c_double, C.c_int, C.POINTER(_dp), C.POINTER(C.c_int)]
        L.orc_num_threads.restype = C.c_int
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


# ---------------------------------------------------------------- random / quadrature

def uniform01(seed: int, n: int) -> np.ndarray:
    out = np.empty(n)
    lib().orc_uniform01(C.c_uint64(seed), _p(out), n)
    return out


def gl_nodes(n: int):
    x, w = np.empty(n), np.empty(n)
    if lib().orc_gl_nodes(n, _p(x), _p(w)) != 0:
        raise ValueError("bad order")
    return x, w


def imtqlx(d, e, z):
    d, e, z = _f64(d).copy(), _f64(e).copy(), _f64(z).copy()
    if lib().orc_imtqlx(len(d), _p(d), _p(e), _p(z)) != 0:
        raise RuntimeError("IMTQLX iteration limit")
    return d, z


def make_quadrature(a, b, mu0):
    a, b = _f64(a).copy(), _f64(b).copy()
    n = len(a)
    x, w = np.empty(n), np.empty(n)
    if lib().orc_make_quadrature(n, _p(a), _p(b), mu0, _p(x), _p(w)) != 0:
        raise RuntimeError("IMTQLX iteration limit")
    return x, w


def matsubara_predict_size(T, E, step=2, vacuum_quad_size=48, precision_factor=0):
    return lib().orc_matsubara_predict_size(T, E, step, vacuum_quad_size, precision_factor)


def matsubara(T, E, step=2, min_size=0, max_size=1024, vacuum_quad_size=48, precision_factor=1):
    """MatsubaraQuadrature(T, E, step, min, max, vacuum, precision) -> (x, w, T_rule)."""
    cap = max(max_size, vacuum_quad_size + 2, min_size) + 2
    x, w = np.empty(cap), np.empty(cap)
    Teff = C.c_double()
    n = lib().orc_matsubara(T, E, step, min_size, max_size, vacuum_quad_size, precision_factor, _p(x), _p(w),
                            cap, C.byref(Teff))
    if n < 0:
        raise RuntimeError("matsubara rule failed")
    return x[:n].copy(), w[:n].copy(), Teff.value


def matsubara_storage_rule(T, E, min_size=32, vacuum_quad_size=48, add_matsubara_size=0):
    """The rule QuadratureProvider hands out for (T, E) (quadrature_provider.cu:31-60)."""
    return matsubara(T, E, 2, min_size, 128, vacuum_quad_size, add_matsubara_size)


def matsubara_sum(x, w, T, f):
    return T * f(0.0) + np.sum(w * (f(x) + f(-x)))


# ---------------------------------------------------------------- regulators / coordinates

def exp_RB(b, k2, q2):
    return lib().orc_exp_RB(b, k2, q2)


def exp_RBdot(b, k2, q2):
    return lib().orc_exp_RBdot(b, k2, q2)


def exp_dq2RB(b, k2, q2):
    return lib().orc_exp_dq2RB(b, k2, q2)


def litim_RB(k2, q2):
    return lib().orc_litim_RB(k2, q2)


def litim_RBdot(k2, q2):
    return lib().orc_litim_RBdot(k2, q2)


def logcoord(n, start, stop, bias):
    return LogCoord(start, stop, bias, n)


def logcoord_forward(c: LogCoord, i: int) -> float:
    return lib().orc_logcoord_forward(C.byref(c), i)


def logcoord_backward(c: LogCoord, y: float) -> float:
    return lib().orc_logcoord_backward(C.byref(c), y)


def lerp_log(data, c: LogCoord, y):
    data = _f64(data)
    return lib().orc_lerp_log(_p(data), C.byref(c), y)


def lerp_lin(data, start, stop, y):
    data = _f64(data)
    return lib().orc_lerp_lin(_p(data), len(data), start, stop, y)


# ---------------------------------------------------------------- integrals

def c0_threshold(m2, x, w, x_extent=20.0, k=1.0, b=1.0, d=4):
    m2, x, w = _f64(m2), _f64(x), _f64(w)
    out = np.empty(len(m2))
    lib().orc_c0_threshold(d, _p(m2), len(m2), _p(x), _p(w), len(x), x_extent, k, b, _p(out))
    return out


def c1_selfenergy(p, tables, coords, xq_list, xa_list, x_extent=20.0, k=1.0, b=1.0):
    """p[i], tables[i] (array), coords[i] (LogCoord), xq_list[i] = (x, w), xa_list[i] = (x, w)."""
    n = len(p)
    keep = []
    arr = (SEIntegral * n)()
    for i in range(n):
        xq, wq = (_f64(a) for a in xq_list[i])
        xa, wa = (_f64(a) for a in xa_list[i])
        tab = _f64(tables[i])
        keep += [xq, wq, xa, wa, tab]
        arr[i] = SEIntegral(_p(xq), _p(wq), len(xq), _p(xa), _p(wa), len(xa), _p(tab), coords[i], float(p[i]))
    out = np.empty(n)
    lib().orc_c1_selfenergy(arr, n, x_extent, k, b, _p(out))
    return out


def c2_finiteT(T, p0, pabs, m1, m2, mats, x, w, extent, k=1.0, b=1.0, d=4, radial_mode=1):
    """mats[i] = (mx, mw) Matsubara rule of integral i (positive nodes; T f(0) added when T>0)."""
    n = len(T)
    keep = []
    arr = (FTIntegral * n)()
    for i in range(n):
        mx, mw = (_f64(a) for a in mats[i])
        keep += [mx, mw]
        arr[i] = FTIntegral(float(T[i]), float(p0[i]), float(pabs[i]), float(m1[i]), float(m2[i]), _p(mx),
                            _p(mw), len(mx))
    x, w = _f64(x), _f64(w)
    out = np.empty(n)
    lib().orc_c2_finiteT(d, radial_mode, arr, n, _p(x), _p(w), len(x), extent, k, b, _p(out))
    return out


def kat_radial(d, x, w, x_extent, k, c6, constant):
    x, w, c6 = _f64(x), _f64(w), _f64(c6)
    return lib().orc_kat_radial(d, _p(x), _p(w), len(x), x_extent, k, _p(c6), constant)


def kat_4d2ang(xq, wq, xa, wa, x_extent, k, c6, a4, b4, constant):
    xq, wq, xa, wa, c6, a4, b4 = (_f64(v) for v in (xq, wq, xa, wa, c6, a4, b4))
    return lib().orc_kat_4d2ang(_p(xq), _p(wq), len(xq), _p(xa), _p(wa), len(xa), x_extent, k, _p(c6), _p(a4),
                                _p(b4), constant)


def kat_4d(xq, wq, xa, wa, x_extent, k, c6, a4, b4, p4, constant):
    xq, wq, xa, wa, c6, a4, b4, p4 = (_f64(v) for v in (xq, wq, xa, wa, c6, a4, b4, p4))
    return lib().orc_kat_4d(_p(xq), _p(wq), len(xq), _p(xa), _p(wa), len(xa), x_extent, k, _p(c6), _p(a4),
                            _p(b4), _p(p4), constant)


def kat_finiteT(d, x, w, x_extent, k, mx, mw, T, c6, d4, constant):
    x, w, mx, mw, c6, d4 = (_f64(v) for v in (x, w, mx, mw, c6, d4))
    return lib().orc_kat_finiteT(d, _p(x), _p(w), len(x), x_extent, k, _p(mx), _p(mw), len(mx), T, _p(c6), _p(d4),
                                 constant)


def kat_angle(d, xq, wq, xa, wa, x_extent, k, c6, a4, constant):
    xq, wq, xa, wa, c6, a4 = (_f64(v) for v in (xq, wq, xa, wa, c6, a4))
    return lib().orc_kat_angle(d, _p(xq), _p(wq), len(xq), _p(xa), _p(wa), len(xa), x_extent, k, _p(c6), _p(a4),
                               constant)


def kat_cartesian(rules, lo, hi, c, constant):
    """rules: [(x, w)] per axis (1 to 3); c: 4 cubic coefficients per axis."""
    D = len(rules)
    xs = [_f64(r[0]) for r in rules]
    ws = [_f64(r[1]) for r in rules]
    xp = (_dp * D)(*[_p(v) for v in xs])
    wp = (_dp * D)(*[_p(v) for v in ws])
    n = (C.c_int * D)(*[len(v) for v in xs])
    lo, hi, c = _f64(lo), _f64(hi), _f64(c)
    assert len(lo) == D and len(hi) == D and len(c) == 4 * D
    return lib().orc_kat_cartesian(D, xp, wp, n, _p(lo), _p(hi), _p(c), constant)


def kat_freq(dim, q0_axis, px0, xq, wq, xa, wa, xb, wb, x_extent, ext, k, c6, a4, p4, z4, constant):
    """Integrator3D (dim 3, no q0 axis) and Integrator{2,3}D{pq0,px0}: see orc_kat_freq."""
    xq, wq, xa, wa, xb, wb, c6, a4, p4, z4 = (_f64(v) for v in (xq, wq, xa, wa, xb, wb, c6, a4, p4, z4))
    return lib().orc_kat_freq(dim, int(q0_axis), int(px0), _p(xq), _p(wq), len(xq), _p(xa), _p(wa), len(xa), _p(xb),
                              _p(wb), len(xb), x_extent, ext, k, _p(c6), _p(a4), _p(p4), _p(z4), constant)


def kat_finiteTx0(kind, d, xq, wq, xa, wa, x0, w0, x_extent, x0_extent, summands, T, k, c6, a4, p4, z4, constant):
    """IntegratorFiniteTx0 (kind 0), IntegratorAngleFiniteTx0 (1), Integrator4DFiniteTx0 (2): see
    orc_kat_finiteTx0; x0, w0 is the integral's rule (its order is the grid's plus the summands at T = 0)."""
    xq, wq, xa, wa, x0, w0, c6, a4, p4, z4 = (_f64(v) for v in (xq, wq, xa, wa, x0, w0, c6, a4, p4, z4))
    return lib().orc_kat_finiteTx0(kind, d, _p(xq), _p(wq), len(xq), _p(xa), _p(wa), len(xa), _p(x0), _p(w0), len(x0),
                                   x_extent, x0_extent, summands, T, k, _p(c6), _p(a4), _p(p4), _p(z4), constant)


def kat_4d_finiteT(xq, wq, xa, wa, x_extent, k, mx, mw, T, c6, a4, p4, d5, constant):
    xq, wq, xa, wa, mx, mw, c6, a4, p4, d5 = (_f64(v) for v in (xq, wq, xa, wa, mx, mw, c6, a4, p4, d5))
    return lib().orc_kat_4d_finiteT(_p(xq), _p(wq), len(xq), _p(xa), _p(wa), len(xa), x_extent, k, _p(mx), _p(mw),
                                    len(mx), T, _p(c6), _p(a4), _p(p4), _p(d5), constant)


def kat_angle_finiteT(d, xq, wq, xa, wa, x_extent, k, mx, mw, T, c6, a4, d4, constant):
    xq, wq, xa, wa, mx, mw, c6, a4, d4 = (_f64(v) for v in (xq, wq, xa, wa, mx, mw, c6, a4, d4))
    return lib().orc_kat_angle_finiteT(d, _p(xq), _p(wq), len(xq), _p(xa), _p(wa), len(xa), x_extent, k, _p(mx),
                                       _p(mw), len(mx), T, _p(c6), _p(a4), _p(d4), constant)


def lpa_threshold(m2, T, k, x, w, x_extent=20.0, b=1.0):
    x, w = _f64(x), _f64(w)
    return lib().orc_lpa_threshold(m2, T, k, _p(x), _p(w), len(x), x_extent, b)


def flow_params(n_grid, rho_max, T, mu, N=4.0, m2_0=0.2, lam=1.0):
    return FlowParams(n_grid, rho_max, T, mu, N, m2_0, lam)


def flow_rhs(fp: FlowParams, u, k, x, w, x_extent=20.0, b=1.0):
    u, x, w = _f64(u), _f64(x), _f64(w)
    out = np.empty(fp.n_grid)
    lib().orc_flow_rhs(C.byref(fp), _p(u), k, _p(x), _p(w), len(x), x_extent, b, _p(out))
    return out


def flow_lpa(fp: FlowParams, x, w, x_extent=20.0, b=1.0, k_end=1e-3, n_steps=2000):
    x, w = _f64(x), _f64(w)
    u = np.empty(fp.n_grid)
    rc = lib().orc_flow_lpa(C.byref(fp), _p(x), _p(w), len(x), x_extent, b, k_end, n_steps, _p(u))
    if rc != 0:
        raise FloatingPointError("flow diverged")
    return u


def flow_lpa_many(fps, x, w, x_extent=20.0, b=1.0, k_end=1e-3, n_steps=2000):
    """Final u of several flows, one flow per OpenMP thread."""
    x, w = _f64(x), _f64(w)
    n = len(fps)
    arr = (FlowParams * n)(*fps)
    outs = [np.empty(fp.n_grid) for fp in fps]
    ptrs = (_dp * n)(*[_p(u) for u in outs])
    rc = (C.c_int * n)()
    lib().orc_flow_lpa_many(arr, n, _p(x), _p(w), len(x), x_extent, b, k_end, n_steps, ptrs, rc)
    if any(rc):
        raise FloatingPointError("flow diverged")
    return outs


def num_threads() -> int:
    return lib().orc_num_threads()
