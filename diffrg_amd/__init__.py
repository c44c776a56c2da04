"""diffrg_amd -- MI355X (gfx950) runtime for DiFfRG's momentum-loop integrators.

Python view of libdfrg.so's C ABI (include/dfrg.h).  The C++ drop-in headers under
include/DiFfRG/ are the primary interface for compiled call sites; this module serves the
tests, the benchmark and Python callers.  There is no CPU fallback: if libdfrg.so is missing
or no GPU is visible, the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Sequence

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libdfrg.so")

FP64 = 0
MIXED = 1

OK = 0
ERR_INVALID = -1
ERR_RUNTIME = -2
ERR_HIP = -3
ERR_NO_DEVICE = -4
ERR_NOT_INIT = -5


class DfrgError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class DfrgInvalidArgument(DfrgError, ValueError):
    pass


_lib = None
_vp = C.c_void_p
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

# (name, argtypes) of every symbol include/dfrg.h declares
SIGNATURES = {
    "dfrg_init": [C.c_int],
    "dfrg_finalize": [],
    "dfrg_last_error": [],
    "dfrg_device_count": [_ip],
    "dfrg_device_props": [C.c_char_p, C.c_int, _ip, _ip, C.POINTER(C.c_ulonglong)],
    "dfrg_malloc": [C.POINTER(_vp), C.c_size_t],
    "dfrg_free": [_vp],
    "dfrg_host_alloc": [C.POINTER(_vp), C.c_size_t],
    "dfrg_host_free": [_vp],
    "dfrg_memcpy_h2d": [_vp, _vp, C.c_size_t, _vp],
    "dfrg_memcpy_d2h": [_vp, _vp, C.c_size_t, _vp],
    "dfrg_memcpy_d2d": [_vp, _vp, C.c_size_t, _vp],
    "dfrg_thread_stream": [C.POINTER(_vp)],
    "dfrg_stream_create": [C.POINTER(_vp)],
    "dfrg_stream_destroy": [_vp],
    "dfrg_stream_sync": [_vp],
    "dfrg_event_create": [C.POINTER(_vp)],
    "dfrg_event_destroy": [_vp],
    "dfrg_event_record": [_vp, _vp],
    "dfrg_event_sync": [_vp],
    "dfrg_event_elapsed_ms": [_vp, _vp, C.POINTER(C.c_float)],
    "dfrg_graph_begin": [_vp],
    "dfrg_graph_end": [_vp, C.POINTER(_vp)],
    "dfrg_graph_launch": [_vp, _vp],
    "dfrg_graph_destroy": [_vp],
    "dfrg_pool_alloc": [C.POINTER(_vp), C.c_size_t],
    "dfrg_pool_free": [_vp],
    "dfrg_pool_host_alloc": [C.POINTER(_vp), C.c_size_t],
    "dfrg_pool_host_free": [_vp],
    "dfrg_gl_generate": [C.c_int, _vp, _vp, _vp],
    "dfrg_matsubara_generate": [C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp,
                                C.c_int, _ip, _dp, _vp],
    "dfrg_qp_create": [C.POINTER(_vp), C.c_int, C.c_int, C.c_int],
    "dfrg_qp_destroy": [_vp],
    "dfrg_qp_gl": [_vp, C.c_int, C.c_int, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp)],
    "dfrg_qp_matsubara": [_vp, C.c_double, C.c_double, C.c_int, _ip, _dp, C.POINTER(_vp), C.POINTER(_vp),
                          C.POINTER(_vp), C.POINTER(_vp)],
    "dfrg_valu_peak": [C.c_int, C.POINTER(C.c_double)],
    "dfrg_threshold_batch": [C.c_int, _vp, C.c_int, _vp, _vp, C.c_int, C.c_double, C.c_double, C.c_int, _vp, _vp],
    "dfrg_se_batch_create": [C.POINTER(_vp), _vp, C.c_int, _dp, _ip, _ip, _ip, _ip, _dp, _dp, _dp, _dp, C.c_int,
                             C.c_double, C.c_double],
    "dfrg_se_batch_destroy": [_vp],
    "dfrg_se_batch_set_shape": [_vp, C.c_int, C.c_int, C.c_int],
    "dfrg_se_batch_get_shape": [_vp, _ip, _ip, _ip, _ip],
    "dfrg_se_batch_run": [_vp, C.c_int, _vp, _vp],
    "dfrg_se_batch_autotune": [_vp, C.c_int, C.c_int, C.c_char_p, _ip, _ip, _ip, C.POINTER(C.c_float)],
    "dfrg_se_batch_graph_launch": [_vp, C.c_int, _vp, C.c_int, _vp],
    "dfrg_se_batch_evals": [_vp, C.POINTER(C.c_ulonglong)],
    "dfrg_se_batch_variants": [_vp, C.c_int, _ip],
    "dfrg_se_batch_unroll": [_vp, _ip],
    "dfrg_se_batch_run_noalloc": [_vp, C.c_int, _vp, _vp],
    "dfrg_ft_set_block": [C.c_int],
    "dfrg_ft_batch": [C.c_int, C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_int,
                      C.c_double, C.c_double, C.c_int, _vp, _vp],
    "dfrg_fs_create": [C.POINTER(_vp), _vp, C.c_size_t, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int],
    "dfrg_fs_destroy": [_vp],
    "dfrg_fs_submit": [_vp, _vp, C.POINTER(C.c_longlong)],
    "dfrg_fs_progress": [_vp, _ip, _ip, _ip],
    "dfrg_fs_poll": [_vp, C.c_longlong, _ip],
    "dfrg_fs_result": [_vp, C.c_longlong, _dp, C.c_int, _dp],
    "dfrg_fs_stats": [_vp, _ip, _dp, _dp, _dp],
    "dfrg_fs_timing": [_vp, C.POINTER(C.c_longlong), _dp, C.POINTER(C.c_longlong)],
    "dfrg_table_create": [C.POINTER(_vp), C.c_int, C.c_double, C.c_double, C.c_double, _dp],
    "dfrg_table_update": [_vp, _dp, _vp],
    "dfrg_table_view": [_vp, C.POINTER(_vp), _ip],
    "dfrg_table_eval": [_vp, _vp, C.c_int, _vp, _vp],
    "dfrg_table_destroy": [_vp],
}


def lib() -> C.CDLL:
    """Load libdfrg.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m diffrg_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        L.dfrg_last_error.restype = C.c_char_p
        L.dfrg_fs_flow_bytes.restype = C.c_size_t
        L.dfrg_fs_flow_bytes.argtypes = [C.c_int]
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != OK:
        msg = lib().dfrg_last_error().decode(errors="replace")
        if rc == ERR_INVALID:
            raise DfrgInvalidArgument(rc, msg)
        raise DfrgError(rc, msg)


def init(device: int = 0) -> None:
    check(lib().dfrg_init(device))


def finalize() -> None:
    check(lib().dfrg_finalize())


def device_props() -> dict:
    name = C.create_string_buffer(256)
    cu, clk, hbm = C.c_int(), C.c_int(), C.c_ulonglong()
    check(lib().dfrg_device_props(name, 256, C.byref(cu), C.byref(clk), C.byref(hbm)))
    return {"name": name.value.decode(), "cu_count": cu.value, "clock_khz": clk.value, "hbm_bytes": hbm.value}


def valu_peak(fp64: bool = True) -> float:
    """Measured vector FMA rate of the device in FLOP/s (independent FMA chains, all SIMDs busy)."""
    r = C.c_double()
    check(lib().dfrg_valu_peak(1 if fp64 else 0, C.byref(r)))
    return r.value


# ------------------------------------------------------------------------------- device memory

class DeviceArray:
    """HBM buffer owned by Python (dfrg_malloc / dfrg_free)."""

    def __init__(self, n: int, dtype=np.float64):
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        self.nbytes = max(1, self.n * self.dtype.itemsize)
        p = _vp()
        check(lib().dfrg_malloc(C.byref(p), self.nbytes))
        self.ptr = p.value

    @classmethod
    def from_host(cls, a, dtype=None, stream=None) -> "DeviceArray":
        a = np.ascontiguousarray(a, dtype=dtype)
        d = cls(a.size, a.dtype)
        if a.size:
            check(lib().dfrg_memcpy_h2d(d.ptr, a.ctypes.data, a.nbytes, stream))
            check(lib().dfrg_stream_sync(stream))
        return d

    def to_host(self, stream=None) -> np.ndarray:
        out = np.empty(self.n, self.dtype)
        if self.n:
            check(lib().dfrg_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes, stream))
            check(lib().dfrg_stream_sync(stream))
        return out

    def free(self) -> None:
        if self.ptr:
            lib().dfrg_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _need(a: "DeviceArray", n: int, dtype, what: str) -> None:
    """A device argument must hold at least n values of dtype: the kernels index it without bounds."""
    if not isinstance(a, DeviceArray) or a.ptr is None:
        raise ValueError(f"{what}: a live DeviceArray is required")
    if a.dtype != np.dtype(dtype):
        raise ValueError(f"{what}: dtype {a.dtype} where {np.dtype(dtype)} is required")
    if a.n < n:
        raise ValueError(f"{what}: holds {a.n} values, {n} are required")


def memcpy_d2d(dst: "DeviceArray", src: "DeviceArray", stream=None) -> None:
    """Device-to-device copy of src into dst (enqueued on `stream`)."""
    if dst.dtype != src.dtype or dst.n < src.n:
        raise ValueError("memcpy_d2d: destination smaller than source or of another dtype")
    check(lib().dfrg_memcpy_d2d(dst.ptr, src.ptr, src.n * src.dtype.itemsize, stream))


def stream_sync(stream=None) -> None:
    check(lib().dfrg_stream_sync(stream))


class Stream:
    """A non-blocking HIP stream owned by the library (dfrg_stream_create)."""

    def __init__(self):
        h = _vp()
        check(lib().dfrg_stream_create(C.byref(h)))
        self.ptr = h.value

    def synchronize(self):
        check(lib().dfrg_stream_sync(self.ptr))

    def close(self):
        if self.ptr:
            lib().dfrg_stream_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Event:
    def __init__(self):
        p = _vp()
        check(lib().dfrg_event_create(C.byref(p)))
        self.ptr = p.value

    def record(self, stream=None):
        check(lib().dfrg_event_record(self.ptr, stream))

    def synchronize(self):
        check(lib().dfrg_event_sync(self.ptr))

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        check(lib().dfrg_event_elapsed_ms(self.ptr, end.ptr, C.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            lib().dfrg_event_destroy(self.ptr)
        except Exception:
            pass


class Graph:
    """Work enqueued by `fn()` (called `times` times on `stream`, NULL: the thread's stream)
    captured into a hipGraph; launch() replays all of it with one launch.  fn must only enqueue
    (the batched entry points do); a call that synchronises fails the capture."""

    def __init__(self, fn, times=1, stream=None):
        self.ptr = None
        check(lib().dfrg_graph_begin(stream))
        err = None
        try:
            for _ in range(times):
                fn()
        except Exception as e:  # the capture must end either way
            err = e
        g = _vp()
        rc = lib().dfrg_graph_end(stream, C.byref(g))
        if err is not None:
            if g.value:
                lib().dfrg_graph_destroy(g)
            raise err
        check(rc)
        self.ptr = g.value

    def launch(self, stream=None):
        check(lib().dfrg_graph_launch(self.ptr, stream))

    def __del__(self):
        try:
            if self.ptr:
                lib().dfrg_graph_destroy(self.ptr)
        except Exception:
            pass


# ------------------------------------------------------------------------------- quadrature

def gl_generate(order: int):
    """FP64 Gauss-Legendre rule on [0,1] generated on the GPU -> (x, w) host copies."""
    x, w = DeviceArray(order), DeviceArray(order)
    check(lib().dfrg_gl_generate(order, x.ptr, w.ptr, None))
    return x.to_host(), w.to_host()


def matsubara_generate(T, E, step=2, min_size=0, max_size=1024, vacuum_quad_size=48, precision_factor=1):
    """MatsubaraQuadrature(T, E, ...) generated on the GPU -> (x, w, T_rule)."""
    cap = max(max_size, vacuum_quad_size + 2, min_size) + 2
    x, w = DeviceArray(cap), DeviceArray(cap)
    n, Tr = C.c_int(), C.c_double()
    check(lib().dfrg_matsubara_generate(T, E, step, min_size, max_size, vacuum_quad_size, precision_factor, x.ptr,
                                        w.ptr, cap, C.byref(n), C.byref(Tr), None))
    return x.to_host()[: n.value], w.to_host()[: n.value], Tr.value


class Table:
    """Interpolation table on logarithmic coordinates in HBM (dfrg_table_*; the data of a
    LinearInterpolator1D<double, LogarithmicCoordinates1D<double>>)."""

    def __init__(self, values, start: float, stop: float, bias: float):
        v = np.ascontiguousarray(values, dtype=np.float64)
        h = _vp()
        check(lib().dfrg_table_create(C.byref(h), len(v), start, stop, bias, v.ctypes.data_as(_dp)))
        self.ptr, self.n = h, len(v)

    def update(self, values, stream=None) -> None:
        v = np.ascontiguousarray(values, dtype=np.float64)
        if len(v) != self.n:
            raise ValueError("update needs as many values as the table holds")
        check(lib().dfrg_table_update(self.ptr, v.ctypes.data_as(_dp), stream))
        check(lib().dfrg_stream_sync(stream))

    def view(self) -> int:
        d, n = _vp(), C.c_int()
        check(lib().dfrg_table_view(self.ptr, C.byref(d), C.byref(n)))
        return d.value

    def eval(self, d_y: "DeviceArray", d_out: "DeviceArray", stream=None) -> None:
        _need(d_y, 0, np.float64, "y")
        _need(d_out, d_y.n, np.float64, "out")
        check(lib().dfrg_table_eval(self.ptr, d_y.ptr, d_y.n, d_out.ptr, stream))

    def close(self) -> None:
        if self.ptr:
            lib().dfrg_table_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class QuadratureProvider:
    """Mirror of DiFfRG's QuadratureProvider (quadrature_provider.hh:138-248)."""

    def __init__(self, vacuum_quad_size=48, min_matsubara_size=32, add_matsubara_size=0):
        p = _vp()
        check(lib().dfrg_qp_create(C.byref(p), vacuum_quad_size, min_matsubara_size, add_matsubara_size))
        self.ptr = p.value

    def _gl(self, order, fp32=0):
        dx, dw, hx, hw = _vp(), _vp(), _vp(), _vp()
        check(lib().dfrg_qp_gl(self.ptr, order, fp32, C.byref(dx), C.byref(dw), C.byref(hx), C.byref(hw)))
        return dx.value, dw.value, hx.value, hw.value

    def get_points(self, order: int) -> np.ndarray:
        return np.ctypeslib.as_array(C.cast(self._gl(order)[2], _dp), (order,)).copy()

    def get_weights(self, order: int) -> np.ndarray:
        return np.ctypeslib.as_array(C.cast(self._gl(order)[3], _dp), (order,)).copy()

    def get_device_points(self, order: int) -> int:
        return self._gl(order)[0]

    def get_device_weights(self, order: int) -> int:
        return self._gl(order)[1]

    def _mats(self, T, E):
        n, Tr = C.c_int(), C.c_double()
        dx, dw, hx, hw = _vp(), _vp(), _vp(), _vp()
        check(lib().dfrg_qp_matsubara(self.ptr, T, E, 0, C.byref(n), C.byref(Tr), C.byref(dx), C.byref(dw),
                                      C.byref(hx), C.byref(hw)))
        return n.value, Tr.value, dx.value, dw.value, hx.value, hw.value

    def get_matsubara_points(self, T, E) -> np.ndarray:
        n, _, _, _, hx, _ = self._mats(T, E)
        return np.ctypeslib.as_array(C.cast(hx, _dp), (n,)).copy()

    def get_matsubara_weights(self, T, E) -> np.ndarray:
        n, _, _, _, _, hw = self._mats(T, E)
        return np.ctypeslib.as_array(C.cast(hw, _dp), (n,)).copy()

    def get_device_matsubara(self, T, E):
        n, Tr, dx, dw, _, _ = self._mats(T, E)
        return n, Tr, dx, dw

    def close(self):
        if self.ptr:
            lib().dfrg_qp_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------------------- batched integrals

def threshold_batch(d_m2: DeviceArray, d_x: DeviceArray, d_w: DeviceArray, d_out: DeviceArray, x_extent=20.0,
                    k=1.0, mode=FP64, d=4, stream=None) -> None:
    # only the output is checked here (one check per call keeps this launch-sized call cheap)
    _need(d_out, d_m2.n, np.float64, "out")
    check(lib().dfrg_threshold_batch(d, d_m2.ptr, d_m2.n, d_x.ptr, d_w.ptr, d_x.n, x_extent, k, mode, d_out.ptr,
                                     stream))


def ft_batch(dev: dict, d_out: DeviceArray, extent, k=1.0, mode=FP64, d=4, radial_mode=1, stream=None) -> None:
    """dev holds DeviceArrays T, p0, pabs, m1, m2, mx, mw, m_off (int32), m_n (int32), x, w."""
    # only the output is checked here (one check per call keeps this launch-sized call cheap); the
    # inputs are the caller's: per-external arrays of n values, m_off/m_n int32 into mx/mw
    _need(d_out, dev["T"].n, np.float64, "out")
    check(lib().dfrg_ft_batch(d, radial_mode, dev["T"].n, dev["T"].ptr, dev["p0"].ptr, dev["pabs"].ptr,
                              dev["m1"].ptr, dev["m2"].ptr, dev["mx"].ptr, dev["mw"].ptr, dev["m_off"].ptr,
                              dev["m_n"].ptr, dev["x"].ptr, dev["w"].ptr, dev["x"].n, extent, k, mode, d_out.ptr,
                              stream))


def ft_set_block(block: int = 0) -> None:
    """Threads per integral of ft_batch (64/128/256/512; 0 = automatic)."""
    check(lib().dfrg_ft_set_block(block))


class SEBatch:
    """Two-angle self-energy integrals of uneven orders evaluated in one launch."""

    def __init__(self, qp: QuadratureProvider, p, nq, na, tables: Sequence[np.ndarray], start, stop, bias,
                 x_extent=20.0, k=1.0):
        self.ptr = None  # until the library object exists
        n = len(p)
        arr = lambda v, dt: np.ascontiguousarray(np.broadcast_to(np.asarray(v, dt), (n,)))
        # one copy of each distinct table: integrals that pass the same table object on the same
        # coordinates share its cells (one staging per table in HBM and L2)
        start_a, stop_a, bias_a = arr(start, np.float64), arr(stop, np.float64), arr(bias, np.float64)
        tab_n = np.array([len(t) for t in tables], np.int32)
        tab_off = np.zeros(n, np.int32)
        seen, parts, size = {}, [], 0
        for i, t in enumerate(tables):
            key = (id(t), float(start_a[i]), float(stop_a[i]), float(bias_a[i]))
            if key not in seen:
                seen[key] = size
                parts.append(np.asarray(t, np.float64))
                size += len(t)
            tab_off[i] = seen[key]
        # an empty call set reaches the library, which rejects it with its own message
        flat = np.ascontiguousarray(np.concatenate(parts or [np.zeros(0)]))
        self._keep = [arr(p, np.float64), arr(nq, np.int32), arr(na, np.int32), tab_off, tab_n, start_a, stop_a,
                      bias_a, flat]
        P = lambda a, t: a.ctypes.data_as(C.POINTER(t))
        k_ = self._keep
        h = _vp()
        check(lib().dfrg_se_batch_create(C.byref(h), qp.ptr, n, P(k_[0], C.c_double), P(k_[1], C.c_int),
                                         P(k_[2], C.c_int), P(k_[3], C.c_int), P(k_[4], C.c_int),
                                         P(k_[5], C.c_double), P(k_[6], C.c_double), P(k_[7], C.c_double),
                                         P(k_[8], C.c_double), len(flat), x_extent, k))
        self.ptr = h.value
        self.n = n

    @property
    def evals(self) -> int:
        e = C.c_ulonglong()
        check(lib().dfrg_se_batch_evals(self.ptr, C.byref(e)))
        return e.value

    def variants(self, mode: int) -> np.ndarray:
        """Integrand form a run in `mode` uses per integral (0 FP64, 1 mixed, 2 mixed narrow)."""
        v = np.empty(self.n, dtype=np.int32)
        check(lib().dfrg_se_batch_variants(self.ptr, mode, v.ctypes.data_as(C.POINTER(C.c_int))))
        return v

    def unroll(self, mode: int) -> int:
        """Radial unroll of the paired evaluation loop a run in `mode` uses at the current shape."""
        if mode == FP64:
            return 1
        u = C.c_int()
        check(lib().dfrg_se_batch_unroll(self.ptr, C.byref(u)))
        return u.value

    def set_shape(self, block=256, tile_evals=8192, grid=0):
        check(lib().dfrg_se_batch_set_shape(self.ptr, block, tile_evals, grid))

    def shape(self) -> dict:
        b, t, g, n = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        check(lib().dfrg_se_batch_get_shape(self.ptr, C.byref(b), C.byref(t), C.byref(g), C.byref(n)))
        return {"block": b.value, "tile_evals": t.value, "grid": g.value, "n_tiles": n.value}

    def run(self, d_out: DeviceArray, mode=FP64, stream=None):
        _need(d_out, self.n, np.float64, "out")
        check(lib().dfrg_se_batch_run(self.ptr, mode, d_out.ptr, stream))

    def run_noalloc(self, d_out: DeviceArray, mode, stream):
        """run() with the library's static scratch: allocates nothing; `stream` (a Stream or a
        raw handle) is required.  Calls that overlap on one GPU race."""
        _need(d_out, self.n, np.float64, "out")
        h = stream.ptr if isinstance(stream, Stream) else stream
        check(lib().dfrg_se_batch_run_noalloc(self.ptr, mode, d_out.ptr, h))

    def graph_launch(self, d_out: DeviceArray, mode=FP64, launches=1, stream=None):
        _need(d_out, self.n, np.float64, "out")
        check(lib().dfrg_se_batch_graph_launch(self.ptr, mode, d_out.ptr, launches, stream))

    def autotune(self, mode=FP64, reps=3, cache_path=None) -> dict:
        b, t, g, ms = C.c_int(), C.c_int(), C.c_int(), C.c_float()
        cp = cache_path.encode() if cache_path else None
        check(lib().dfrg_se_batch_autotune(self.ptr, mode, reps, cp, C.byref(b), C.byref(t), C.byref(g),
                                           C.byref(ms)))
        return {"block": b.value, "tile_evals": t.value, "grid": g.value, "ms": ms.value}

    def close(self):
        if self.ptr:
            lib().dfrg_se_batch_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------------------- flow server

class FlowDesc(C.Structure):
    _fields_ = [("n_grid", C.c_int), ("rho_max", C.c_double), ("T", C.c_double), ("mu", C.c_double),
                ("N", C.c_double), ("m2_0", C.c_double), ("lambda_", C.c_double)]


QUEUED, RESIDENT, FINISHED, FAILED = 0, 1, 2, 3


def flow_bytes(n_grid: int) -> int:
    return lib().dfrg_fs_flow_bytes(n_grid)


class FlowServer:
    """Continuous-batching server of O(N) LPA flows (dfrg_fs_*)."""

    def __init__(self, qp: QuadratureProvider, hbm_budget: int, max_resident: int = 64, nx: int = 64,
                 n_steps: int = 2000, k_end: float = 1e-3, steps_per_launch: int = 10):
        h = _vp()
        check(lib().dfrg_fs_create(C.byref(h), qp.ptr, hbm_budget, max_resident, nx, n_steps, k_end,
                                   steps_per_launch))
        self.ptr = h.value
        self.n_grid = {}

    def submit(self, n_grid, T, mu=0.0, rho_max=4.0, N=4.0, m2_0=0.2, lam=1.0) -> int:
        d = FlowDesc(int(n_grid), rho_max, T, mu, N, m2_0, lam)
        fid = C.c_longlong()
        check(lib().dfrg_fs_submit(self.ptr, C.byref(d), C.byref(fid)))
        self.n_grid[fid.value] = int(n_grid)
        return fid.value

    def progress(self):
        r, q, d = C.c_int(), C.c_int(), C.c_int()
        check(lib().dfrg_fs_progress(self.ptr, C.byref(r), C.byref(q), C.byref(d)))
        return r.value, q.value, d.value

    def poll(self, fid: int) -> int:
        s = C.c_int()
        check(lib().dfrg_fs_poll(self.ptr, fid, C.byref(s)))
        return s.value

    def result(self, fid: int):
        n = self.n_grid[fid]
        u = np.empty(n)
        lat = C.c_double()
        check(lib().dfrg_fs_result(self.ptr, fid, u.ctypes.data_as(_dp), n, C.byref(lat)))
        return u, lat.value

    def stats(self) -> dict:
        f, p50, p99, busy = C.c_int(), C.c_double(), C.c_double(), C.c_double()
        check(lib().dfrg_fs_stats(self.ptr, C.byref(f), C.byref(p50), C.byref(p99), C.byref(busy)))
        return {"finished": f.value, "p50_s": p50.value, "p99_s": p99.value, "busy_s": busy.value}

    def timing(self) -> dict:
        """Graph launches, their device time (HIP events) and the grid-point steps they advanced."""
        n, ps, t = C.c_longlong(), C.c_longlong(), C.c_double()
        check(lib().dfrg_fs_timing(self.ptr, C.byref(n), C.byref(t), C.byref(ps)))
        return {"launches": n.value, "device_s": t.value, "point_steps": ps.value}

    def close(self):
        if self.ptr:
            lib().dfrg_fs_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
