"""The C ABI: libdfrg.so loads and exports every symbol include/dfrg.h declares; host-side
error behaviour that needs no GPU."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "dfrg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dfrg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import diffrg_amd as D
    lib = C.CDLL(D.LIB_PATH)
    names = declared_symbols()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python view binds each of them
    for n in names:
        if n not in ("dfrg_last_error", "dfrg_fs_flow_bytes"):
            assert n in D.SIGNATURES, n


def test_calls_before_init_fail_cleanly():
    import diffrg_amd as D
    lib = D.lib()
    p = C.c_void_p()
    rc = lib.dfrg_malloc(C.byref(p), 16)
    assert rc in (D.ERR_NOT_INIT, D.ERR_NO_DEVICE, D.OK)
    if rc != D.OK:
        assert lib.dfrg_last_error()


def test_no_device_reported_without_gpu():
    import diffrg_amd as D
    n = C.c_int(-1)
    rc = D.lib().dfrg_device_count(C.byref(n))
    if rc == D.ERR_NO_DEVICE:
        assert n.value == 0
        with pytest.raises(D.DfrgError):
            D.init(0)
    else:
        assert rc == D.OK and n.value >= 1


def test_native_test_library_loads_without_undefined_symbols():
    """The template test library (tests/native/libdfrg_kat.so) resolves every symbol at load time;
    loading runs no GPU code."""
    import diffrg_amd as D
    kat = os.path.join(ROOT, "tests", "native", "libdfrg_kat.so")
    if not os.path.exists(kat):
        pytest.skip("libdfrg_kat.so not built")
    C.CDLL(D.LIB_PATH, mode=C.RTLD_GLOBAL)
    lib = C.CDLL(kat, mode=os.RTLD_NOW)
    for name in ("kat_radial", "kat_regulator_eval", "kat_thermo_eval", "kat_selfenergy_request_bench"):
        assert hasattr(lib, name), name
