"""The C ABI library: built, loadable without a GPU, and exporting every entry point that
include/cnhip.h declares (no compute calls here)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "cnhip.h").read_text()
    return sorted(set(re.findall(r"\b(cn_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_python_binding_set():
    from conatus_amd import _lib
    assert set(declared_symbols()) == set(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    from conatus_amd import _lib
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_library_loads_and_reports():
    from conatus_amd import _lib
    L = _lib.lib()
    assert b"gfx950" in L.cn_version()
    assert _lib.device_count() >= 0
    assert L.cn_last_error() is not None


def test_null_handle_is_an_error_not_a_crash():
    from conatus_amd import _lib
    from conatus_amd.errors import LayoutError
    L = _lib.lib()
    with pytest.raises(LayoutError):
        _lib.check(L.cn_bb_reserve(None, 1), "reserve")
    with pytest.raises(LayoutError):
        _lib.check(L.cn_bb_solve(None, None, 1, 0, 0, None), "solve")
    L.cn_bb_destroy(None)


def test_create_argument_checks_before_any_device_work():
    """Argument errors are reported before the device is touched (works without a GPU)."""
    import numpy as np
    from conatus_amd import _lib
    from conatus_amd.errors import LayoutError, UnsupportedError
    L = _lib.lib()
    h = ctypes.c_void_p()
    P = ctypes.c_void_p * 1
    ip = np.array([0, 1, 2], np.int32)
    ix = np.array([0, 1], np.int32)
    dv = np.array([1.0, 1.0])
    args = (P(ip.ctypes.data), P(ix.ctypes.data), P(dv.ctypes.data), (ctypes.c_int64 * 1)(2))
    with pytest.raises(LayoutError, match="empty batch"):
        _lib.check(L.cn_bb_create(ctypes.byref(h), 0, 0, 2, *args, None, 0, 1, 0), "create")
    with pytest.raises(UnsupportedError, match="column border"):
        _lib.check(L.cn_bb_create(ctypes.byref(h), 0, 1, 2, *args, None, 0, 1, 1), "create")
    bad = np.array([5], np.int32)
    with pytest.raises(LayoutError, match="border row out of range"):
        _lib.check(L.cn_bb_create(ctypes.byref(h), 0, 1, 2, *args,
                                  bad.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), 1, 1, 0), "create")
    ixb = np.array([0, 7], np.int32)
    args_b = (P(ip.ctypes.data), P(ixb.ctypes.data), P(dv.ctypes.data), (ctypes.c_int64 * 1)(2))
    with pytest.raises(LayoutError, match="column index 7"):
        _lib.check(L.cn_bb_create(ctypes.byref(h), 0, 1, 2, *args_b, None, 0, 1, 0), "create")
    dnan = np.array([1.0, np.nan])
    args_n = (P(ip.ctypes.data), P(ix.ctypes.data), P(dnan.ctypes.data), (ctypes.c_int64 * 1)(2))
    with pytest.raises(LayoutError, match="non-finite"):
        _lib.check(L.cn_bb_create(ctypes.byref(h), 0, 1, 2, *args_n, None, 0, 1, 0), "create")
