"""ctypes binding of libcnhip.so (the C ABI declared in include/cnhip.h).

The library is loaded from this package directory only.  There is no fallback: a missing or
unloadable library raises ``BackendError`` at the first use, so nothing silently runs elsewhere.
"""
from __future__ import annotations

import ctypes
import threading
from pathlib import Path

from .errors import BackendError, LayoutError, SingularError, UnsupportedError

LIB_PATH = Path(__file__).resolve().with_name("libcnhip.so")

CN_OK, CN_EINVAL, CN_ESINGULAR, CN_EOVERFLOW, CN_ELOGIC, CN_EHIP, CN_EUNSUPPORTED = range(7)
CN_F64, CN_C128 = 0, 1

#: every symbol include/cnhip.h declares (tests check the library exports all of them)
EXPORTS = ("cn_last_error", "cn_device_count", "cn_version", "cn_bb_create", "cn_bb_factor",
           "cn_bb_reserve", "cn_bb_solve", "cn_bb_solve_mapped", "cn_bb_solve_split",
           "cn_bb_prepare_transpose", "cn_bb_props_get", "cn_bb_cond", "cn_bb_destroy",
           "cn_csr_create", "cn_csr_destroy", "cn_csr_matvec", "cn_sbdf_stage", "cn_rk_rhs",
           "cn_pointwise", "cn_rb2d_explicit", "cn_rb2d_explicit_many", "cn_bb_request_mode",
           "cn_bb_copy_systems", "cn_ch3d_workspace_bytes", "cn_ch3d_explicit", "cn_ch3d_explicit_many",
           "cn_bb_solve_noalloc", "cn_chpt_workspace_bytes", "cn_chpt_explicit", "cn_chpt_sbdf_stage",
           "cn_tr_rows", "cn_imex_create", "cn_imex_set_pointwise", "cn_imex_run", "cn_imex_reset",
           "cn_imex_info", "cn_imex_destroy")


class Props(ctypes.Structure):
    _fields_ = [("dtype", ctypes.c_int32), ("batch", ctypes.c_int32), ("n", ctypes.c_int32),
                ("kl", ctypes.c_int32), ("ku", ctypes.c_int32), ("rank", ctypes.c_int32),
                ("pivot", ctypes.c_int32), ("kl_kernel", ctypes.c_int32), ("wu_kernel", ctypes.c_int32),
                ("lanes", ctypes.c_int32), ("nrhs_reserved", ctypes.c_int32),
                ("transpose_ready", ctypes.c_int32), ("cond_max", ctypes.c_double),
                ("factor_ms", ctypes.c_double), ("factor_bytes", ctypes.c_int64),
                ("solve_mode", ctypes.c_int32)]


_lock = threading.Lock()
_lib = None

_vp = ctypes.c_void_p
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_dp = ctypes.POINTER(ctypes.c_double)


def _declare(lib):
    sig = {
        "cn_last_error": (ctypes.c_char_p, []),
        "cn_device_count": (ctypes.c_int, []),
        "cn_version": (ctypes.c_char_p, []),
        "cn_bb_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                        _i64p, _i32p, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "cn_bb_factor": (ctypes.c_int, [_vp, _vp, ctypes.c_int]),
        "cn_bb_reserve": (ctypes.c_int, [_vp, ctypes.c_int]),
        "cn_bb_solve": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int64, ctypes.c_int, _vp]),
        "cn_bb_solve_mapped": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
        "cn_bb_solve_noalloc": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int64, _vp]),
        "cn_bb_solve_split": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
        "cn_bb_prepare_transpose": (ctypes.c_int, [_vp]),
        "cn_bb_props_get": (ctypes.c_int, [_vp, ctypes.POINTER(Props)]),
        "cn_bb_cond": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_double)]),
        "cn_bb_destroy": (None, [_vp]),
        "cn_csr_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp), _i64p]),
        "cn_csr_destroy": (None, [_vp]),
        "cn_csr_matvec": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_double, ctypes.c_int, _vp]),
        "cn_sbdf_stage": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(_vp),
                                         ctypes.POINTER(_vp), ctypes.c_int, _dp, _dp, _vp, _vp]),
        "cn_rk_rhs": (ctypes.c_int, [ctypes.c_int, ctypes.c_int64, _vp, ctypes.POINTER(_vp),
                                     ctypes.POINTER(_vp), _dp, _dp, ctypes.c_int, _vp, _vp]),
        "cn_pointwise": (ctypes.c_int, [ctypes.c_int, ctypes.c_int64, _vp, _vp, _vp, _vp, _vp]),
        "cn_rb2d_explicit": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_double, _vp]),
        "cn_bb_request_mode": (ctypes.c_int, [ctypes.c_int]),
        "cn_bb_copy_systems": (ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int, _vp]),
        "cn_rb2d_explicit_many": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int,
                                                 ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                 ctypes.c_double, ctypes.c_int, _vp]),
        "cn_ch3d_workspace_bytes": (ctypes.c_size_t, [ctypes.c_int] * 6),
        "cn_chpt_workspace_bytes": (ctypes.c_size_t, [ctypes.c_int] * 6),
        "cn_chpt_explicit": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_size_t, _vp, _vp, _vp, _vp, _vp]
                             + [ctypes.c_int] * 6 + [ctypes.c_double] * 3 + [_vp]),
        "cn_chpt_sbdf_stage": (ctypes.c_int, [_vp, ctypes.POINTER(_vp), _vp, ctypes.POINTER(_vp), ctypes.c_int,
                                              _dp, _dp, _vp, _vp, _vp] + [ctypes.c_int] * 3
                               + [ctypes.c_double] * 2 + [_vp]),
        "cn_ch3d_explicit": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, ctypes.c_size_t, _vp, _vp, _vp, _vp]
                             + [ctypes.c_int] * 8 + [ctypes.c_double, ctypes.c_double, _vp]),
        "cn_ch3d_explicit_many": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, ctypes.c_size_t, _vp, _vp, _vp, _vp]
                                  + [ctypes.c_int] * 8 + [ctypes.c_double, ctypes.c_double, ctypes.c_int, _vp]),
        "cn_tr_rows": (ctypes.c_int, [_vp, _vp, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp,
                                      _vp]),
        "cn_imex_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, ctypes.c_int]
                           + [ctypes.POINTER(_vp)] * 3 + [_i64p] + [ctypes.POINTER(_vp)] * 3 + [_i64p]
                           + [_i32p, ctypes.c_int, ctypes.c_char_p, ctypes.c_double, ctypes.c_int]),
        "cn_imex_set_pointwise": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
        "cn_imex_run": (ctypes.c_int, [_vp, _vp, ctypes.c_double, ctypes.c_int, _dp, _vp]),
        "cn_imex_reset": (ctypes.c_int, [_vp]),
        "cn_imex_info": (ctypes.c_int, [_vp, _i32p, _i32p, _i64p]),
        "cn_imex_destroy": (None, [_vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args


def lib():
    """The loaded library (loaded once).  Raises BackendError when it is not built."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not LIB_PATH.exists():
                    raise BackendError(f"{LIB_PATH.name} is not built",
                                       where=str(LIB_PATH),
                                       fix="python -m conatus_amd.build (or __graft_entry__.build())")
                try:
                    # torch ships its own HIP runtime; loading it first makes libcnhip bind to the same
                    # runtime (same soname), so that both see the device and share streams.
                    import torch  # noqa: F401
                except ImportError:
                    pass
                try:
                    l = ctypes.CDLL(str(LIB_PATH))
                except OSError as e:
                    raise BackendError(f"cannot load {LIB_PATH}: {e}") from e
                _declare(l)
                _lib = l
    return _lib


def last_error() -> str:
    return lib().cn_last_error().decode(errors="replace")


def check(status: int, where: str) -> None:
    """Map a C status to the reference's exception types (include/cnhip.h)."""
    if status == CN_OK:
        return
    msg = last_error() or f"status {status}"
    if status == CN_EINVAL:
        raise LayoutError(msg, where=where)
    if status == CN_ESINGULAR:
        raise SingularError(msg, where=where)
    if status == CN_EOVERFLOW:
        raise OverflowError(f"{where}: {msg}")
    if status == CN_EUNSUPPORTED:
        raise UnsupportedError(msg, where=where)
    if status == CN_EHIP:
        raise BackendError(msg, where=where)
    raise BackendError(f"internal error: {msg}", where=where)


def device_count() -> int:
    return int(lib().cn_device_count())
