"""Bordered banded solvers on the MI355X: the host-array API of the reference's compiled core.

The classes ``BorderedBanded`` and ``BorderedBandedBatch`` and the helpers ``csr_args``,
``solve_operand``, ``threads_for`` and ``solve_threads`` keep the names, signatures and error
types of python/conatus/runtime/banded.py:28-213, so code written against the reference runs
unchanged; underneath, every factorization and solve is a libcnhip.so launch on the GPU.  Host
arrays are staged to device memory, solved there and copied back (the reference returns fresh
arrays as well).  Without a GPU the constructors raise ``BackendError``; there is no CPU path.
``BatchHandle`` owns one ``cn_bb`` handle and is what the device-side API builds on.
"""
from __future__ import annotations

import ctypes
from collections.abc import Sequence

import numpy as np
import scipy.sparse as sp

from .. import _lib
from ..errors import InputTypeError, OptionError
from .checks import (SOLVE_THREAD_UNKNOWNS, check_rhs_shape, csr_args, factor_dtype, solve_operand,  # noqa: F401
                     solve_threads, threads_for)

__all__ = ["BorderedBanded", "BorderedBandedBatch", "csr_args", "solve_operand", "solve_threads",
           "threads_for", "BatchHandle"]

# the private names other modules of this package import
_factor_dtype = factor_dtype
_check_rhs_shape = check_rhs_shape


#: kernel modes a handle can be asked for (cn_bb_request_mode) and the solve modes it reports
#: (cn_bb_props.solve_mode): one lane per (system, rhs); one wavefront per (system, rhs) (wave8,
#: wave32: with staged blocks of 8 or 32 steps, else by the launch size); one workgroup per
#: (system, rhs) with the sweeps partitioned over its wavefronts
MODE_REQUEST = {None: 0, "lane": 1, "wave": 2, "part": 3, "wave8": 4, "wave32": 5}
SOLVE_MODES = {0: "lane", 1: "wave", 2: "part"}


class BatchHandle:
    """Owner of one ``cn_bb`` handle: a factored batch resident on the current HIP device."""

    def __init__(self, csrs: Sequence[tuple], border: Sequence[int], dtype, pivot: bool = True,
                 col_border: bool = False, mode: str | None = None):
        L = _lib.lib()
        if mode not in MODE_REQUEST:
            raise OptionError(f"BandedBatch: mode must be None or one of {sorted(k for k in MODE_REQUEST if k)}, "
                              f"got {mode!r}")
        self.dtype = np.dtype(dtype)
        B = len(csrs)
        if B == 0:
            raise InputTypeError("BandedBatch: empty batch", fix="pass at least one matrix (one per mode/group)")
        n = int(csrs[0][0])
        for k, c in enumerate(csrs):
            if int(c[0]) != n:
                raise OptionError("BandedBatch: every mode must have the same size n",
                                  where=f"matrix {k} has n = {c[0]}, matrix 0 has n = {n}")
        self._keep = [(c[1], c[2], np.ascontiguousarray(c[3], dtype=self.dtype)) for c in csrs]
        P = ctypes.c_void_p * B
        ip = P(*[a[0].ctypes.data for a in self._keep])
        ix = P(*[a[1].ctypes.data for a in self._keep])
        dv = P(*[a[2].ctypes.data for a in self._keep])
        nnz = (ctypes.c_int64 * B)(*[int(a[1].size) for a in self._keep])
        bord = np.ascontiguousarray(np.asarray([int(i) for i in border], dtype=np.int32))
        h = ctypes.c_void_p()
        code = _lib.CN_C128 if self.dtype == np.complex128 else _lib.CN_F64
        _lib.check(L.cn_bb_request_mode(MODE_REQUEST[mode]), "BandedBatch")
        try:
            st = L.cn_bb_create(ctypes.byref(h), code, B, n, ip, ix, dv, nnz,
                                bord.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), int(bord.size),
                                1 if pivot else 0, 1 if col_border else 0)
        finally:
            L.cn_bb_request_mode(0)
        self._keep = None          # the library copied what it needs
        _lib.check(st, "BandedBatch")
        self._h = h
        self.border = [int(i) for i in border]
        self.col_border = bool(col_border)

    def props(self) -> _lib.Props:
        p = _lib.Props()
        _lib.check(_lib.lib().cn_bb_props_get(self._h, ctypes.byref(p)), "BandedBatch.props")
        return p

    def cond(self) -> np.ndarray:
        p = self.props()
        out = np.empty(p.batch, np.float64)
        _lib.check(_lib.lib().cn_bb_cond(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                   "BandedBatch.cond")
        return out

    def reserve(self, nrhs: int) -> None:
        _lib.check(_lib.lib().cn_bb_reserve(self._h, int(nrhs)), "BandedBatch.reserve")

    def factor(self, stream: int = 0, check: bool = True) -> None:
        _lib.check(_lib.lib().cn_bb_factor(self._h, ctypes.c_void_p(stream), 1 if check else 0),
                   "BandedBatch.factor")

    def solve_ptr(self, ptr: int, nrhs: int, rhs_stride: int, transpose: bool, stream: int) -> None:
        _lib.check(_lib.lib().cn_bb_solve(self._h, ctypes.c_void_p(ptr), int(nrhs), int(rhs_stride),
                                          1 if transpose else 0, ctypes.c_void_p(stream)),
                   "BandedBatch.solve_ptr")

    def copy_systems_from(self, src: "BatchHandle", src_first: int, dst_first: int, count: int, stream: int) -> None:
        """Device copy of factored systems src_first .. of ``src`` into systems dst_first .. here
        (cn_bb_copy_systems): a run's mode systems moved into a slot without refactorization."""
        _lib.check(_lib.lib().cn_bb_copy_systems(self._h, int(dst_first), src._h, int(src_first), int(count),
                                                 ctypes.c_void_p(stream)), "BandedBatch.copy_systems_from")

    def solve_mapped_ptr(self, rhs: int, rows: int, x: int, cols: int, stream: int) -> None:
        v = ctypes.c_void_p
        _lib.check(_lib.lib().cn_bb_solve_mapped(self._h, v(rhs), v(rows or None), v(x), v(cols or None),
                                                 v(stream)), "BandedBatch.solve_mapped_ptr")

    def solve_split_ptr(self, re: int, rows_c: int, rows_s: int, x: int, cols_c: int, cols_s: int,
                        stream: int) -> None:
        v = ctypes.c_void_p
        _lib.check(_lib.lib().cn_bb_solve_split(self._h, v(re), v(rows_c), v(rows_s), v(x), v(cols_c),
                                                v(cols_s), v(stream)), "BandedBatch.solve_split_ptr")

    def prepare_transpose(self) -> None:
        _lib.check(_lib.lib().cn_bb_prepare_transpose(self._h), "BandedBatch.prepare_transpose")

    def close(self) -> None:
        h, self._h = getattr(self, "_h", None), None
        if h:
            _lib.lib().cn_bb_destroy(h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _device_solve(h: "BatchHandle", b: np.ndarray, transpose: bool) -> np.ndarray:
    """Solve (nrhs, batch, n) host right-hand sides on the device: one upload, one launch over
    all right-hand sides, one download."""
    import torch
    if not torch.cuda.is_available():
        from ..errors import BackendError
        raise BackendError("no HIP device visible", why="the bordered banded solves run on the GPU only",
                           fix="run on an MI355X node")
    nrhs, batch, n = b.shape
    dev = torch.from_numpy(np.ascontiguousarray(b)).to("cuda")
    h.solve_ptr(dev.data_ptr(), nrhs, batch * n, transpose, torch.cuda.current_stream().cuda_stream)
    return dev.cpu().numpy()


class BorderedBanded:
    """One factored system; ``border_rows`` are the rows outside the band (boundary conditions),
    as python/conatus/runtime/banded.py:130-166 declares it."""

    def __init__(self, A, border_rows: Sequence[int], dtype=None, col_border: bool = False):
        A = sp.csr_matrix(A)
        self.dtype = factor_dtype([A], dtype)
        self.border_rows = [int(i) for i in border_rows]
        self.col_border = bool(col_border)
        self._h = BatchHandle([csr_args(A, self.dtype, "BorderedBanded")], self.border_rows, self.dtype,
                              True, self.col_border)
        p = self._h.props()
        self.n, self.kl, self.ku, self.rank = int(p.n), int(p.kl), int(p.ku), int(p.rank)
        self.capacitance_cond = float(self._h.cond()[0])

    def solve(self, b: np.ndarray, transpose: bool = False) -> np.ndarray:
        """x with A x = b (A^T x = b with ``transpose``, no conjugation); b is (n,) or (n, nrhs)."""
        b = solve_operand(b, self.dtype, "BorderedBanded.solve")
        what = f"({self.n},) or ({self.n}, nrhs)"
        check_rhs_shape(b, (self.n,) if b.ndim == 1 else (self.n, None), "BorderedBanded.solve", what)
        cols = np.ascontiguousarray(b.reshape(self.n, -1).T)        # one right-hand side per row
        x = _device_solve(self._h, cols[:, None, :], transpose)[:, 0, :]
        return np.ascontiguousarray(x.T).reshape(b.shape)

    def explain(self) -> str:
        return (f"BorderedBanded(n={self.n}, band kl={self.kl} ku={self.ku}, border rank={self.rank}, "
                f"capacitance cond_1={self.capacitance_cond:.3g}, dtype={np.dtype(self.dtype).name}); "
                "gfx950 banded LU + Woodbury on the boundary rows")


class BorderedBandedBatch:
    """Systems of one size n (one per mode/group), factored once and solved together in one
    device launch (python/conatus/runtime/banded.py:169-213)."""

    def __init__(self, mats: Sequence, border_rows: Sequence[int], dtype=None, col_border: bool = False):
        mats = [sp.csr_matrix(A) for A in mats]
        if not mats:
            raise InputTypeError("BorderedBandedBatch: at least one matrix is needed (one per mode/group)",
                                 fix="pass a non-empty list of matrices")
        self.dtype = factor_dtype(mats, dtype)
        self.col_border = bool(col_border)
        triples = [csr_args(A, self.dtype, f"BorderedBandedBatch (matrix {k})") for k, A in enumerate(mats)]
        self._h = BatchHandle(triples, [int(i) for i in border_rows], self.dtype, True, self.col_border)
        p = self._h.props()
        self.n, self._B, self.kl, self.ku = int(p.n), int(p.batch), int(p.kl), int(p.ku)
        self.threads = solve_threads()

    def __len__(self) -> int:
        return self._B

    def explain(self) -> str:
        kind = "complex" if self.dtype == np.complex128 else "real"
        return (f"BorderedBandedBatch: {len(self)} {kind} bordered-banded systems (one per mode/group), "
                f"factored once on the GPU (boundary rows first, banded LU kl={self.kl} ku={self.ku}, "
                "Woodbury on the border); every solve is one gfx950 launch over the whole batch")

    def solve(self, b: np.ndarray, transpose: bool = False) -> np.ndarray:
        """(batch, n) right-hand sides -> (batch, n) solutions; ``transpose`` solves A_k^T x = b_k."""
        b = solve_operand(b, self.dtype, "BorderedBandedBatch.solve")
        check_rhs_shape(b, (len(self), self.n), "BorderedBandedBatch.solve", f"({len(self)}, {self.n})")
        return _device_solve(self._h, b[None], bool(transpose))[0]
