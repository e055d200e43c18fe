"""Bordered banded solvers on the MI355X: the host-array API of the reference's compiled core.

Same classes, signatures, dtype rules and errors as python/conatus/runtime/banded.py:74-213
(``BorderedBanded``, ``BorderedBandedBatch``, ``csr_args``, ``solve_operand``, ``threads_for``,
``solve_threads``), but every factorization and solve runs in the HIP kernels of libcnhip.so:
host arrays are staged to device memory (torch tensors), solved on the GPU and copied back.
There is no CPU path: without a GPU the constructors raise ``BackendError``.
"""
from __future__ import annotations

import ctypes
import math
import os
from collections.abc import Sequence

import numpy as np
import scipy.sparse as sp

from .. import _lib
from ..errors import DtypeError, InputTypeError, NumericalError, OptionError

__all__ = ["BorderedBanded", "BorderedBandedBatch", "csr_args", "solve_operand", "solve_threads",
           "threads_for", "BatchHandle"]

#: unknowns per thread squared of the reference's CPU thread rule (runtime/banded.py:19-25)
SOLVE_THREAD_UNKNOWNS = 900


def threads_for(batch: int, n: int, threads: int) -> int:
    """min(threads, batch, isqrt(batch n / 900)), or 1 below 2 (runtime/banded.py:28-33).  Kept for
    API parity; the device solve does not use host threads."""
    t = min(threads, batch, math.isqrt(batch * n // SOLVE_THREAD_UNKNOWNS))
    return t if t >= 2 else 1


def solve_threads() -> int:
    """``$CONATUS_SOLVE_THREADS`` / ``$OMP_NUM_THREADS`` / 1, capped by the usable CPUs
    (runtime/banded.py:36-59)."""
    for name in ("CONATUS_SOLVE_THREADS", "OMP_NUM_THREADS"):
        raw = os.environ.get(name, "").strip()
        if raw:
            try:
                t = int(raw)
            except ValueError:
                raise OptionError(f"{name} = {raw!r}: expected a whole number of threads",
                                  fix=f"export {name}=4 (a whole number), or unset it") from None
            if t < 1:
                raise OptionError(f"{name} = {t}: at least one thread", fix=f"export {name}=1")
            break
    else:
        return 1
    try:
        cpus = len(os.sched_getaffinity(0))
    except AttributeError:
        cpus = os.cpu_count() or 1
    return max(1, min(t, cpus))


def _kind(dtype) -> str:
    return "Z" if np.issubdtype(np.dtype(dtype), np.complexfloating) else "D"


def _factor_dtype(mats: Sequence, dtype=None):
    """complex128 if ``dtype`` or any matrix is complex, else float64 (runtime/banded.py:66-71)."""
    if dtype is not None:
        return np.complex128 if _kind(dtype) == "Z" else np.float64
    return np.complex128 if any(_kind(m.dtype) == "Z" for m in mats) else np.float64


def solve_operand(b, dtype, where: str, xp=np):
    """Right-hand side converted to the factor dtype by an exact widening only; a narrowing
    (complex into real factors) is refused (runtime/banded.py:74-89)."""
    b = xp.asarray(b)
    dtype = np.dtype(dtype)
    if b.dtype != dtype:
        if not np.can_cast(b.dtype, dtype, "safe"):
            raise DtypeError(f"{where}: right-hand side of dtype {b.dtype} for {dtype.name} factors",
                             why="the conversion would lose information (e.g. the imaginary part)",
                             fix=f"pass a {dtype.name} right-hand side (take .real explicitly if "
                                 "the imaginary part is known to vanish)")
        b = b.astype(dtype)
    return xp.ascontiguousarray(b)


def csr_args(A, dtype=None, where: str = "csr_args") -> tuple:
    """(n, indptr int32, indices int32, data) of a square, finite CSR matrix; complex entries with
    a non-zero imaginary part refused for real factors (runtime/banded.py:92-117)."""
    A = sp.csr_matrix(A)
    if A.shape[0] != A.shape[1]:
        raise OptionError(f"square matrix required, got {A.shape}", fix="a square matrix")
    A.sum_duplicates()
    A.sort_indices()
    if not np.all(np.isfinite(A.data)):
        k = int(np.flatnonzero(~np.isfinite(A.data))[0])
        row = int(np.searchsorted(A.indptr, k, side="right") - 1)
        raise NumericalError(f"{where}: non-finite matrix entry {A.data[k]!r} at ({row}, {int(A.indices[k])})",
                             why="a factorization of a matrix with NaN/Inf has no meaning",
                             fix="find where the NaN/Inf enters the operator")
    dt = np.complex128 if _kind(dtype or A.dtype) == "Z" else np.float64
    if dt == np.float64 and np.iscomplexobj(A.data) and np.any(A.data.imag != 0):
        raise DtypeError(f"{where}: complex matrix for real (float64) factors",
                         why="the conversion would drop the imaginary part of the entries",
                         fix="pass dtype=np.complex128 (or no dtype: it follows the matrices)")
    return (A.shape[0], np.ascontiguousarray(A.indptr, dtype=np.int32),
            np.ascontiguousarray(A.indices, dtype=np.int32),
            np.ascontiguousarray(A.data.real if dt == np.float64 else A.data, dtype=dt))


def _check_rhs_shape(b, expected: tuple, where: str, what: str) -> None:
    ok = b.ndim == len(expected) and all(e is None or s == e for s, e in zip(b.shape, expected))
    if not ok:
        raise InputTypeError(f"{where}: right-hand side of shape {tuple(b.shape)}, expected {what}",
                             why="any other shape would be reinterpreted, not solved",
                             fix=f"pass an array of shape {what}")


class BatchHandle:
    """Owner of one ``cn_bb`` handle: a factored batch resident on the current HIP device."""

    def __init__(self, csrs: Sequence[tuple], border: Sequence[int], dtype, pivot: bool = True,
                 col_border: bool = False, mode: str | None = None):
        L = _lib.lib()
        if mode not in (None, "lane", "wave"):
            raise OptionError(f"BandedBatch: mode must be None, 'lane' or 'wave', got {mode!r}")
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
        _lib.check(L.cn_bb_request_mode({None: 0, "lane": 1, "wave": 2}[mode]), "BandedBatch")
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


def _torch():
    import torch
    if not torch.cuda.is_available():
        from ..errors import BackendError
        raise BackendError("no HIP device visible", why="the bordered banded solves run on the GPU only",
                           fix="run on an MI355X node")
    return torch


def _device_solve(h: BatchHandle, b: np.ndarray, transpose: bool) -> np.ndarray:
    """Stage (nrhs, batch, n) host data to the device, solve in place, copy back."""
    torch = _torch()
    d = torch.from_numpy(np.ascontiguousarray(b)).to("cuda")
    nrhs = b.shape[0]
    stride = b.shape[1] * b.shape[2]
    stream = torch.cuda.current_stream().cuda_stream
    h.solve_ptr(d.data_ptr(), nrhs, stride, transpose, stream)
    return d.cpu().numpy()


class BorderedBanded:
    """Factor A once; ``border_rows`` are the rows that break the band (runtime/banded.py:130-166)."""

    def __init__(self, A, border_rows: Sequence[int], dtype=None, col_border: bool = False):
        A = sp.csr_matrix(A)
        self.dtype = _factor_dtype([A], dtype)
        self.border_rows = [int(i) for i in border_rows]
        self._h = BatchHandle([csr_args(A, self.dtype, "BorderedBanded")], self.border_rows, self.dtype,
                              True, col_border)
        p = self._h.props()
        self.n, self.kl, self.ku, self.rank = p.n, p.kl, p.ku, p.rank
        self.capacitance_cond = float(self._h.cond()[0])
        self.col_border = bool(col_border)

    def solve(self, b: np.ndarray, transpose: bool = False) -> np.ndarray:
        """Solve A x = b (or A^T x = b). b: (n,) or (n, nrhs)."""
        b = solve_operand(b, self.dtype, "BorderedBanded.solve")
        if b.ndim == 1:
            _check_rhs_shape(b, (self.n,), "BorderedBanded.solve", f"({self.n},) or ({self.n}, nrhs)")
        else:
            _check_rhs_shape(b, (self.n, None), "BorderedBanded.solve", f"({self.n},) or ({self.n}, nrhs)")
        cols = b.reshape(self.n, -1).T.copy()                # (nrhs, n): nrhs blocks of (1, n)
        x = _device_solve(self._h, cols[:, None, :], transpose)[:, 0, :].T
        return x.reshape(b.shape)

    def explain(self) -> str:
        return (f"BorderedBanded(n={self.n}, band kl={self.kl} ku={self.ku}, border rank={self.rank}, "
                f"capacitance cond_1={self.capacitance_cond:.3g}, dtype={np.dtype(self.dtype).name}); "
                "gfx950 banded LU + Woodbury on the boundary rows")


class BorderedBandedBatch:
    """One factorization per mode/group (same n), all solved in one device launch
    (runtime/banded.py:169-213)."""

    def __init__(self, mats: Sequence, border_rows: Sequence[int], dtype=None, col_border: bool = False):
        mats = [sp.csr_matrix(A) for A in mats]
        if not mats:
            raise InputTypeError("BorderedBandedBatch: no matrix given",
                                 fix="pass at least one matrix (one per mode/group)")
        self.dtype = _factor_dtype(mats, dtype)
        self.col_border = bool(col_border)
        self._h = BatchHandle([csr_args(A, self.dtype, f"BorderedBandedBatch (matrix {k})")
                               for k, A in enumerate(mats)], [int(i) for i in border_rows], self.dtype,
                              True, col_border)
        self.n = mats[0].shape[0]
        self.threads = solve_threads()
        self._B = len(mats)
        p = self._h.props()
        self.kl, self.ku = p.kl, p.ku

    def __len__(self) -> int:
        return self._B

    def explain(self) -> str:
        kind = "complex" if self.dtype == np.complex128 else "real"
        return (f"BorderedBandedBatch: {len(self)} {kind} bordered-banded systems (one per mode/group), "
                f"each factorized once on the GPU: boundary rows first, banded LU (kl={self.kl}, "
                f"ku={self.ku}), Woodbury on the border; all solves in one gfx950 launch")

    def solve(self, b: np.ndarray, transpose: bool = False) -> np.ndarray:
        """b: (batch, n) -> x: (batch, n); ``transpose=True`` solves A_k^T x = b_k."""
        b = solve_operand(b, self.dtype, "BorderedBandedBatch.solve")
        _check_rhs_shape(b, (len(self), self.n), "BorderedBandedBatch.solve", f"({len(self)}, {self.n})")
        return _device_solve(self._h, b[None], bool(transpose))[0]
