"""Input checks of the banded-solver boundary.

Every operand that crosses into libcnhip.so is described here once: matrices become canonical
CSR triples of the factor dtype, right-hand sides are converted only when no information is lost,
device index maps are validated against the arrays they index, and the CPU thread count of the
host API is read from the environment.  The kernels read raw pointers, so anything they cannot
interpret safely is refused on the host with the exception type the reference's Python layer
raises (python/conatus/runtime/banded.py:28-127, python/conatus/runtime/cuda.py:186-232): the
names, signatures, exception classes and the words tests match on are that interface.
"""
from __future__ import annotations

import math
import os
from collections.abc import Iterable

import numpy as np
import scipy.sparse as sp

from ..errors import DtypeError, InputTypeError, NumericalError, OptionError

__all__ = ["SOLVE_THREAD_UNKNOWNS", "threads_for", "solve_threads", "factor_dtype", "solve_operand",
           "csr_args", "check_rhs_shape", "check_index_map", "check_io_overlap"]

#: A batched CPU solve pays ~30 ns per unknown and ~27 us per thread it starts, so t threads are
#: worth it near t = sqrt(B n / 900); the constant is the reference's measured one.
SOLVE_THREAD_UNKNOWNS = 900

_THREAD_VARS = ("CONATUS_SOLVE_THREADS", "OMP_NUM_THREADS")


def threads_for(batch: int, n: int, threads: int) -> int:
    """Threads worth starting for one batched solve of ``batch`` systems of size ``n`` with
    ``threads`` allowed: the square-root rule capped by ``threads`` and ``batch``; a single
    thread whenever the rule gives fewer than two."""
    best = math.isqrt((batch * n) // SOLVE_THREAD_UNKNOWNS)
    t = min(best, batch, threads)
    return 1 if t < 2 else t


def _env_threads(var: str) -> int | None:
    text = os.environ.get(var, "").strip()
    if not text:
        return None
    if not text.lstrip("+-").isdigit():
        raise OptionError(f"{var} must name a number of threads, it holds {text!r}",
                          fix=f"export {var}=4, or unset {var}")
    value = int(text)
    if value < 1:
        raise OptionError(f"{var} = {value} asks for no thread at all", fix=f"export {var}=1 (or more)")
    return value


def solve_threads() -> int:
    """CPU threads the host API may use: the first of $CONATUS_SOLVE_THREADS, $OMP_NUM_THREADS that
    is set, never more than the CPUs this process may run on; 1 when neither is set.  The device
    solve itself uses no host threads; the value is kept for the reference's explain() text."""
    wanted = next((t for t in map(_env_threads, _THREAD_VARS) if t is not None), None)
    if wanted is None:
        return 1
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return max(1, min(wanted, usable))


def _is_complex(dt) -> bool:
    return np.dtype(dt).kind == "c"


def factor_dtype(mats: Iterable, dtype=None) -> type:
    """The factor type of a batch: complex128 as soon as ``dtype`` says complex or, without
    ``dtype``, as soon as one matrix of the batch is complex; float64 otherwise."""
    if dtype is not None:
        return np.complex128 if _is_complex(dtype) else np.float64
    return np.complex128 if any(_is_complex(m.dtype) for m in mats) else np.float64


def solve_operand(b, dtype, where: str, xp=np):
    """``b`` as a contiguous array of the factor dtype.  Only a lossless conversion is made
    (float32 to float64, real to complex); complex data for real factors would lose its imaginary
    part and is refused, since the kernels would read interleaved complex memory as reals."""
    b = xp.asarray(b)
    want = np.dtype(dtype)
    if b.dtype == want:
        return xp.ascontiguousarray(b)
    if np.promote_types(b.dtype, want) != want:
        raise DtypeError(f"{where}: cannot solve a {b.dtype} right-hand side with {want.name} factors",
                         why="converting it would drop information such as the imaginary part",
                         fix=f"pass {want.name} data; take .real yourself if the imaginary part is "
                             "known to be zero")
    return xp.ascontiguousarray(b.astype(want))


def csr_args(A, dtype=None, where: str = "csr_args") -> tuple:
    """(n, indptr, indices, data) of ``A`` for the library: square, canonical (duplicates summed,
    columns sorted), int32 indices, every entry finite, data in the factor dtype.  A complex
    matrix is accepted for real factors only when all its imaginary parts are zero."""
    A = sp.csr_matrix(A, copy=True)
    rows, cols = A.shape
    if rows != cols:
        raise OptionError(f"{where}: the operator is {rows} x {cols}, a square matrix is required",
                          fix="pass the n x n matrix of one mode/group")
    A.sum_duplicates()            # also sorts the column indices of every row
    finite = np.isfinite(A.data)
    if not finite.all():
        k = int(np.argmin(finite))
        row_of = np.repeat(np.arange(rows), np.diff(A.indptr))
        raise NumericalError(f"{where}: non-finite entry {A.data[k]!r} in row {int(row_of[k])}, "
                             f"column {int(A.indices[k])}",
                             why="a matrix holding NaN or Inf cannot be factored meaningfully",
                             fix="trace the NaN/Inf back to the coefficient or parameter it comes from")
    target = factor_dtype([A], dtype if dtype is not None else A.dtype)
    data = A.data
    if target is np.float64 and _is_complex(data.dtype):
        if np.any(data.imag):
            raise DtypeError(f"{where}: a complex operator cannot be factored as float64",
                             why="its imaginary part would be discarded",
                             fix="factor it as complex128 (dtype=np.complex128, or no dtype at all)")
        data = data.real
    return (rows, np.ascontiguousarray(A.indptr, np.int32), np.ascontiguousarray(A.indices, np.int32),
            np.ascontiguousarray(data, target))


def check_rhs_shape(b, expected: tuple, where: str, what: str) -> None:
    """Refuse a right-hand side whose shape differs from ``expected`` (None: any extent there);
    a reshape would silently reinterpret it as another problem."""
    shape = tuple(b.shape)
    fits = len(shape) == len(expected) and all(e is None or e == s for s, e in zip(shape, expected))
    if not fits:
        raise InputTypeError(f"{where}: the right-hand side has shape {shape}, {what} is needed",
                             why="another shape would be reinterpreted rather than solved",
                             fix=f"reshape it to {what} explicitly")


def check_index_map(name: str, idx, length: int, batch: int, n: int, where: str, unique: bool = False) -> None:
    """A device gather/scatter map of a fused solve: an int64 (batch, n) contiguous tensor whose
    entries index an array of ``length`` entries; ``unique`` for a scatter map (two systems
    writing one entry would race).  ``idx`` None means the contiguous (batch, n) layout, which
    needs ``length`` >= batch n."""
    import torch
    if idx is None:
        if length < batch * n:
            raise ValueError(f"{where}: {name}=None (contiguous layout) needs {batch * n} entries, the "
                             f"array has {length}")
        return
    if idx.dtype != torch.int64 or tuple(idx.shape) != (batch, n) or not idx.is_contiguous():
        raise ValueError(f"{where}: {name} must be a contiguous int64 tensor of shape ({batch}, {n}); "
                         f"got {idx.dtype} of shape {tuple(idx.shape)}")
    if idx.numel() == 0:
        return
    lo, hi = torch.aminmax(idx)
    lo, hi = int(lo), int(hi)
    if lo < 0 or hi >= length:
        raise ValueError(f"{where}: {name} has indices out of range: [{lo}, {hi}] against {length} entries")
    if unique and torch.sort(idx.reshape(-1)).values.diff().eq(0).any():
        raise ValueError(f"{where}: {name} repeats an index, so two systems would write one entry")


def check_io_overlap(rhs, rows, x, cols, where: str) -> None:
    """rhs and x may share memory only as one array solved in place with one map for both: any
    other overlap lets a system read entries another one has already overwritten."""
    import torch
    a0, a1 = rhs.data_ptr(), rhs.data_ptr() + rhs.numel() * rhs.element_size()
    b0, b1 = x.data_ptr(), x.data_ptr() + x.numel() * x.element_size()
    if a1 <= b0 or b1 <= a0:
        return
    same_span = (a0, a1) == (b0, b1)
    same_map = (rows is None and cols is None) or (
        rows is not None and cols is not None and bool(torch.equal(rows, cols)))
    if not (same_span and same_map):
        raise ValueError(f"{where}: rhs and x overlap in memory with different maps, so systems would read "
                         "entries already overwritten; pass a separate x, or one map for both rows and cols")
