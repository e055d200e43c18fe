"""MI355X device backend: batched bordered banded solves on caller-owned device memory.

The device-side twin of python/conatus/runtime/cuda.py:127-264 (``cuda.BandedBatch``): same
constructor arguments, properties and solve modes (in place, fused gather/scatter maps, split
complex), with torch tensors as the device arrays and the current torch stream as the launch
stream.  Solves allocate nothing once the work space is reserved and never synchronise the host,
so they can be captured in a HIP graph.
"""
from __future__ import annotations

from collections.abc import Sequence

import numpy as np
import scipy.sparse as sp

from ..errors import BackendError, OptionError
from .banded import SOLVE_MODES, BatchHandle, _factor_dtype, csr_args
from .checks import check_index_map, check_io_overlap

__all__ = ["available", "device_name", "synchronize", "to_device", "BandedBatch"]


def _torch():
    import torch
    return torch


def available() -> bool:
    """True when a HIP device is visible and libcnhip.so loads."""
    try:
        from .. import _lib
        return _lib.device_count() > 0
    except Exception:
        return False


def _need():
    torch = _torch()
    if not torch.cuda.is_available():
        raise BackendError("HIP backend not available: no GPU visible to this process",
                           fix="run on an MI355X node")
    return torch


def device_name() -> str:
    torch = _torch()
    return torch.cuda.get_device_name(0) if torch.cuda.is_available() else "none"


def synchronize() -> None:
    _need().cuda.synchronize()


def to_device(a: np.ndarray):
    """Explicit host -> device copy of a 2D (batch, n) array."""
    torch = _need()
    a = np.asarray(a)
    if a.ndim != 2:
        raise ValueError(f"to_device expects a 2D (batch, n) array, got shape {a.shape}")
    dt = np.complex128 if np.iscomplexobj(a) else np.float64
    return torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to("cuda")


def _stream() -> int:
    return _torch().cuda.current_stream().cuda_stream


_TORCH_DT = {np.dtype(np.float64): "float64", np.dtype(np.complex128): "complex128"}


class BandedBatch:
    """One bordered banded system per mode (same n and border rows), factored and solved on the GPU.

    ``pivot=False`` skips row interchanges (declared diagonally dominant).  ``level_schedule`` and
    ``level_merge`` are accepted for signature parity with the reference and ignored: the gfx950
    solve streams the interleaved band factors directly.
    """

    def __init__(self, mats: Sequence, border_rows: Sequence[int], pivot: bool = True, dtype=None,
                 level_schedule: bool = True, level_merge: bool = True, col_border: bool = False,
                 mode: str | None = None):
        _need()
        mats = [sp.csr_matrix(A) for A in mats]
        self.dtype = _factor_dtype(mats, dtype)
        self.border_rows = [int(i) for i in border_rows]
        self._h = BatchHandle([csr_args(A, self.dtype) for A in mats], self.border_rows, self.dtype,
                              bool(pivot), bool(col_border), mode=mode)
        p = self._h.props()
        self.n, self.kl, self.ku, self.rank, self.batch = p.n, p.kl, p.ku, p.rank, p.batch
        self.pivot = bool(p.pivot)
        self.factor_ms = float(p.factor_ms)
        self.capacitance_cond_max = float(p.cond_max)
        self.col_border = bool(col_border)
        self.kernel = (p.kl_kernel, p.wu_kernel, p.lanes)
        self.solve_mode = SOLVE_MODES[p.solve_mode]
        self.factor_bytes = int(p.factor_bytes)

    level_scheduled = False
    level_merged = False
    level_note = "interleaved band factors (gfx950 kernel); no level schedule"

    def _check_tensor(self, d, name: str, shape: tuple | None = None):
        torch = _torch()
        if not isinstance(d, torch.Tensor) or d.device.type != "cuda":
            raise ValueError(f"BandedBatch: {name} must be a device tensor")
        if str(d.dtype).replace("torch.", "") != _TORCH_DT[np.dtype(self.dtype)]:
            raise ValueError(f"BandedBatch: {name} must be {np.dtype(self.dtype).name}, got {d.dtype}")
        if not d.is_contiguous():
            raise ValueError(f"BandedBatch: {name} must be contiguous")
        if shape is not None and tuple(d.shape) != shape:
            raise ValueError(f"BandedBatch: expected a {shape} tensor for {name}, got {tuple(d.shape)}")

    def reserve(self, nrhs: int) -> None:
        """Size the solve work space for ``nrhs`` right-hand sides (before graph capture)."""
        self._h.reserve(nrhs)

    def refactor(self, check: bool = True) -> None:
        """Run the device factorization again from the stored layout (asynchronous if not check)."""
        self._h.factor(_stream(), check)
        if check:
            self.factor_ms = float(self._h.props().factor_ms)

    def solve(self, d) -> None:
        """Solve in place; ``d`` is a (batch, n) device tensor of the factor dtype."""
        self._check_tensor(d, "d", (self.batch, self.n))
        self._h.solve_ptr(d.data_ptr(), 1, self.batch * self.n, False, _stream())

    def solve_many(self, d) -> None:
        """Solve in place for nrhs right-hand sides per system: ``d`` is (nrhs, batch, n); each system's
        factors are read once for all of its right-hand sides."""
        if d.dim() != 3:
            raise ValueError("BandedBatch.solve_many: expected a (nrhs, batch, n) tensor")
        self._check_tensor(d, "d", (d.shape[0], self.batch, self.n))
        self._h.solve_ptr(d.data_ptr(), d.shape[0], self.batch * self.n, False, _stream())

    def solve_ptr(self, ptr: int, batch: int, n: int, stream: int = 0) -> None:
        """In place on caller-owned device memory (C-contiguous (batch, n))."""
        if batch != self.batch or n != self.n:
            raise ValueError("BandedBatch.solve_ptr: expected (batch, n) matching the factorization")
        if not ptr:
            raise ValueError("BandedBatch.solve_ptr: null device pointer")
        self._h.solve_ptr(ptr, 1, self.batch * self.n, False, stream)

    def solve_transpose(self, d) -> None:
        """Solve A_k^T x = b_k in place (plain transpose, same factors)."""
        self._check_tensor(d, "d", (self.batch, self.n))
        self._h.solve_ptr(d.data_ptr(), 1, self.batch * self.n, True, _stream())

    def prepare_transpose(self) -> None:
        self._h.prepare_transpose()

    def solve_mapped(self, rhs, rows, x, cols) -> None:
        """Out-of-place solve with the gather/scatter fused in the kernel: b_m[i] = rhs[rows[m, i]],
        x[cols[m, i]] = x_m[i]; ``rows``/``cols`` int64 (batch, n) device tensors or None (the
        contiguous layout).  The kernel reads raw pointers, so the operands, the maps' ranges, a
        repeated scatter index and overlapping rhs/x are refused here (the refusals of
        python/conatus/runtime/cuda.py:186-232)."""
        where = "BandedBatch.solve_mapped"
        self._check_tensor(rhs, "rhs")
        self._check_tensor(x, "x")
        check_index_map("rows", rows, rhs.numel(), self.batch, self.n, where)
        check_index_map("cols", cols, x.numel(), self.batch, self.n, where, unique=True)
        check_io_overlap(rhs, rows, x, cols, where)
        self._h.solve_mapped_ptr(rhs.data_ptr(), 0 if rows is None else rows.data_ptr(), x.data_ptr(),
                                 0 if cols is None else cols.data_ptr(), _stream())

    def solve_mapped_ptr(self, rhs: int, rows: int, x: int, cols: int, stream: int = 0) -> None:
        if not rhs or not x:
            raise ValueError("BandedBatch.solve_mapped_ptr: null device pointer")
        self._h.solve_mapped_ptr(rhs, rows, x, cols, stream)

    def solve_split(self, re, rows_c, rows_s, x, cols_c, cols_s) -> None:
        """Complex systems of real data: b = re[rows_c] + i re[rows_s], x[cols_c] = Re, x[cols_s] = Im
        (four int64 (batch, n) maps)."""
        torch = _torch()
        where = "BandedBatch.solve_split"
        if self.dtype != np.complex128:
            raise OptionError(f"{where}: complex128 factors required")
        for name, a in (("re", re), ("x", x)):
            if not (isinstance(a, torch.Tensor) and a.device.type == "cuda" and a.dtype == torch.float64
                    and a.is_contiguous()):
                raise ValueError(f"{where}: {name} must be a contiguous float64 device tensor")
        for name, idx, target in (("rows_c", rows_c, re), ("rows_s", rows_s, re), ("cols_c", cols_c, x),
                                  ("cols_s", cols_s, x)):
            if idx is None:
                raise ValueError(f"{where}: {name} is required (an int64 ({self.batch}, {self.n}) tensor)")
            check_index_map(name, idx, target.numel(), self.batch, self.n, where)
        self._h.solve_split_ptr(re.data_ptr(), rows_c.data_ptr(), rows_s.data_ptr(), x.data_ptr(),
                                cols_c.data_ptr(), cols_s.data_ptr(), _stream())

    def solve_split_ptr(self, re: int, rows_c: int, rows_s: int, x: int, cols_c: int, cols_s: int,
                        stream: int = 0) -> None:
        self._h.solve_split_ptr(re, rows_c, rows_s, x, cols_c, cols_s, stream)

    def explain(self) -> str:
        kl, wu, lanes = self.kernel
        return (f"hip.BandedBatch(batch={self.batch}, n={self.n}, common band kl={self.kl} ku={self.ku}, "
                f"border rank={self.rank}, pivot={'partial (in band)' if self.pivot else 'none (declared)'}, "
                f"capacitance cond_1 max={self.capacitance_cond_max:.3g}, factor {self.factor_ms:.3g} ms); "
                f"gfx950: factor {lanes} lanes per system (kernel kl={kl}), "
                + ("solve one lane per (system, rhs) over 16-system interleaved factors (window "
                   f"{wu}), streamed through LDS" if self.solve_mode == "lane" else
                   "solve one wavefront per (system, rhs), band window across lanes, U stored by columns"
                   if self.solve_mode == "wave" else
                   "solve one workgroup per (system, rhs), its sweeps partitioned over 2-8 wavefronts "
                   "coupled through factor-time spikes, band window across lanes"))
