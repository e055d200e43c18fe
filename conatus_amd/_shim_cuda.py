"""The ``cuda`` submodule of the ``conatus._core`` stand-in (core_shim): the names and signatures
of the reference's device bindings (core/bindings/module.cpp:261-443) on libcnhip.so, with torch
tensors as the device memory.

    available() device_name() synchronize() pool_stats() pool_trim() plan_cache_size()
    DeviceArray{D,Z}.from_host(a) / .zeros(rows, cols) / .to_host() / .copy() / .shape / .ptr
    cheb_forward(grid, N) cheb_backward(coef, M) (DeviceArrayD or Z),
    realfourier_forward(grid, N) realfourier_backward(coef, M) (DeviceArrayD),
    complexfourier_forward(grid, N) complexfourier_backward(coef, M) (DeviceArrayZ)
    BandedBatch{D,Z}(mats, border_rows, pivot=True, level_schedule=True, level_merge=True,
                     col_border=False)
        .solve(DeviceArray) .solve_transpose(DeviceArray) .prepare_transpose()
        .solve_ptr(ptr, batch, n, stream=0) .solve_transpose_ptr(ptr, batch, n, stream=0)
        .solve_mapped_ptr(rhs, rows, x, cols, stream=0)
        .solve_split_ptr(re, rows_c, rows_s, x, cols_c, cols_s, stream=0)
        .n .kl .ku .rank .batch .pivot .factor_ms .capacitance_cond_max .col_border ...
    Imex{D,Z}(M, L, border_rows, scheme, dt, pivot=True)
        .set_pointwise_rhs(g, h) .run(DeviceArray, t0, nsteps) -> t .reset()
        .history_length .n_factorizations .iteration .scheme

Pointer entry points take device addresses as integers (a torch tensor's ``data_ptr()``) and a
stream handle (``torch.cuda.current_stream().cuda_stream``; 0 is the null stream), exactly as the
reference's take CuPy pointers; they allocate nothing once the work space is reserved (the
constructor reserves one right-hand side), so they can be captured in a HIP graph.  The
level-scheduled solve of the reference is an experimental CUDA latency trick this library does
not have: its flags are accepted and its properties report that no level schedule exists.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from .runtime.banded import BatchHandle

__all__ = ["available", "device_name", "synchronize", "pool_stats", "pool_trim", "plan_cache_size",
           "DeviceArrayD", "DeviceArrayZ", "BandedBatchD", "BandedBatchZ", "ImexD", "ImexZ",
           "cheb_forward", "cheb_backward", "realfourier_forward", "realfourier_backward",
           "complexfourier_forward", "complexfourier_backward"]


def _torch():
    import torch
    return torch


def available() -> bool:
    try:
        from . import _lib
        return _lib.device_count() > 0 and _torch().cuda.is_available()
    except Exception:
        return False


def device_name() -> str:
    t = _torch()
    return t.cuda.get_device_name(0) if t.cuda.is_available() else "none"


def synchronize() -> None:
    _torch().cuda.synchronize()


def pool_stats() -> tuple:
    """(live bytes, cached bytes, allocations, reuses) of the device memory pool (torch's)."""
    st = _torch().cuda.memory_stats()
    live = int(st.get("allocated_bytes.all.current", 0))
    cached = int(st.get("reserved_bytes.all.current", 0)) - live
    n_alloc = int(st.get("segment.all.allocated", 0))
    n_reuse = max(0, int(st.get("allocation.all.allocated", 0)) - n_alloc)
    return live, cached, n_alloc, n_reuse


def pool_trim() -> None:
    _torch().cuda.empty_cache()


def plan_cache_size() -> int:
    """Transform plans cached by the backend: none (the library's transforms need no plans)."""
    return 0


class _DeviceArray:
    _dtype = np.float64

    def __init__(self, tensor):
        self._t = tensor

    @classmethod
    def from_host(cls, array):
        a = np.asarray(array)
        if a.ndim != 2:
            raise ValueError(f"DeviceArray.from_host: a 2D array is needed, got shape {a.shape}")
        return cls(_torch().from_numpy(np.ascontiguousarray(a, dtype=cls._dtype)).to("cuda"))

    @classmethod
    def zeros(cls, rows, cols):
        torch = _torch()
        dt = torch.complex128 if cls._dtype == np.complex128 else torch.float64
        return cls(torch.zeros((int(rows), int(cols)), dtype=dt, device="cuda"))

    def to_host(self) -> np.ndarray:
        return self._t.cpu().numpy()

    def copy(self):
        return type(self)(self._t.clone())

    @property
    def shape(self) -> tuple:
        return tuple(self._t.shape)

    @property
    def ptr(self) -> int:
        return int(self._t.data_ptr())


class DeviceArrayD(_DeviceArray):
    _dtype = np.float64


class DeviceArrayZ(_DeviceArray):
    _dtype = np.complex128


# ---- transforms (core/bindings/module.cpp:297-311, 476-496): rows of a DeviceArray along the last
# axis, the grid size taken from the input (forward) or given (backward); a new DeviceArray returned
def _transform(a, kind: str, direction: str, size: int, kinds: tuple):
    from . import transforms as TR
    if not isinstance(a, kinds):
        raise TypeError(f"{kind} transforms take {' or '.join(k.__name__ for k in kinds)}, got {type(a).__name__}")
    size = int(size)
    rows, n = a.shape
    M, N = (n, size) if direction == "forward" else (size, n)
    if N < 1 or N > M:
        raise ValueError(f"{kind} {direction}: {N} coefficients and {M} grid points (need 1 <= N <= M)")
    if kind in ("rf", "cf") and N % 2:
        raise ValueError(f"{kind} {direction}: Fourier sizes must be even (Nyquist convention)")
    Ar, Ai = TR._maps(kind, direction, M, N, a._t.device)
    return type(a)(TR._apply(a._t, Ar, Ai, size))


def cheb_forward(grid, N):
    return _transform(grid, "cheb", "forward", N, (DeviceArrayD, DeviceArrayZ))


def cheb_backward(coef, M):
    return _transform(coef, "cheb", "backward", M, (DeviceArrayD, DeviceArrayZ))


def realfourier_forward(grid, N):
    return _transform(grid, "rf", "forward", N, (DeviceArrayD,))


def realfourier_backward(coef, M):
    return _transform(coef, "rf", "backward", M, (DeviceArrayD,))


def complexfourier_forward(grid, N):
    return _transform(grid, "cf", "forward", N, (DeviceArrayZ,))


def complexfourier_backward(coef, M):
    return _transform(coef, "cf", "backward", M, (DeviceArrayZ,))


def _triples(mats, dtype):
    from .core_shim import csr_triple
    return [csr_triple(n, ip, ix, d, dtype) for (n, ip, ix, d) in mats]


def _stream(stream) -> int:
    return int(stream or 0)


class _BandedBatch:
    _dtype = np.float64

    def __init__(self, mats, border_rows, pivot=True, level_schedule=True, level_merge=True, col_border=False):
        self._h = BatchHandle(_triples(mats, self._dtype), [int(i) for i in border_rows], self._dtype,
                              bool(pivot), bool(col_border))
        p = self._h.props()
        self.n, self.kl, self.ku, self.rank, self.batch = int(p.n), int(p.kl), int(p.ku), int(p.rank), int(p.batch)
        self.pivot = bool(p.pivot)
        self.factor_ms = float(p.factor_ms)
        self.capacitance_cond_max = float(p.cond_max)
        self.col_border = bool(col_border)
        self.col_border_nnz = 0

    level_scheduled = False
    level_merged = False
    level_depth_forward = level_depth_backward = 0
    level_steps_forward = level_steps_backward = 0
    level_stream_bytes = 0.0
    level_note = "no level schedule: gfx950 band kernels (one lane or one wavefront per system)"
    level_transposed = False
    band_released = False
    level_t_stream_bytes = 0.0
    level_t_note = "no level schedule"
    level_t_steps = (0, 0)

    def _shape_ok(self, batch, n, what):
        if int(batch) != self.batch or int(n) != self.n:
            raise ValueError(f"BandedBatch.{what}: expected (batch, n) matching the factorization")

    def solve(self, b) -> None:
        if b.shape != (self.batch, self.n):
            raise ValueError("BandedBatch.solve: expected a (batch, n) DeviceArray")
        self._h.solve_ptr(b.ptr, 1, self.batch * self.n, False, _torch().cuda.current_stream().cuda_stream)

    def solve_transpose(self, b) -> None:
        if b.shape != (self.batch, self.n):
            raise ValueError("BandedBatch.solve_transpose: expected a (batch, n) DeviceArray")
        self._h.solve_ptr(b.ptr, 1, self.batch * self.n, True, _torch().cuda.current_stream().cuda_stream)

    def prepare_transpose(self) -> None:
        self._h.prepare_transpose()

    def solve_ptr(self, ptr, batch, n, stream=0) -> None:
        self._shape_ok(batch, n, "solve_ptr")
        if not ptr:
            raise ValueError("BandedBatch.solve_ptr: null device pointer")
        self._h.solve_ptr(int(ptr), 1, self.batch * self.n, False, _stream(stream))

    def solve_transpose_ptr(self, ptr, batch, n, stream=0) -> None:
        self._shape_ok(batch, n, "solve_transpose_ptr")
        if not ptr:
            raise ValueError("BandedBatch.solve_transpose_ptr: null device pointer")
        self._h.solve_ptr(int(ptr), 1, self.batch * self.n, True, _stream(stream))

    def solve_mapped_ptr(self, rhs, rows, x, cols, stream=0) -> None:
        if not rhs or not x:
            raise ValueError("BandedBatch.solve_mapped_ptr: null device pointer")
        self._h.solve_mapped_ptr(int(rhs), int(rows or 0), int(x), int(cols or 0), _stream(stream))

    def solve_split_ptr(self, re, rows_c, rows_s, x, cols_c, cols_s, stream=0) -> None:
        self._h.solve_split_ptr(int(re), int(rows_c), int(rows_s), int(x), int(cols_c), int(cols_s),
                                _stream(stream))


class BandedBatchD(_BandedBatch):
    _dtype = np.float64


class BandedBatchZ(_BandedBatch):
    _dtype = np.complex128


class _DeviceImex:
    """The reference's cuda.Imex binding: pointwise F = g * X + h (or 0), SBDF / RK schemes."""
    _dtype = np.float64

    def __init__(self, M, L, border_rows, scheme, dt, pivot=True):
        from .runtime.imex import Imex
        Mt, Lt = _triples(M, self._dtype), _triples(L, self._dtype)
        mats = lambda ts: [sp.csr_matrix((d, ix, ip), shape=(n, n)) for (n, ip, ix, d) in ts]   # noqa: E731
        self._im = Imex(mats(Mt), mats(Lt), [int(b) for b in border_rows], str(scheme), float(dt),
                        pivot=bool(pivot), dtype=self._dtype)

    def set_pointwise_rhs(self, g, h) -> None:
        """F = g * X + h with (batch, n) host arrays (None: absent), copied to the device once."""
        N = self._im.batch * self._im.n
        for a in (g, h):
            if a is not None and np.asarray(a).size != N:
                raise ValueError("pointwise rhs: expected batch*n values")
        self._im.set_pointwise(g, h)

    def run(self, x, t0, nsteps) -> float:
        if x.shape != (self._im.batch, self._im.n):
            raise ValueError("Imex.run: expected a (batch, n) DeviceArray")
        return self._im.run(x._t, float(t0), int(nsteps))

    def reset(self) -> None:
        self._im.reset()

    history_length = property(lambda self: self._im.history_length)
    n_factorizations = property(lambda self: self._im.n_factorizations)
    iteration = property(lambda self: self._im.iteration)
    scheme = property(lambda self: self._im.scheme)


class ImexD(_DeviceImex):
    _dtype = np.float64


class ImexZ(_DeviceImex):
    _dtype = np.complex128
