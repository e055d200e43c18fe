"""Batched grid <-> coefficient transforms on the device, along the last axis of (batch, n) rows:
the reference's ``runtime.cuda.forward`` / ``backward`` (python/conatus/runtime/cuda.py:93-125),
with torch tensors as the device arrays.

    forward(d, basis, dealias=True)   (batch, M) grid values   -> (batch, N) coefficients
    backward(d, basis, dealias=True)  (batch, N) coefficients  -> (batch, M) grid values

Bases carry the reference's attributes (``size`` N, ``bounds``, ``dealias``; ``grid_size`` M =
floor(N dealias)) and conventions (python/conatus/reference/transforms.py:1-20):
* ``Chebyshev``: Gauss-Chebyshev nodes in ascending order, coefficients of T_n, n < N;
* ``RealFourier``: real data, coefficients as pairs [cos_0, msin_0, cos_1, msin_1, ...] (N even,
  msin_0 = 0, no Nyquist mode);
* ``ComplexFourier``: complex data, coefficients in fftfreq order, the Nyquist entry zero.
Chebyshev rows may be real or complex (the same real map on both parts).

Each transform is one launch of ``cn_tr_rows`` (libcnhip): the product of the rows with the
transform's matrix, built once per (basis kind, direction, M, N) on the host in float64 and kept
on the device.  The first call for a size uploads its matrix (synchronously); later calls only
launch, so they can be captured in a HIP graph.
"""
from __future__ import annotations

import threading
from dataclasses import dataclass

import numpy as np

from . import _lib
from . import spectral as S
from .errors import OptionError

__all__ = ["Chebyshev", "RealFourier", "ComplexFourier", "grid_size", "coeff_size", "grid", "forward",
           "backward"]


def _torch():
    import torch
    return torch


@dataclass(frozen=True)
class _Basis:
    size: int
    bounds: tuple = (-1.0, 1.0)
    dealias: float = 1.5

    def __post_init__(self):
        if int(self.size) != self.size or self.size < 1:
            raise ValueError(f"{type(self).__name__}: size must be a whole number >= 1, got {self.size!r}")
        if not self.dealias >= 1.0:
            raise ValueError(f"{type(self).__name__}: dealias must be >= 1, got {self.dealias!r}")

    @property
    def grid_size(self) -> int:
        return S.dealias_size(int(self.size), self.dealias)

    @property
    def length(self) -> float:
        return float(self.bounds[1] - self.bounds[0])


class Chebyshev(_Basis):
    """Chebyshev basis on [a, b] (Gauss-Chebyshev grid)."""


@dataclass(frozen=True)
class RealFourier(_Basis):
    """Real Fourier basis of period b - a (pairs [cos_k, msin_k]); size even."""
    bounds: tuple = (0.0, 2 * np.pi)


@dataclass(frozen=True)
class ComplexFourier(_Basis):
    """Complex Fourier basis of period b - a (fftfreq order); size even."""
    bounds: tuple = (0.0, 2 * np.pi)


def grid_size(b: _Basis, dealias: bool = True) -> int:
    """Number of grid points of ``b`` (dealiased or not)."""
    return b.grid_size if dealias else int(b.size)


def coeff_size(b: _Basis) -> int:
    return int(b.size)


def grid(b: _Basis, dealias: bool = True) -> np.ndarray:
    """1D grid points of ``b`` (reference/transforms.py:86-101)."""
    M = grid_size(b, dealias)
    a = float(b.bounds[0])
    if isinstance(b, Chebyshev):
        return S.cheb_grid(M, a, float(b.bounds[1]))
    if isinstance(b, (RealFourier, ComplexFourier)):
        return a + b.length * np.arange(M) / M
    raise NotImplementedError(f"transforms: {type(b).__name__} not supported")


def _check_even(b: _Basis) -> None:
    if isinstance(b, (RealFourier, ComplexFourier)) and b.size % 2:
        raise ValueError(f"{b!r}: Fourier sizes must be even (Nyquist convention)")


_cache: dict = {}
_cache_lock = threading.Lock()


def _maps(kind: str, direction: str, M: int, N: int, device):
    """(Ar, Ai or None) of one transform on ``device``: (nout, nin) float64 row-major tensors."""
    key = (kind, direction, M, N, str(device))
    got = _cache.get(key)
    if got is not None:
        return got
    if kind == "cheb":
        A = S.cheb_forward_matrix(N, M) if direction == "forward" else S.cheb_backward_matrix(N, M)
    elif kind == "rf":
        A = S.realfourier_forward_matrix(N, M) if direction == "forward" else S.realfourier_backward_matrix(N, M)
    else:
        A = S.complexfourier_forward_matrix(N, M) if direction == "forward" else S.complexfourier_backward_matrix(N, M)
    torch = _torch()
    Ar = torch.from_numpy(np.ascontiguousarray(A.real, dtype=np.float64)).to(device)
    Ai = torch.from_numpy(np.ascontiguousarray(A.imag, dtype=np.float64)).to(device) if np.iscomplexobj(A) else None
    with _cache_lock:
        _cache.setdefault(key, (Ar, Ai))
    return _cache[key]


def _check_rows(d, n: int, what: str) -> None:
    torch = _torch()
    if not isinstance(d, torch.Tensor) or not d.is_cuda:
        raise TypeError("transforms take device tensors (torch, on the GPU)")
    if d.ndim != 2:
        raise ValueError(f"expected a 2D (batch, n) array, got shape {tuple(d.shape)}")
    if d.shape[1] != n:
        raise ValueError(f"expected {n} {what} per row, got {d.shape[1]}")
    if d.dtype not in (torch.float64, torch.complex128):
        raise TypeError(f"transforms compute in float64 / complex128, got {d.dtype}")


def _apply(d, Ar, Ai, nout: int):
    torch = _torch()
    x = d.contiguous()
    cplx = x.dtype == torch.complex128
    out = torch.empty((x.shape[0], nout), dtype=x.dtype, device=x.device)
    if x.shape[0]:
        _lib.check(_lib.lib().cn_tr_rows(x.data_ptr(), out.data_ptr(), int(x.shape[0]), int(x.shape[1]), nout,
                                         2 if cplx else 1, Ar.data_ptr(), Ai.data_ptr() if Ai is not None else None,
                                         torch.cuda.current_stream(x.device).cuda_stream), "cn_tr_rows")
    return out


def forward(d, b: _Basis, dealias: bool = True):
    """Device rows of grid values (batch, M) -> device coefficients (batch, N), along the last axis."""
    torch = _torch()
    _check_even(b)
    M, N = grid_size(b, dealias), int(b.size)
    if N > M:
        raise OptionError(f"transforms: {N} coefficients from {M} grid points", fix="a grid of at least N points")
    _check_rows(d, M, "grid points")
    if isinstance(b, Chebyshev):
        Ar, Ai = _maps("cheb", "forward", M, N, d.device)
    elif isinstance(b, RealFourier):
        if d.dtype == torch.complex128:
            raise TypeError("RealFourier takes real data")
        Ar, Ai = _maps("rf", "forward", M, N, d.device)
    elif isinstance(b, ComplexFourier):
        if d.dtype != torch.complex128:
            raise TypeError("ComplexFourier takes complex data (convert on the host explicitly)")
        Ar, Ai = _maps("cf", "forward", M, N, d.device)
    else:
        raise NotImplementedError(f"transforms: {type(b).__name__} not supported")
    return _apply(d, Ar, Ai, N)


def backward(d, b: _Basis, dealias: bool = True):
    """Device coefficients (batch, N) -> device grid values (batch, M)."""
    torch = _torch()
    _check_even(b)
    M, N = grid_size(b, dealias), int(b.size)
    if N > M:
        raise OptionError(f"transforms: {N} coefficients onto {M} grid points", fix="a grid of at least N points")
    _check_rows(d, N, "coefficients")
    if isinstance(b, Chebyshev):
        Ar, Ai = _maps("cheb", "backward", M, N, d.device)
    elif isinstance(b, RealFourier):
        if d.dtype == torch.complex128:
            raise TypeError("RealFourier backward takes real coefficients (transform Re and Im apart)")
        Ar, Ai = _maps("rf", "backward", M, N, d.device)
    elif isinstance(b, ComplexFourier):
        if d.dtype != torch.complex128:
            raise TypeError("ComplexFourier takes complex data (convert on the host explicitly)")
        Ar, Ai = _maps("cf", "backward", M, N, d.device)
    else:
        raise NotImplementedError(f"transforms: {type(b).__name__} not supported")
    return _apply(d, Ar, Ai, M)
