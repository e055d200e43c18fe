"""Drop-in stand-in for the reference's compiled module ``conatus._core`` (banded part).

The reference's Python layer loads ``conatus._core`` and accepts a module already present in
``sys.modules`` (python/conatus/runtime/__init__.py:55-59).  This module exposes the names and call
signatures of the nanobind bindings (core/bindings/module.cpp:128-206) on top of libcnhip.so:

    BorderedBanded{D,Z}(n, indptr, indices, data, border, col_border=False)
        .solve(b (nrhs, n), transpose=False) -> new (nrhs, n) array; .n .kl .ku .rank
        .col_border .col_border_nnz .capacitance_cond
    BorderedBandedBatch{D,Z}(mats: list[(n, indptr, indices, data)], border, col_border=False)
        .solve(b (batch, n), transpose=False, threads=1) -> new (batch, n) array; len(); .kl .ku

so that ``sys.modules["conatus._core"] = conatus_amd.core_shim`` before importing ``conatus``
routes the reference's BorderedBanded / BorderedBandedBatch (python/conatus/runtime/banded.py:130-213)
to the gfx950 kernels.  Errors keep the reference's exception types: ValueError for bad CSR /
shapes / non-finite entries, RuntimeError containing "singular", OverflowError for sizes.
``threads`` is accepted and ignored (every system is one lane of one device launch; results do
not depend on it, as the reference guarantees for its thread count).
"""
from __future__ import annotations

import numpy as np

from .runtime.banded import BatchHandle, _device_solve

__all__ = ["BorderedBandedD", "BorderedBandedZ", "BorderedBandedBatchD", "BorderedBandedBatchZ",
           "cuda_compiled"]

cuda_compiled = False   # the reference's CUDA submodule is not provided by this shim


def _csr_tuple(n, indptr, indices, data, dtype):
    indptr = np.ascontiguousarray(indptr, dtype=np.int32)
    indices = np.ascontiguousarray(indices, dtype=np.int32)
    data = np.ascontiguousarray(data, dtype=dtype)
    if int(n) < 0:
        raise ValueError(f"CSR: n must be >= 0, got {n}")
    if indptr.shape[0] != int(n) + 1:
        raise ValueError("indptr must have n+1 entries")
    return (int(n), indptr, indices, data)


class _Single:
    _dtype = np.float64

    def __init__(self, n, indptr, indices, data, border, col_border=False):
        self._h = BatchHandle([_csr_tuple(n, indptr, indices, data, self._dtype)], [int(i) for i in border],
                              self._dtype, True, bool(col_border))
        p = self._h.props()
        self.n, self.kl, self.ku, self.rank = int(p.n), int(p.kl), int(p.ku), int(p.rank)
        self.col_border = bool(col_border)
        self.col_border_nnz = 0
        self.capacitance_cond = float(self._h.cond()[0])

    def solve(self, b, transpose=False):
        b = np.ascontiguousarray(b, dtype=self._dtype)
        if b.ndim != 2 or b.shape[1] != self.n:
            raise ValueError("rhs length != n")
        return _device_solve(self._h, b[:, None, :], bool(transpose))[:, 0, :]


class _Batch:
    _dtype = np.float64

    def __init__(self, mats, border, col_border=False):
        csrs = [_csr_tuple(n, ip, ix, d, self._dtype) for (n, ip, ix, d) in mats]
        self._h = BatchHandle(csrs, [int(i) for i in border], self._dtype, True, bool(col_border))
        p = self._h.props()
        self._B, self._n = int(p.batch), int(p.n)
        self.kl, self.ku = int(p.kl), int(p.ku)

    def __len__(self):
        return self._B

    def solve(self, b, transpose=False, threads=1):
        b = np.ascontiguousarray(b, dtype=self._dtype)
        if b.ndim != 2 or b.shape[0] != self._B:
            raise ValueError("batch size mismatch")
        if b.shape[1] != self._n:
            raise ValueError("rhs length != n")
        return _device_solve(self._h, b[None], bool(transpose))[0]


class BorderedBandedD(_Single):
    _dtype = np.float64


class BorderedBandedZ(_Single):
    _dtype = np.complex128


class BorderedBandedBatchD(_Batch):
    _dtype = np.float64


class BorderedBandedBatchZ(_Batch):
    _dtype = np.complex128
