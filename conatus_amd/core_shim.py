"""Drop-in stand-in for the reference's compiled module ``conatus._core``.

The reference's Python layer imports ``conatus._core`` and takes a module already present in
``sys.modules`` as it is (python/conatus/runtime/__init__.py:55-59).  This module carries the
names and call signatures of the nanobind bindings (core/bindings/module.cpp:128-255, the host
part, and 321-443, the ``cuda`` submodule), implemented on libcnhip.so:

    BorderedBanded{D,Z}(n, indptr, indices, data, border, col_border=False)
        .solve(b (nrhs, n), transpose=False) -> new (nrhs, n) array
        .n .kl .ku .rank .col_border .col_border_nnz .capacitance_cond
    BorderedBandedBatch{D,Z}(mats: [(n, indptr, indices, data)], border, col_border=False)
        .solve(b (batch, n), transpose=False, threads=1) -> new (batch, n) array; len(); .kl .ku
    Imex{D,Z}(n, M_indptr, M_indices, M_data, L_indptr, L_indices, L_data, border, scheme, dt)
        .set_rhs(F(x, t) -> (n,) array, or None) .run(x0, t0, nsteps) -> (x, t) .reset()
        .history_length .n_factorizations .iteration .scheme
    cuda (submodule): BandedBatch{D,Z}, Imex{D,Z}, DeviceArray{D,Z}, available, device_name, ...
    cuda_compiled() -> True

so that ``sys.modules["conatus._core"] = conatus_amd.core_shim`` before ``import conatus`` routes
the reference's BorderedBanded, BorderedBandedBatch, ImexIntegrator and runtime.cuda classes to
the gfx950 kernels.  The host classes stage their NumPy operands through device memory: the
factorization, every solve and every integrator step run on the GPU; only the Python callback F
of ``Imex`` runs on the host, between the device stages of a step.  Errors keep the binding's
exception types: ValueError for bad CSR, shapes, non-finite entries and foreign (x0, t0) with a
history (the message says to call reset()), RuntimeError containing "singular", OverflowError
for sizes.  ``threads`` is accepted and changes nothing (every system is one lane or wavefront of
one launch; the reference's results do not depend on it either).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from . import _shim_cuda as cuda  # noqa: F401  (the reference's `_core.cuda` submodule)
from .runtime.banded import BatchHandle, _device_solve

__all__ = ["BorderedBandedD", "BorderedBandedZ", "BorderedBandedBatchD", "BorderedBandedBatchZ",
           "ImexD", "ImexZ", "cuda", "cuda_compiled", "__version__"]

__version__ = "0.5.0"          # the reference version whose binding surface this module carries


def cuda_compiled() -> bool:
    """The device backend is part of this module (libcnhip.so is gfx950 code)."""
    return True


def csr_triple(n, indptr, indices, data, dtype) -> tuple:
    """(n, indptr, indices, data) as contiguous arrays, with the structure checks of the binding's
    to_csr (core/bindings/module.cpp:75-95): n >= 0, n + 1 row pointers starting at 0, rising,
    ending at nnz, and every column inside [0, n)."""
    n = int(n)
    if n < 0:
        raise ValueError(f"CSR: n must be >= 0, got {n}")
    ip = np.ascontiguousarray(indptr, dtype=np.int32)
    ix = np.ascontiguousarray(indices, dtype=np.int32)
    dv = np.ascontiguousarray(data, dtype=dtype)
    if ip.ndim != 1 or ip.shape[0] != n + 1:
        raise ValueError("indptr must have n+1 entries")
    if ix.shape != dv.shape or ip[0] != 0 or ip[-1] != dv.size:
        raise ValueError("inconsistent CSR arrays (indptr[0] = 0, indptr[n] = nnz = len(indices) = len(data))")
    steps = np.diff(ip)
    if (steps < 0).any():
        raise ValueError(f"CSR: indptr decreases at row {int(np.argmax(steps < 0))}")
    if ix.size and (ix.min() < 0 or ix.max() >= n):
        bad = int(ix[(ix < 0) | (ix >= n)][0])
        raise ValueError(f"CSR: column index {bad} outside [0, {n})")
    return (n, ip, ix, dv)


class _Single:
    _dtype = np.float64

    def __init__(self, n, indptr, indices, data, border, col_border=False):
        self._h = BatchHandle([csr_triple(n, indptr, indices, data, self._dtype)], [int(i) for i in border],
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
        triples = [csr_triple(n, ip, ix, d, self._dtype) for (n, ip, ix, d) in mats]
        self._h = BatchHandle(triples, [int(i) for i in border], self._dtype, True, bool(col_border))
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


class _HostF:
    """The explicit term of a host integrator: the Python callback F(x, t) -> (n,) array (None: 0),
    evaluated on a host copy of the state and written back into the device buffer."""

    wants_time = True

    def __init__(self, n: int, dtype):
        import torch
        self.n, self.dtype, self.fn = n, np.dtype(dtype), None
        tdt = torch.complex128 if self.dtype == np.complex128 else torch.float64
        self._zero = torch.zeros(n, dtype=tdt, device="cuda")

    def base(self):
        return self._zero

    def write(self, x, out, t: float) -> None:
        import torch
        if self.fn is None:
            out.zero_()
            return
        r = self.fn(x.reshape(-1).cpu().numpy(), float(t))
        r = np.asarray(r)
        if r.ndim != 1 or r.shape[0] != self.n:
            raise ValueError("rhs callback: wrong length")
        if r.dtype != self.dtype and np.promote_types(r.dtype, self.dtype) != self.dtype:
            raise TypeError(f"rhs callback: a {r.dtype} result for a {self.dtype.name} integrator")
        out.copy_(torch.from_numpy(np.ascontiguousarray(r, dtype=self.dtype)))


class _Imex:
    """Host IMEX integrator (the reference's ImexIntegrator binding): the state and every step on
    the device (runtime.imex.Imex with one system), F from a Python callback."""
    _dtype = np.float64

    def __init__(self, n, mp, mi, md, lp, li, ld, border, scheme, dt):
        from .runtime.imex import SCHEMES, Imex
        if not float(dt) > 0:
            raise ValueError("dt must be > 0")
        if scheme not in SCHEMES:
            raise ValueError(f"unknown IMEX scheme: {scheme}")
        Mt = csr_triple(n, mp, mi, md, self._dtype)
        Lt = csr_triple(n, lp, li, ld, self._dtype)
        M = sp.csr_matrix((Mt[3], Mt[2], Mt[1]), shape=(Mt[0], Mt[0]))
        L = sp.csr_matrix((Lt[3], Lt[2], Lt[1]), shape=(Lt[0], Lt[0]))
        self._n = Mt[0]
        self._F = _HostF(self._n, self._dtype)
        self._im = Imex([M], [L], [int(b) for b in border], scheme, float(dt), dtype=self._dtype, explicit=self._F)

    def set_rhs(self, f):
        if f is not None and not callable(f):
            raise TypeError("set_rhs: a callable F(x, t) or None")
        self._F.fn = f

    def reset(self):
        """Forget the multistep history and the iteration count (factorizations kept); the next
        run() starts a new trajectory."""
        self._im.reset()

    def run(self, x0, t0, nsteps):
        import torch
        x0 = np.ascontiguousarray(x0, dtype=self._dtype)
        if x0.ndim != 1 or x0.shape[0] != self._n:
            raise ValueError("state length != n")
        d = torch.from_numpy(x0.copy()).to("cuda").reshape(1, self._n)
        t = self._im.run(d, float(t0), int(nsteps))
        return d.reshape(-1).cpu().numpy(), t

    history_length = property(lambda self: self._im.history_length)
    n_factorizations = property(lambda self: self._im.n_factorizations)
    iteration = property(lambda self: self._im.iteration)
    scheme = property(lambda self: self._im.scheme)


class BorderedBandedD(_Single):
    _dtype = np.float64


class BorderedBandedZ(_Single):
    _dtype = np.complex128


class BorderedBandedBatchD(_Batch):
    _dtype = np.float64


class BorderedBandedBatchZ(_Batch):
    _dtype = np.complex128


class ImexD(_Imex):
    _dtype = np.float64


class ImexZ(_Imex):
    _dtype = np.complex128
