"""Factors shared between mode systems that are equal up to signs (the reference's
``SharedFactor``, python/conatus/timestep/shared_factors.py:1-60).

Two systems B and A with B = D_r A D_c (D_r, D_c diagonal with entries +1 or -1) have factors
equal up to signs: partial pivoting compares magnitudes, which D leaves unchanged, and every
multiply-add of the elimination and of the sweeps maps sign-flipped operands to the sign-flipped
result exactly.  So B x = b is solved with A's factors as x = D_c A^-1 (D_r b), bit for bit what
B's own factors give, with half the factor memory and, when both right-hand sides are solved in
one call, one factor stream for the pair.  The bordered layout replaces border row i by s e_i
(s > 0): the pair is shared only when d_r[i] d_c[i] = +1 on every border row, so that the band
part of B is exactly D_r (band of A) D_c.

``sign_similarity`` checks the relation exactly (never assumes it): same CSR pattern, no explicit
zeros, every entry of B equal to +- the entry of A, and the signs factor as d_r[i] d_c[j], found
by a breadth-first walk of the bipartite graph rows - columns of the pattern and then verified on
every entry.  ``SharedBandedBatch`` factors the primary systems only and solves each pair as two
right-hand sides of one system.
"""
from __future__ import annotations

from collections import deque
from dataclasses import dataclass
from typing import Sequence

import numpy as np
import scipy.sparse as sp

__all__ = ["SharedFactor", "sign_similarity", "SharedBandedBatch"]


@dataclass
class SharedFactor:
    """System ``secondary`` is solved with the factors of system ``primary``: B = diag(dr) A diag(dc)."""
    primary: int
    secondary: int
    dr: np.ndarray
    dc: np.ndarray


def _canonical(A) -> sp.csr_matrix:
    A = sp.csr_matrix(A, copy=True)
    A.sum_duplicates()
    A.sort_indices()
    return A


def sign_similarity(A, B, border: Sequence[int] = ()) -> tuple[np.ndarray, np.ndarray] | None:
    """(dr, dc) with B = diag(dr) A diag(dc) exactly, entries +-1, and dr[i] dc[i] = +1 on the
    border rows; None when no such pair exists."""
    A, B = _canonical(A), _canonical(B)
    if A.shape != B.shape or A.shape[0] != A.shape[1]:
        return None
    if not (np.array_equal(A.indptr, B.indptr) and np.array_equal(A.indices, B.indices)):
        return None
    a, b = A.data, B.data
    if np.any(a == 0) or np.any(b == 0):
        return None
    plus, minus = b == a, b == -a
    if not np.all(plus | minus):
        return None
    sgn = np.where(plus, 1, -1).astype(np.int8)
    n = A.shape[0]
    rows = np.repeat(np.arange(n), np.diff(A.indptr))
    cols = A.indices
    csc_order = np.argsort(cols, kind="stable")
    col_ptr = np.zeros(n + 1, np.int64)
    np.add.at(col_ptr, cols + 1, 1)
    col_ptr = np.cumsum(col_ptr)
    dr = np.zeros(n, np.int8)
    dc = np.zeros(n, np.int8)
    for start in range(n):
        if dr[start] or A.indptr[start] == A.indptr[start + 1]:
            continue
        dr[start] = 1
        todo = deque([("r", start)])
        while todo:
            kind, v = todo.popleft()
            if kind == "r":
                for e in range(A.indptr[v], A.indptr[v + 1]):
                    j = cols[e]
                    if not dc[j]:
                        dc[j] = dr[v] * sgn[e]
                        todo.append(("c", j))
            else:
                for q in range(col_ptr[v], col_ptr[v + 1]):
                    e = csc_order[q]
                    i = rows[e]
                    if not dr[i]:
                        dr[i] = dc[v] * sgn[e]
                        todo.append(("r", i))
    dr[dr == 0] = 1
    dc[dc == 0] = 1
    if not np.array_equal(dr[rows] * dc[cols], sgn):
        return None
    for i in border:
        if dr[i] * dc[i] != 1:
            return None
    return dr.astype(np.float64), dc.astype(np.float64)


class SharedBandedBatch:
    """Device batch of systems where the pairs of ``shares`` are solved with one factorization:
    the primaries (every system that is not a secondary) are factored, and a solve of the whole
    batch runs the primaries with two right-hand sides, the second holding D_r b of each
    secondary (zero for a primary without a partner); x of a secondary is D_c times that column.

    Solves take the batch's (batch, n) device rows, in place, on the current stream; they
    allocate nothing after the first call and can be captured in a HIP graph."""

    def __init__(self, mats: Sequence, border: Sequence[int], shares: Sequence[SharedFactor], dtype, pivot=True,
                 mode: str | None = None):
        import torch
        from .banded import BatchHandle, csr_args
        self.batch = len(mats)
        self.n = mats[0].shape[0]
        second = {s.secondary for s in shares}
        if len(second) != len(shares) or any(s.primary in second for s in shares):
            raise ValueError("SharedBandedBatch: a secondary must be shared once, with a primary that is not one")
        self.prim = [m for m in range(self.batch) if m not in second]
        pos = {m: p for p, m in enumerate(self.prim)}
        self.shares = list(shares)
        self.handle = BatchHandle([csr_args(mats[m], dtype) for m in self.prim], list(border), dtype, pivot,
                                  mode=mode)
        self.handle.reserve(2)
        self.dtype = np.dtype(dtype)
        tdt = torch.complex128 if self.dtype == np.complex128 else torch.float64
        P = len(self.prim)
        self._prim = torch.tensor(self.prim, dtype=torch.int64, device="cuda")
        self._pslot = torch.tensor([pos[s.primary] for s in shares], dtype=torch.int64, device="cuda")
        self._sec = torch.tensor([s.secondary for s in shares], dtype=torch.int64, device="cuda")
        self._dr = torch.from_numpy(np.stack([s.dr for s in shares]) if shares else np.zeros((0, self.n))).to("cuda", tdt)
        self._dc = torch.from_numpy(np.stack([s.dc for s in shares]) if shares else np.zeros((0, self.n))).to("cuda", tdt)
        self._rhs = torch.zeros((2, P, self.n), dtype=tdt, device="cuda")

    @property
    def factor_bytes(self) -> int:
        return int(self.handle.props().factor_bytes)

    def props(self):
        return self.handle.props()

    def factor(self, stream, check: bool = True) -> None:
        self.handle.factor(stream, check=check)

    def solve(self, d) -> None:
        """Solve the (batch, n) device rows ``d`` in place."""
        import torch
        r = self._rhs
        r[0].copy_(d.index_select(0, self._prim))
        r[1].zero_()
        if len(self.shares):
            r[1].index_copy_(0, self._pslot, d.index_select(0, self._sec) * self._dr)
        self.handle.solve_ptr(r.data_ptr(), 2, r[0].numel(), False, torch.cuda.current_stream().cuda_stream)
        d.index_copy_(0, self._prim, r[0])
        if len(self.shares):
            d.index_copy_(0, self._sec, r[1].index_select(0, self._pslot) * self._dc)
