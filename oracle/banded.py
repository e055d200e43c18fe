"""CPU oracle of the bordered-banded factor/solve — TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import this module, and only as the checker / the timed CPU baseline.  The product path
(``conatus_amd``) never calls it: it runs the HIP kernels of ``libcnhip.so`` or raises.

It restates the reference algorithm of ``core/include/conatus/banded.hpp`` in numpy + LAPACK
(``scipy.linalg.lapack.{z,d}gbtrf/gbtrs/getrf/getrs``, the same LAPACK routines the reference calls
through ``core/include/conatus/lapack.hpp:30-72``):

* layout (``make_bordered_layout``, banded.hpp:170-251): border rows first (``perm``), band of the
  bulk rows measured on their non-zero entries (banded.hpp:144-168), border row i replaced in B by
  s_i e_i with s_i = max |row entry| (1 for an empty row), V^T = border rows of P A - diag(s);
* factor (banded.hpp:257-305): gbtrf of B, W = B^-1 U (U = first r unit vectors) by gbtrs,
  S = I + V^T W (banded.hpp:375-387), getrf(S) and cond_1(S) from the explicit inverse
  (banded.hpp:403-432);
* solve (banded.hpp:316-341): y = P b, y = B^-1 y, z = S^-1 V^T y, x = y - W z;
* transposed solve (banded.hpp:346-372, 452-462): Wt = B^-T V, St = I + U^T Wt,
  y = B^-T b, z = St^-1 y[:r], y -= Wt z, x = P^T y;
* the batch thread rule ``threads_for`` (python/conatus/runtime/banded.py:28-33).

Pinning (what checks this restatement, tests/test_oracle_banded.py): dense ``numpy.linalg.solve``
on the reference's own generators (tests/core/test_core_banded.py:9-39,
tests/cross_backend/test_banded_level.py:19-38), the backward-error bound of those tests
(residual < 1e-12 max|b| n), the singular-band refusal (tests/core/test_core_banded.py:63-67), the
closed-form layout of a hand-built 6x6 system, and an independent dense Woodbury evaluation.
The column-border variant (banded.hpp:10-27) is not restated: it is out of scope here.
"""
from __future__ import annotations

import math
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import scipy.linalg.lapack as lapack
import scipy.sparse as sp


class OracleError(Exception):
    pass


def _lapack(dtype):
    z = np.dtype(dtype) == np.complex128
    return ((lapack.zgbtrf, lapack.zgbtrs, lapack.zgetrf, lapack.zgetrs) if z else
            (lapack.dgbtrf, lapack.dgbtrs, lapack.dgetrf, lapack.dgetrs))


def measure_band(A: sp.csr_matrix, border):
    """kl, ku of the bulk rows of P A (banded.hpp:144-168): exact zeros do not count."""
    n = A.shape[0]
    r = len(border)
    isb = np.zeros(n, bool)
    for i in border:
        if i < 0 or i >= n:
            raise ValueError("border row out of range")
        if isb[i]:
            raise ValueError("duplicate border row")
        isb[i] = True
    perm = list(border) + [i for i in range(n) if not isb[i]]
    pos = np.empty(n, np.int64)
    pos[perm] = np.arange(n)
    C = A.tocoo()
    keep = (C.data != 0) & (~isb[C.row])
    i = pos[C.row[keep]]
    j = C.col[keep]
    kl = int(max(0, (i - j).max())) if i.size else 0
    ku = int(max(0, (j - i).max())) if i.size else 0
    return kl, ku, np.asarray(perm, np.int64)


def layout(A, border, dtype, kl_min=0, ku_min=0):
    """Unfactored bordered layout (banded.hpp:170-251): perm, LAPACK band ab (ldab x n), V^T, s."""
    A = sp.csr_matrix(A)
    if not np.all(np.isfinite(A.data)):
        raise ValueError("BorderedBanded: non-finite matrix entry: the operator must be finite")
    n = A.shape[0]
    r = len(border)
    kl, ku, perm = measure_band(A, border)
    kl, ku = max(kl, kl_min), max(ku, ku_min)
    ldab = 2 * kl + ku + 1
    ab = np.zeros((ldab, n), dtype)
    Vt = np.zeros((r, n), dtype)
    s = np.ones(r)
    for i in range(n):
        o = perm[i]
        lo, hi = A.indptr[o], A.indptr[o + 1]
        cols, vals = A.indices[lo:hi], A.data[lo:hi]
        if i < r:
            si = float(np.max(np.abs(vals))) if vals.size else 0.0
            if si == 0.0:
                si = 1.0
            np.add.at(Vt[i], cols, vals)
            ab[kl + ku, i] = si                 # AB(i, i) = s
            Vt[i, i] -= si
            s[i] = si
        else:
            nz = vals != 0
            for j, v in zip(cols[nz], vals[nz]):
                ab[kl + ku + i - j, j] += v
    return dict(n=n, r=r, kl=kl, ku=ku, ldab=ldab, perm=perm, ab=ab, Vt=Vt, s=s)


class BorderedBanded:
    """One bordered-banded system factored as the reference does (banded.hpp:253-305)."""

    def __init__(self, A, border, dtype=None, kl_min=0, ku_min=0):
        A = sp.csr_matrix(A)
        if A.shape[0] != A.shape[1]:
            raise ValueError("square matrix required")
        if dtype is None:
            dtype = np.complex128 if np.iscomplexobj(A.data) else np.float64
        self.dtype = np.dtype(dtype)
        Lo = layout(A.astype(self.dtype), [int(i) for i in border], self.dtype, kl_min, ku_min)
        self.__dict__.update(Lo)
        gbtrf, gbtrs, getrf, getrs = _lapack(self.dtype)
        lu, ipiv, info = gbtrf(self.ab, self.kl, self.ku)
        if info < 0:
            raise OracleError(f"gbtrf rejected argument {-info}")
        if info > 0:
            raise RuntimeError(f"BorderedBanded: banded part is singular at pivot {info}")
        self.lu, self.ipiv = lu, ipiv
        n, r = self.n, self.r
        self.cond = 1.0
        if r > 0:
            U = np.zeros((n, r), self.dtype)
            U[np.arange(r), np.arange(r)] = 1.0
            W, info = gbtrs(lu, self.kl, self.ku, U, ipiv)
            assert info == 0
            self.W = W.T.copy()                                  # (r, n) rows
            S = np.eye(r, dtype=self.dtype) + self.Vt @ W          # S = I + V^T W
            self.S_unfactored = S
            nS = np.abs(S).sum(axis=0).max()
            slu, spiv, info = getrf(S)
            if info > 0:
                raise RuntimeError("BorderedBanded: capacitance matrix is singular: the boundary "
                                   "rows are not independent of the banded operator")
            inv, info = getrs(slu, spiv, np.eye(r, dtype=self.dtype))
            self.cond = float(nS * np.abs(inv).sum(axis=0).max())
            self.slu, self.spiv = slu, spiv
        self._tr = None

    def solve(self, b):
        """x with A x = b; b: (n,) or (nrhs, n) rows (banded.hpp:316-341)."""
        gbtrf, gbtrs, getrf, getrs = _lapack(self.dtype)
        b = np.asarray(b, self.dtype)
        one = b.ndim == 1
        Bm = np.atleast_2d(b)
        y = Bm[:, self.perm].T.copy()                            # (n, nrhs)
        y, info = gbtrs(self.lu, self.kl, self.ku, y, self.ipiv)
        assert info == 0
        if self.r > 0:
            z = self.Vt @ y
            z, info = getrs(self.slu, self.spiv, z)
            y = y - self.W.T @ z
        x = y.T
        return x[0] if one else x

    def _prepare_transpose(self):
        if self._tr is not None or self.r == 0:
            return
        gbtrf, gbtrs, getrf, getrs = _lapack(self.dtype)
        Wt, info = gbtrs(self.lu, self.kl, self.ku, self.Vt.T.copy(), self.ipiv, trans=1)
        # St(a, b) = delta_ab + (column b of Wt)[a]  (banded.hpp:457-461)
        St = np.eye(self.r, dtype=self.dtype) + Wt[: self.r, :]
        slu, spiv, info = getrf(St)
        if info > 0:
            raise RuntimeError("BorderedBanded: transposed capacitance matrix is singular")
        self._tr = (Wt, slu, spiv)

    def solve_transpose(self, b):
        """x with A^T x = b (plain transpose; banded.hpp:350-372)."""
        gbtrf, gbtrs, getrf, getrs = _lapack(self.dtype)
        b = np.asarray(b, self.dtype)
        one = b.ndim == 1
        y = np.atleast_2d(b).T.copy()
        y, info = gbtrs(self.lu, self.kl, self.ku, y, self.ipiv, trans=1)
        assert info == 0
        if self.r > 0:
            self._prepare_transpose()
            Wt, slu, spiv = self._tr
            z, info = getrs(slu, spiv, y[: self.r].copy())
            y = y - Wt @ z
        x = np.empty_like(y)
        x[self.perm] = y
        x = x.T
        return x[0] if one else x


def threads_for(batch: int, n: int, threads: int) -> int:
    """min(threads, batch, isqrt(batch n / 900)), 1 below 2 (runtime/banded.py:28-33)."""
    t = min(threads, batch, math.isqrt(batch * n // 900))
    return t if t >= 2 else 1


class BorderedBandedBatch:
    """One oracle factorization per system; solves split over threads by the reference's rule.
    ``common_band=True`` lays every system out with the batch's common (kl, ku), as the device
    batch does (core/src/cuda/banded.cu:1149-1170)."""

    def __init__(self, mats, border, dtype=None, threads=1, common_band=False):
        mats = [sp.csr_matrix(A) for A in mats]
        if dtype is None:
            dtype = np.complex128 if any(np.iscomplexobj(A.data) for A in mats) else np.float64
        kl = ku = 0
        if common_band:
            for A in mats:
                a, b_, _ = measure_band(A, border)
                kl, ku = max(kl, a), max(ku, b_)
        self.sys = [BorderedBanded(A, border, dtype, kl, ku) for A in mats]
        self.n = mats[0].shape[0]
        self.threads = threads_for(len(mats), self.n, threads)

    def solve(self, b, transpose=False):
        b = np.asarray(b)
        out = np.empty_like(b, dtype=self.sys[0].dtype)

        def one(k):
            s = self.sys[k]
            out[k] = s.solve_transpose(b[k]) if transpose else s.solve(b[k])

        if self.threads > 1:
            with ThreadPoolExecutor(self.threads) as ex:
                list(ex.map(one, range(len(self.sys))))
        else:
            for k in range(len(self.sys)):
                one(k)
        return out


def dense_woodbury_solve(A, border, b):
    """Independent dense evaluation of the bordered Woodbury solve (same algebra, numpy.linalg
    only): P A = B + U V^T, x = y - W S^-1 V^T y with y = B^-1 P b.  Used to pin the oracle's
    layout and capacitance arithmetic without LAPACK band storage."""
    A = np.asarray(A.todense() if sp.issparse(A) else A)
    n = A.shape[0]
    r = len(border)
    isb = np.zeros(n, bool)
    isb[list(border)] = True
    perm = list(border) + [i for i in range(n) if not isb[i]]
    PA = A[perm]
    Bm = PA.copy()
    s = np.array([np.max(np.abs(PA[i])) or 1.0 for i in range(r)])
    for i in range(r):
        Bm[i] = 0
        Bm[i, i] = s[i]
    Vt = PA[:r] - np.eye(r, n) * s[:, None]
    U = np.eye(n, r)
    y = np.linalg.solve(Bm, np.asarray(b)[perm])
    W = np.linalg.solve(Bm, U)
    S = np.eye(r) + Vt @ W
    return y - W @ np.linalg.solve(S, Vt @ y)
