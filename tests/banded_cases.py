"""Generators of the reference's bordered-banded tests (restated from their description):
``almost_banded`` (tests/core/test_core_banded.py:9-18), ``sparse_banded`` / ``conditioned``
(tests/cross_backend/test_banded_level.py:19-38), and the microbench systems (B systems of size N,
band kl = ku = 4 with diagonal + 8, two dense boundary rows of +-1, Dirichlet Chebyshev values)."""
import numpy as np
import scipy.sparse as sp


def almost_banded(rng, n, kl, ku, border, dtype):
    A = np.zeros((n, n), dtype=dtype)
    cplx = np.dtype(dtype) == np.complex128
    for i in range(n):
        for j in range(max(0, i - kl), min(n, i + ku + 1)):
            A[i, j] = rng.standard_normal() + (1j * rng.standard_normal() if cplx else 0)
        A[i, i] += 4 * (kl + ku + 1)
    for i in border:
        A[i] = rng.standard_normal(n) + (1j * rng.standard_normal(n) if cplx else 0)
    return A


def bordered_case(rng, n, kl, ku, r, dtype):
    """almost_banded built on the permuted layout, scattered back so that the border rows sit at
    random positions (tests/core/test_core_banded.py:24-31)."""
    border = sorted(rng.choice(n, size=r, replace=False).tolist())
    Ap = almost_banded(rng, n, kl, ku, range(r), dtype)
    order = border + [i for i in range(n) if i not in border]
    A = np.empty_like(Ap)
    A[order] = Ap
    return A, border


def sparse_banded(rng, n, kl, ku, r, dtype, density=0.4):
    A = np.zeros((n, n), dtype=dtype)
    cplx = np.dtype(dtype) == np.complex128
    for i in range(n):
        for j in range(max(0, i - kl), min(n, i + ku + 1)):
            if j == i or rng.random() < density:
                A[i, j] = rng.standard_normal() + (1j * rng.standard_normal() if cplx else 0)
    for i in range(r):
        A[i] = rng.standard_normal(n) + (1j * rng.standard_normal(n) if cplx else 0)
    return A


def conditioned(rng, n, kl, ku, r, shift=8.0, dtype=float):
    A = sparse_banded(rng, n, kl, ku, r, dtype)
    A[np.arange(r, n), np.arange(r, n)] += shift
    return sp.csr_matrix(A)


def microbench_csr(rng, B, N, kl=4, ku=4, r=2, dtype=np.complex128):
    """B systems (CSR pieces as numpy arrays) of the microbench: rows 0..r-1 are the boundary rows
    (row 0: T_n(1) = 1, row 1: T_n(-1) = (-1)^n), rows r..N-1 banded with N(0,1) (+ i N(0,1))
    entries and the diagonal shifted by +8.  Returns (indptr, indices, data) lists."""
    cplx = np.dtype(dtype) == np.complex128
    rows, cols = [], []
    for i in range(N):
        if i < r:
            js = np.arange(N)
        else:
            js = np.arange(max(0, i - kl), min(N, i + ku + 1))
        rows.append(np.full(js.size, i))
        cols.append(js)
    rows = np.concatenate(rows)
    cols = np.concatenate(cols)
    nnz = rows.size
    indptr = np.zeros(N + 1, np.int32)
    np.add.at(indptr, rows + 1, 1)
    indptr = np.cumsum(indptr).astype(np.int32)
    indices = cols.astype(np.int32)
    bulk = rows >= r
    diag = bulk & (rows == cols)
    datas = []
    vals = rng.standard_normal((B, nnz)) + (1j * rng.standard_normal((B, nnz)) if cplx else 0)
    vals = vals.astype(dtype)
    for a in range(r):
        sel = rows == a
        vals[:, sel] = 1.0 if a % 2 == 0 else (-1.0) ** cols[sel]
    vals[:, diag] += 8.0
    for m in range(B):
        datas.append(np.ascontiguousarray(vals[m]))
    return indptr, indices, datas


def csr_of(indptr, indices, data, N):
    return sp.csr_matrix((data, indices, indptr), shape=(N, N))


def backward_error(A, x, b):
    """Normwise backward error ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf)."""
    A = sp.csr_matrix(A)
    r = b - A @ x
    nA = abs(A).sum(axis=1).max()
    return float(np.abs(r).max() / (nA * np.abs(x).max() + np.abs(b).max()))
