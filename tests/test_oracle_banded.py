"""Pins of the CPU oracle (oracle/banded.py) before it is used as the checker.

* dense numpy.linalg on the reference's own generators and bound (tests/core/test_core_banded.py:21-39:
  residual < 1e-12 max|b| n, forward and transposed, r = 1..4, real and complex);
* an independent dense Woodbury evaluation (same algebra, no band storage) — a wrong layout index,
  a dropped s_i or a transposed V^T fails it;
* the hand-computed layout of a 6x6 system (perm, band, s, V^T);
* the singular-band refusal (tests/core/test_core_banded.py:63-67);
* batch = individual systems, and thread count does not change a bit (tests/core/test_core_threads.py:26-36).
"""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle.banded import BorderedBanded, BorderedBandedBatch, dense_woodbury_solve, layout, threads_for
from tests.banded_cases import almost_banded, bordered_case, conditioned


@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
@pytest.mark.parametrize("r", [1, 2, 3, 4])
def test_oracle_matches_dense(dtype, r, rng):
    n, kl, ku = int(rng.integers(20, 80)), int(rng.integers(0, 4)), int(rng.integers(0, 4))
    A, border = bordered_case(rng, n, kl, ku, r, dtype)
    s = BorderedBanded(sp.csr_matrix(A), border)
    assert s.r == r and s.kl <= kl and s.ku <= ku
    cplx = dtype == np.complex128
    b = rng.standard_normal((3, n)) + (1j * rng.standard_normal((3, n)) if cplx else 0)
    x = s.solve(b)
    assert np.max(abs(A @ x.T - b.T)) < 1e-12 * np.max(abs(b)) * n
    xt = s.solve_transpose(b)
    assert np.max(abs(A.T @ xt.T - b.T)) < 1e-12 * np.max(abs(b)) * n
    xw = dense_woodbury_solve(A, border, b[0])
    assert np.max(abs(x[0] - xw)) < 1e-12 * np.max(abs(xw))


def test_oracle_conditioned_generator(rng):
    """The 'conditioned' generator of tests/cross_backend/test_banded_level.py:32-38 (shift 8)."""
    for dtype in (float, complex):
        A = conditioned(rng, 300, 4, 6, 3, dtype=dtype)
        s = BorderedBanded(A, [0, 1, 2])
        b = rng.standard_normal(300) + (1j * rng.standard_normal(300) if dtype is complex else 0)
        x = s.solve(b)
        xd = np.linalg.solve(A.toarray(), b)
        assert np.max(abs(x - xd)) < 1e-11 * np.max(abs(xd))


def test_layout_closed_form():
    """6x6 system, border rows [4, 1]: perm = [4, 1, 0, 2, 3, 5]; band of the bulk rows of P A;
    s = max |row| (1 for an empty row); V^T = border rows - diag(s) (banded.hpp:204-228)."""
    A = np.zeros((6, 6))
    A[0, 0], A[0, 1] = 2.0, -1.0
    A[2, 1], A[2, 2], A[2, 3] = 1.0, 3.0, 0.5
    A[3, 3], A[3, 2] = 4.0, 0.25
    A[5, 5], A[5, 4] = 5.0, 1.5
    A[4] = [1, -7, 1, 1, 1, 1]       # border row: s = 7
    # row 1 empty border row: s = 1
    Lo = layout(sp.csr_matrix(A), [4, 1], np.float64)
    assert list(Lo["perm"]) == [4, 1, 0, 2, 3, 5]
    # bulk rows of P A at positions 2..5 are original rows 0, 2, 3, 5: entries (pos, col)
    # (2,0),(2,1) (3,1),(3,2),(3,3) (4,3),(4,2) (5,5),(5,4) -> kl = max(pos-col) = 2, ku = max(col-pos) = 0
    assert (Lo["kl"], Lo["ku"]) == (2, 0)
    assert np.allclose(Lo["s"], [7.0, 1.0])
    Vt = Lo["Vt"]
    # diag(s) sits at column i (columns are not permuted): Vt[0, 0] = 1 - 7
    assert np.allclose(Vt[0], [1 - 7, -7, 1, 1, 1, 1]) and np.allclose(Vt[1], [0, -1, 0, 0, 0, 0])
    kl, ku = 2, 0
    ab = Lo["ab"]
    AB = lambda i, j: ab[kl + ku + i - j, j]   # noqa: E731
    assert AB(0, 0) == 7.0 and AB(1, 1) == 1.0 and AB(2, 0) == 2.0 and AB(2, 1) == -1.0
    assert AB(3, 1) == 1.0 and AB(3, 2) == 3.0 and AB(5, 5) == 5.0 and AB(5, 4) == 1.5


def test_singular_band_is_refused():
    A = np.eye(6)
    A[3, 3] = 0.0
    with pytest.raises(RuntimeError, match="singular"):
        BorderedBanded(sp.csr_matrix(A), [0])


def test_batch_matches_individual_and_threads(rng):
    n, mats, bs = 40, [], []
    for _ in range(37):
        A = almost_banded(rng, n, 2, 2, [0, 1], complex)
        mats.append(sp.csr_matrix(A))
        bs.append(rng.standard_normal(n) + 1j * rng.standard_normal(n))
    b = np.array(bs)
    x1 = BorderedBandedBatch(mats, [0, 1], threads=1).solve(b)
    x3 = BorderedBandedBatch(mats, [0, 1], threads=3).solve(b)
    assert np.array_equal(x1, x3)
    for A, b_, x_ in zip(mats, b, x1):
        assert np.max(abs(A @ x_ - b_)) < 1e-11


def test_thread_rule():
    assert threads_for(1, 10, 8) == 1
    assert threads_for(100, 100, 8) == 3          # isqrt(10000 / 900) = 3
    assert threads_for(4096, 256, 8) == 8
