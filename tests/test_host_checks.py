"""Host-side argument rules of the Python layer (python/conatus/runtime/banded.py:74-127): no device."""
import numpy as np
import pytest
import scipy.sparse as sp

from conatus_amd.errors import DtypeError, InputTypeError, NumericalError, OptionError
from conatus_amd.runtime.banded import _check_rhs_shape, csr_args, solve_operand, threads_for


def test_csr_args_refusals():
    with pytest.raises(OptionError, match="square"):
        csr_args(sp.csr_matrix(np.ones((2, 3))))
    A = sp.csr_matrix(np.array([[1.0, np.inf], [0, 1]]))
    with pytest.raises(NumericalError, match=r"\(0, 1\)"):
        csr_args(A)
    Z = sp.csr_matrix(np.array([[1.0 + 1j, 0], [0, 1]]))
    with pytest.raises(DtypeError):
        csr_args(Z, np.float64)
    n, ip, ix, d = csr_args(sp.csr_matrix(np.array([[1.0 + 0j, 0], [0, 2]])), np.float64)
    assert n == 2 and d.dtype == np.float64 and ip.dtype == np.int32


def test_csr_args_sums_duplicates():
    A = sp.csr_matrix((np.array([1.0, 2.0, 3.0]), np.array([0, 0, 1]), np.array([0, 2, 3])), shape=(2, 2))
    n, ip, ix, d = csr_args(A)
    assert list(ip) == [0, 1, 2] and list(d) == [3.0, 3.0]


def test_solve_operand_widening_only():
    b = np.ones(3, np.float32)
    assert solve_operand(b, np.complex128, "x").dtype == np.complex128
    with pytest.raises(DtypeError):
        solve_operand(np.ones(3, complex), np.float64, "x")


def test_rhs_shape_check():
    with pytest.raises(InputTypeError):
        _check_rhs_shape(np.ones((3, 4)), (4, 3), "w", "(4, 3)")
    _check_rhs_shape(np.ones((3, 4)), (3, None), "w", "(3, k)")


def test_threads_rule():
    assert threads_for(64, 262, 8) == 4


def test_empty_batch_is_refused_before_the_library():
    # runtime/banded.py:174-176: an empty list is a type error of the call, raised before the core
    from conatus_amd.runtime.banded import BorderedBandedBatch
    with pytest.raises(InputTypeError, match="no matrix"):
        BorderedBandedBatch([], [])
