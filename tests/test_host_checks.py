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
    with pytest.raises(NumericalError, match=r"non-finite entry .* row 0, column 1"):
        csr_args(A)
    Z = sp.csr_matrix(np.array([[1.0 + 1j, 0], [0, 1]]))
    with pytest.raises(DtypeError, match="imaginary"):     # the keyword the reference's tests match
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
    with pytest.raises(DtypeError, match="imaginary"):
        solve_operand(np.ones(3, complex), np.float64, "x")
    assert solve_operand(np.ones(3, np.float32), np.float64, "x").dtype == np.float64


def test_rhs_shape_check():
    with pytest.raises(InputTypeError, match="right-hand side"):
        _check_rhs_shape(np.ones((3, 4)), (4, 3), "w", "(4, 3)")
    _check_rhs_shape(np.ones((3, 4)), (3, None), "w", "(3, k)")


def test_threads_rule(monkeypatch):
    from conatus_amd.runtime.banded import solve_threads
    assert threads_for(64, 262, 8) == 4
    assert threads_for(1, 10, 8) == 1 and threads_for(100, 100, 8) == 3 and threads_for(4096, 256, 8) == 8
    monkeypatch.setenv("CONATUS_SOLVE_THREADS", "two")
    with pytest.raises(OptionError, match="CONATUS_SOLVE_THREADS"):
        solve_threads()
    monkeypatch.setenv("CONATUS_SOLVE_THREADS", "0")
    with pytest.raises(OptionError, match="CONATUS_SOLVE_THREADS"):
        solve_threads()
    monkeypatch.setenv("CONATUS_SOLVE_THREADS", "1")
    assert solve_threads() == 1
    monkeypatch.delenv("CONATUS_SOLVE_THREADS")
    monkeypatch.delenv("OMP_NUM_THREADS", raising=False)
    assert solve_threads() == 1


def test_empty_batch_is_refused_before_the_library():
    # runtime/banded.py:174-176: an empty list is a type error of the call, raised before the core
    from conatus_amd.runtime.banded import BorderedBandedBatch
    with pytest.raises(InputTypeError, match="at least one"):
        BorderedBandedBatch([], [])


def test_index_map_and_overlap_rules():
    """The refusals of the fused gather/scatter solve (python/conatus/runtime/cuda.py:186-232):
    maps must be contiguous int64 (batch, n) tensors inside the array they index, a scatter map
    must not repeat an index, and rhs/x may overlap only in place with one map for both.  Host
    tensors stand in for device ones: the rules read shapes, values and addresses only."""
    import torch
    from conatus_amd.runtime.checks import check_index_map, check_io_overlap
    ok = torch.arange(6, dtype=torch.int64).reshape(2, 3)
    check_index_map("rows", ok, 6, 2, 3, "w")
    check_index_map("rows", None, 6, 2, 3, "w")
    with pytest.raises(ValueError, match="needs 6 entries"):
        check_index_map("rows", None, 5, 2, 3, "w")
    with pytest.raises(ValueError, match="int64"):
        check_index_map("rows", ok.to(torch.int32), 6, 2, 3, "w")
    with pytest.raises(ValueError, match="contiguous"):
        check_index_map("rows", torch.arange(12, dtype=torch.int64).reshape(2, 6)[:, ::2], 12, 2, 3, "w")
    with pytest.raises(ValueError, match="out of range"):
        check_index_map("rows", ok + 1, 6, 2, 3, "w")
    rep = torch.tensor([[0, 1, 2], [3, 4, 4]], dtype=torch.int64)
    check_index_map("rows", rep, 6, 2, 3, "w")                       # a gather may repeat
    with pytest.raises(ValueError, match="repeats"):
        check_index_map("cols", rep, 6, 2, 3, "w", unique=True)
    buf = torch.zeros(12, dtype=torch.complex128)
    check_io_overlap(buf[:6], None, buf[6:], None, "w")              # disjoint halves
    check_io_overlap(buf, ok, buf, ok.clone(), "w")                   # in place, one map
    with pytest.raises(ValueError, match="overlap"):
        check_io_overlap(buf, ok, buf, ok.flip(0).contiguous(), "w")
    with pytest.raises(ValueError, match="overlap"):
        check_io_overlap(buf[:8], None, buf[4:], None, "w")
