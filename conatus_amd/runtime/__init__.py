"""Runtime layer: the reference's ``conatus.runtime`` surface (runtime/__init__.py:98-102) backed by
libcnhip.so.  ``hip`` is the device backend (the reference's ``runtime.cuda``)."""
from .banded import (BorderedBanded, BorderedBandedBatch, csr_args, solve_operand,  # noqa: F401
                     solve_threads, threads_for)
from . import hip  # noqa: F401

__all__ = ["BorderedBanded", "BorderedBandedBatch", "csr_args", "solve_operand", "solve_threads",
           "threads_for", "hip"]
