"""Typed errors of conatus_amd, named and shaped as the reference's hierarchy
(python/conatus/errors.py:98-344): each class also derives from the built-in exception the
reference binding raises, so ``except ValueError`` / ``pytest.raises(RuntimeError, match=...)``
written against the reference keep working."""
from __future__ import annotations


class ConatusError(Exception):
    """Base class: ``ConatusError(message, where=, why=, fix=)``; the fields are appended to the
    message (python/conatus/errors.py:98-135)."""

    def __init__(self, message: str = "", *, where: str | None = None, why: str | None = None,
                 fix: str | None = None) -> None:
        super().__init__(message)
        self.message, self.where, self.why, self.fix = str(message), where, why, fix

    def __str__(self) -> str:
        lines = [self.message]
        for label, text in (("where", self.where), ("why", self.why), ("fix", self.fix)):
            if text:
                lines.append(f"  {label + ':':6s} {text}")
        return "\n".join(lines)


class OptionError(ConatusError, ValueError):
    """An option or argument value is refused (errors.py:236)."""


class InputTypeError(ConatusError, TypeError):
    """An input has the wrong type or shape (errors.py:285)."""


class DtypeError(InputTypeError):
    """A dtype conversion would lose information (errors.py:292)."""


class NumericalError(ConatusError, ArithmeticError):
    """A non-finite value where a finite one is required (errors.py:321)."""


class BackendError(ConatusError, RuntimeError):
    """The HIP backend is unavailable or failed (errors.py:330)."""


class UnsupportedError(ConatusError, NotImplementedError):
    """A case outside what the compiled kernels cover (errors.py:344)."""


class SingularError(ConatusError, RuntimeError):
    """A singular band pivot or capacitance matrix (the reference raises RuntimeError with
    "singular" in the message: core/include/conatus/banded.hpp:274-281)."""


class LayoutError(ConatusError, ValueError):
    """Invalid CSR structure, border rows or non-finite operator entries (the reference's
    std::invalid_argument, core/bindings/module.cpp:75-95)."""
