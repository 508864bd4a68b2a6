"""Exception classes raised by hsolve_amd.

The class names, their base classes and the ``code`` strings are the reference's Python error
contract (python/hsolve/__init__.py:48-130), so callers that catch the reference's errors catch
these; the code names follow hs::Error::Code (include/hsolve/error.hpp:15-25), whose order is also
the C ABI's error numbering (include/hsg.h: code = 1 + position in ERROR_CLASSES)."""
from __future__ import annotations


class HsolveError(Exception):
    """Root of the hierarchy; ``code`` names the failure category in snake_case."""

    code = "unknown"

    def __init__(self, message: str = "", code: str | None = None):
        super().__init__(message)
        if code is not None:
            self.code = code


class InvalidInputError(HsolveError, ValueError):
    """The request itself is wrong: a non-square or non-numeric matrix, an out-of-range k, a
    configuration key with a bad value."""

    code = "invalid_input"


class HsolveIOError(HsolveError, OSError):
    """Reading a matrix or config file failed, or its contents are not in a readable format."""

    code = "io"


class NotHermitianError(HsolveError, ValueError):
    """A or B is further from Hermitian than hermitian_tol allows (relative Frobenius defect)."""

    code = "not_hermitian"


class UnsupportedError(HsolveError, NotImplementedError):
    """The request is well formed but outside what this engine solves, for instance a B that is
    not positive definite."""

    code = "unsupported"


class FactorizationFailedError(HsolveError, RuntimeError):
    """The sparse LDL^T of A - sigma B could not be completed."""

    code = "factorization_failed"


class SingularError(HsolveError, ArithmeticError):
    """A - sigma B stayed numerically singular at the shift and at every nearby shift tried."""

    code = "singular"


class OutOfMemoryBudgetError(HsolveError, MemoryError):
    """The factorization needs more memory than the budget (device memory here) provides."""

    code = "out_of_memory_budget"


class NotConvergedError(HsolveError, RuntimeError):
    """The eigensolver ran out of iterations before the requested pairs converged."""

    code = "not_converged"


class InternalError(HsolveError, RuntimeError):
    """An invariant of the engine itself was violated: a defect in this package."""

    code = "internal"


class HsolveWarning(UserWarning):
    """Warns that solve() returned pairs without reaching the requested tolerance
    (report['success'] is False)."""


ERROR_CLASSES = (InvalidInputError, HsolveIOError, NotHermitianError, UnsupportedError, FactorizationFailedError,
                 SingularError, OutOfMemoryBudgetError, NotConvergedError, InternalError)

# C ABI codes (include/hsg.h): 1 + position in ERROR_CLASSES
_BY_CODE = {i + 1: cls for i, cls in enumerate(ERROR_CLASSES)}


def error_from_code(code: int, message: str) -> HsolveError:
    cls = _BY_CODE.get(int(code), InternalError)
    return cls(message)
