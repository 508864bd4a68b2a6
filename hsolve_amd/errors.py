"""Exception classes of the hsolve API (python/hsolve/__init__.py:48-130), keyed by the
hs::Error::Code names (include/hsolve/error.hpp:15-25)."""
from __future__ import annotations


class HsolveError(Exception):
    """Base class of every hsolve error. ``code`` is the error code name (snake_case)."""

    code = "unknown"

    def __init__(self, message: str = "", code: str | None = None):
        super().__init__(message)
        if code is not None:
            self.code = code


class InvalidInputError(HsolveError, ValueError):
    """Malformed or unsupported input: bad shapes, dtypes, configuration values, k, ..."""

    code = "invalid_input"


class HsolveIOError(HsolveError, OSError):
    """File system or file format error (matrix files, config.json)."""

    code = "io"


class NotHermitianError(HsolveError, ValueError):
    """A (or B) fails the Hermitian check: ||M - M^H||_F / ||M||_F > hermitian_tol."""

    code = "not_hermitian"


class UnsupportedError(HsolveError, NotImplementedError):
    """A valid request this build cannot serve (e.g. indefinite B)."""

    code = "unsupported"


class FactorizationFailedError(HsolveError, RuntimeError):
    """The sparse LDL^T factorization failed."""

    code = "factorization_failed"


class SingularError(HsolveError, ArithmeticError):
    """A - sigma B is numerically singular at and near the requested shift."""

    code = "singular"


class OutOfMemoryBudgetError(HsolveError, MemoryError):
    """The predicted memory exceeds the memory budget."""

    code = "out_of_memory_budget"


class NotConvergedError(HsolveError, RuntimeError):
    code = "not_converged"


class InternalError(HsolveError, RuntimeError):
    """A broken internal invariant (a bug)."""

    code = "internal"


class HsolveWarning(UserWarning):
    """Issued when a solve finishes without meeting the tolerance (report['success'] is False)."""


ERROR_CLASSES = (InvalidInputError, HsolveIOError, NotHermitianError, UnsupportedError, FactorizationFailedError,
                 SingularError, OutOfMemoryBudgetError, NotConvergedError, InternalError)

# C ABI codes (include/hsg.h): 1 + position in ERROR_CLASSES
_BY_CODE = {i + 1: cls for i, cls in enumerate(ERROR_CLASSES)}


def error_from_code(code: int, message: str) -> HsolveError:
    cls = _BY_CODE.get(int(code), InternalError)
    return cls(message)
