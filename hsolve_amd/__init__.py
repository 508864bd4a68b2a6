"""hsolve_amd: MI355X-native shift-invert eigensolver engine with the hsolve API.

    import hsolve_amd as hsolve
    res = hsolve.solve(A, B, sigma=10.0, k=21)
    res.eigenvalues, res.eigenvectors, res.report["certificate"]

The numeric path (multifrontal LDL^T, inertia, solves, Krylov kernels) runs in libhsg.so on the
GPU; importing this package does not need a GPU, solving does (there is no CPU fallback).
"""
from .api import SolveResult, default_config, solve, version
from .errors import (FactorizationFailedError, HsolveError, HsolveIOError, HsolveWarning, InternalError,
                     InvalidInputError, NotConvergedError, NotHermitianError, OutOfMemoryBudgetError,
                     SingularError, UnsupportedError)

__all__ = ["solve", "default_config", "version", "SolveResult", "HsolveError", "InvalidInputError", "HsolveIOError",
           "NotHermitianError", "UnsupportedError", "FactorizationFailedError", "SingularError",
           "OutOfMemoryBudgetError", "NotConvergedError", "InternalError", "HsolveWarning"]
__version__ = version()
