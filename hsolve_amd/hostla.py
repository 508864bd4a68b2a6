"""Host dense linear algebra of the driver: projected eigenproblems and the Cholesky factors and
inverses of small Gram matrices (order <= a few hundred). Spectrum slices solved concurrently call
these from several threads at once, each on one BLAS thread (the solver limits the BLAS pool while it
runs). The LAPACK builds bundled with numpy and scipy are reentrant (per-thread work buffers, LAPACK
compiled recursive), so the calls run in parallel: a lock here would make a slice's 8 x 8 Cholesky
factor wait behind another slice's Rayleigh-Ritz eigensolve."""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla


def eigh(a: np.ndarray):
    return np.linalg.eigh(a)


def cholesky_upper(a: np.ndarray) -> np.ndarray:
    return sla.cholesky(a, lower=False)


def inv(a: np.ndarray) -> np.ndarray:
    return np.linalg.inv(a)
