"""Host dense linear algebra of the driver (projected eigenproblems, Cholesky factors of small
Gram matrices), serialized across threads: spectrum slices solved concurrently call it from
several threads, and LAPACK/BLAS builds are not all safe to enter concurrently. The calls are
small (order <= a few hundred) and run while other slices keep the device busy."""
from __future__ import annotations

import threading

import numpy as np
import scipy.linalg as sla

_LOCK = threading.Lock()


def eigh(a: np.ndarray):
    with _LOCK:
        return np.linalg.eigh(a)


def cholesky_upper(a: np.ndarray) -> np.ndarray:
    with _LOCK:
        return sla.cholesky(a, lower=False)


def inv(a: np.ndarray) -> np.ndarray:
    with _LOCK:
        return np.linalg.inv(a)
