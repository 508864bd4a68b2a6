"""Phase timing of the GPU engine (hsg_timing_enable / hsg_timing_read): HIP events recorded on
each call's stream around the numeric factorization, the inverse-panel build and the forward and
backward sweeps of every solve, with their kernel launches and algorithmic work; at detail level 2
also every front-update GEMM launch of the factorizations (phase "update")."""
from __future__ import annotations

import ctypes as C

from . import _lib

PHASES = {"factor": 0, "inverse": 1, "forward": 2, "backward": 3, "update": 4}
UNITS = {"factor": "flop", "inverse": "flop", "forward": "byte", "backward": "byte", "update": "flop"}


def enable(on: bool = True, detail: bool = False) -> None:
    """Starts (or stops) recording; clears what was recorded before. detail: also record each
    front-update GEMM launch (an event pair per launch)."""
    _lib.call("hsg_timing_enable", (2 if detail else 1) if on else 0)


def read(phase: str) -> dict:
    """Sums over the recorded calls of one phase (synchronizes their events)."""
    ms, launches, work, calls = C.c_double(), C.c_int64(), C.c_double(), C.c_int64()
    _lib.call("hsg_timing_read", PHASES[phase], C.byref(ms), C.byref(launches), C.byref(work), C.byref(calls))
    t = ms.value * 1e-3
    return {"phase": phase, "seconds": t, "launches": launches.value, "work": work.value, "unit": UNITS[phase],
            "calls": calls.value, "rate": work.value / t if t > 0 else 0.0,
            "avg_launch_s": t / launches.value if launches.value else 0.0}
