"""config.json and report.json layouts (src/driver/serialization.cpp:146-413).

parse_config follows serialization.cpp:146-245: keys are SolverConfig / Target field names
(include/hsolve/solver.hpp:71-235), enums are snake_case strings, unknown keys, wrong types and
invalid enum values raise InvalidInputError, then the same semantic validation. config_to_json
writes every field (:247-281). new_report() returns a SolveReport with every key of
SolveReport::to_json (:283-413) in the same order, schema_version 1.
"""
from __future__ import annotations

import copy
import json
import math

import numpy as np

from .errors import InvalidInputError

VERSION = "0.1.0"

ENUMS = {
    "target.mode": ("nearest", "interval", "smallest", "largest"),
    "problem": ("auto", "standard", "generalized_definite", "generalized_indefinite"),
    "method": ("auto", "shift_invert_lanczos", "shift_invert_krylov_schur", "lanczos"),
    "factor_precision": ("double", "single", "auto"),
    "blr": ("auto", "on", "off"),
    "out_of_core": ("auto", "on", "off"),
    "inertia_check": ("auto", "on", "off"),
    "inertia_counter": ("auto", "on", "off"),
    "certificate": ("exact", "randomized", "auto", "off"),
    "subspace_bound": ("auto", "on", "off"),
}

ORDERING_OWN_ND = 10  # kOrderingNestedDissection (include/hsolve/factor.hpp)


def default_config() -> dict:
    """Every field with its default (solver.hpp:71-235)."""
    return {
        "target": {"mode": "nearest", "sigma": 0.0, "k": 6, "lower": 0.0, "upper": 0.0},
        "problem": "auto",
        "method": "auto",
        "tol": 1e-10,
        "inner_tol_factor": 0.1,
        "hermitian_tol": 1e-10,
        "ncv": 0,
        "block_size": 0,
        "slice_size": 0,
        "max_restarts": 1000,
        "seed": 2026,
        "threads": 0,
        "memory_budget_gb": "auto",
        "factor_precision": "double",
        "ordering": -1,
        "deterministic": False,
        "blr": "auto",
        "out_of_core": "auto",
        "inertia_check": "auto",
        "inertia_counter": "auto",
        "certificate": "exact",
        "certificate_delta": 1e-12,
        "certificate_fault_drop": 0,
        "inexact_fault_reject": False,
        "certificate_fault_count": 0,
        "structure_candidates": "off",
        "subspace_bound": "auto",
        "secure": False,
        "zero_filter_tol": -1.0,
        "complex_tol": 1e-8,
        "max_shift_retries": 5,
    }


def _is_int(v) -> bool:
    return isinstance(v, int) and not isinstance(v, bool)


def _is_num(v) -> bool:
    return isinstance(v, (int, float)) and not isinstance(v, bool)


def _enum(v, key):
    if not isinstance(v, str):
        raise InvalidInputError(f"config: '{key}' must be a string")
    allowed = ENUMS[key]
    if v not in allowed:
        raise InvalidInputError(f"config: '{key}' = \"{v}\" is not one of {{{', '.join(allowed)}}}")
    return v


def parse_config(doc, base: dict | None = None) -> dict:
    if isinstance(doc, (str, bytes)):
        try:
            doc = json.loads(doc)
        except json.JSONDecodeError as e:
            raise InvalidInputError(f"config: invalid JSON: {e}") from None
    if not isinstance(doc, dict):
        raise InvalidInputError("config: top level must be a JSON object")
    c = copy.deepcopy(base) if base is not None else default_config()
    for key, v in doc.items():
        if key == "target":
            if not isinstance(v, dict):
                raise InvalidInputError("config: 'target' must be an object")
            for tk, tv in v.items():
                full = "target." + tk
                if tk == "mode":
                    c["target"]["mode"] = _enum(tv, "target.mode")
                elif tk in ("sigma", "lower", "upper"):
                    if not _is_num(tv):
                        raise InvalidInputError(f"config: '{full}' must be a number")
                    c["target"][tk] = float(tv)
                elif tk == "k":
                    if not _is_int(tv):
                        raise InvalidInputError(f"config: '{full}' must be an integer")
                    c["target"]["k"] = int(tv)
                else:
                    raise InvalidInputError(f"config: unknown key '{full}'")
        elif key in ("problem", "method", "factor_precision", "blr", "out_of_core", "inertia_check",
                     "inertia_counter", "certificate", "subspace_bound"):
            c[key] = _enum(v, key)
        elif key in ("tol", "inner_tol_factor", "hermitian_tol", "certificate_delta", "zero_filter_tol",
                     "complex_tol"):
            if not _is_num(v):
                raise InvalidInputError(f"config: '{key}' must be a number")
            c[key] = float(v)
        elif key in ("ncv", "block_size", "slice_size", "max_restarts", "seed", "threads", "ordering",
                     "certificate_fault_drop", "certificate_fault_count", "max_shift_retries"):
            if not _is_int(v):
                raise InvalidInputError(f"config: '{key}' must be an integer")
            c[key] = int(v)
        elif key == "memory_budget_gb":
            if isinstance(v, str) and v == "auto":
                c[key] = "auto"
            elif _is_num(v):
                c[key] = float(v) if v > 0 else "auto"
            else:
                raise InvalidInputError("config: 'memory_budget_gb' must be a number of GiB or \"auto\"")
        elif key in ("deterministic", "inexact_fault_reject", "secure"):
            if not isinstance(v, bool):
                raise InvalidInputError(f"config: '{key}' must be true/false")
            c[key] = v
        elif key == "structure_candidates":
            if v not in ("auto", "off"):
                raise InvalidInputError(f"config: 'structure_candidates' = \"{v}\" is not one of {{auto, off}}")
            c[key] = v
        else:
            raise InvalidInputError(f"config: unknown key '{key}'")
    t = c["target"]
    if t["k"] < 1 and t["mode"] != "interval":
        raise InvalidInputError("config: target.k must be >= 1")
    if t["mode"] == "interval" and not (t["lower"] < t["upper"]):
        raise InvalidInputError("config: interval needs target.lower < target.upper")
    if not (0.0 < c["tol"] < 1.0):
        raise InvalidInputError("config: tol must be in (0, 1)")
    if not (0.0 < c["inner_tol_factor"] <= 1.0):
        raise InvalidInputError("config: inner_tol_factor must be in (0, 1]")
    if (c["ncv"] < 0 or c["block_size"] < 0 or c["slice_size"] < 0 or c["max_restarts"] < 1 or c["threads"] < 0
            or c["max_shift_retries"] < 0):
        raise InvalidInputError("config: ncv, block_size, slice_size, threads, max_shift_retries must be >= 0 and "
                                "max_restarts >= 1")
    mb = c["memory_budget_gb"]
    if mb != "auto" and (not math.isfinite(mb) or mb < 0):
        raise InvalidInputError("config: memory_budget_gb must be > 0, or \"auto\" / 0")
    if not (0.0 < c["certificate_delta"] < 1.0):
        raise InvalidInputError("config: certificate_delta must be in (0, 1)")
    if c["certificate_fault_drop"] < 0:
        raise InvalidInputError("config: certificate_fault_drop must be >= 0")
    fc = c["certificate_fault_count"]
    if fc < 0 or fc > 255 or (fc & 48) == 48:
        raise InvalidInputError("config: certificate_fault_count must be a sum of the bits 1, 2, 4, 8, 16 or 32 "
                                "(not both), 64, 128 (0..255)")
    o = c["ordering"]
    if o != ORDERING_OWN_ND and (o < -1 or o > 7 or o == 1):
        raise InvalidInputError(f"config: ordering must be -1 (adaptive), a MUMPS ICNTL(7) value 0 or 2..7, or "
                                f"{ORDERING_OWN_ND} (hsolve's nested dissection), got {o}")
    return c


def config_to_json(c: dict, indent: int = 2) -> str:
    out = copy.deepcopy(c)
    return json.dumps(out, indent=indent)


def _finite_or_none(x):
    return x if (x is not None and math.isfinite(x)) else None


def new_report() -> dict:
    """A SolveReport with every key of serialization.cpp:320-413, in the same order."""
    return {
        "schema_version": 1,
        "solver": "hsolve",
        "version": VERSION,
        "n": 0,
        "nnz_a": 0,
        "nnz_b": 0,
        "complex_scalars": False,
        "problem": "auto",
        "norm2_a": 0.0,
        "norm2_b": 1.0,
        "scale": 0.0,
        "hermitian_defect_a": 0.0,
        "hermitian_defect_b": 0.0,
        "method": "auto",
        "candidate_source": "iterative",
        "sigma_requested": 0.0,
        "sigma_used": 0.0,
        "shift_retries": 0,
        "block_size": 1,
        "factor_options": {"precision": "double", "ordering": ORDERING_OWN_ND, "scaling": 0, "threads": 0,
                           "mem_relax_pct": 0, "max_memory_mb": 0.0, "blr": False, "blr_tol": 0.0,
                           "blr_estimates": False, "out_of_core": False, "ooc_dir": "",
                           "ooc_on_memory_error": False, "null_pivot_tol": 1e-8, "verbosity": 0,
                           "deterministic": True},
        "factor_stats": {"analyze_s": 0.0, "factorize_s": 0.0, "solve_s": 0.0, "factorizations": 0, "solves": 0,
                         "rhs_solved": 0,
                         "estimate": {"factor_entries": 0, "incore_mb": 0.0, "ooc_mb": 0.0, "blr_incore_mb": -1.0,
                                      "blr_ooc_mb": -1.0, "flops": 0.0},
                         "factor_entries": 0, "factor_mb_used": 0.0, "workspace_reserved_mb": 0.0, "null_pivots": 0,
                         "delayed_pivots": 0, "workspace_retries": 0, "ooc_switches": 0, "last_info1": 0,
                         "last_info2": 0},
        "restarts": 0,
        "op_applies": 0,
        "time_total_s": 0.0,
        "time_setup_s": 0.0,
        "time_factor_s": 0.0,
        "time_iterate_s": 0.0,
        "time_verify_s": 0.0,
        "time_certificate_s": 0.0,
        "peak_rss_bytes": 0,
        "memory": {"budget_gb": 0.0, "source": "default", "limit_gb": -1.0, "limit_source": "", "measure": "resident",
                   "reserved_gb": 0.0, "limits": []},
        "precision": {"requested": "double", "used": "double", "blr": False, "decision": "", "fallback": False,
                      "fallback_reason": "", "attempt_s": 0.0},
        "parallel": {"processes": 1, "threads": 0, "helper_threads": 0, "helper_counts": 0, "helper_slices": 0,
                     "fallbacks": 0, "share_s": 0.0, "wait_s": 0.0, "helper_busy_s": 0.0, "notes": []},
        "pairs": [],
        "certificate": {"performed": False, "lower": 0.0, "upper": 0.0, "count_inertia": -1, "count_found": -1,
                        "endpoint_moves": 0, "passed": False, "note": "", "method": "none", "rigorous": False,
                        "failure_probability": 1.0, "solves": 0, "seconds": 0.0, "statement": ""},
        "subspace": {"performed": False, "max_sin_bound": None, "rigorous": False, "failure_probability": 1.0,
                     "residual_norm": "", "b_min": 0.0, "b_min_source": "", "groups": [],
                     "note": "not computed by this engine", "seconds": 0.0},
        "slices": [],
        "zero_cluster_dim": -1,
        "zero_modes_filtered": 0,
        "success": False,
        "warnings": [],
    }


def report_to_json(rep: dict, indent: int = 2) -> str:
    """JSON text of a report; non-finite window ends become null (serialization.cpp:110)."""

    def clean(o):
        if isinstance(o, np.generic):   # numpy scalars (bool_, integers, floats) as Python values
            o = o.item()
        if isinstance(o, float):
            return o if math.isfinite(o) else None
        if isinstance(o, dict):
            return {k: clean(v) for k, v in o.items()}
        if isinstance(o, (list, tuple)):
            return [clean(v) for v in o]
        return o

    return json.dumps(clean(rep), indent=indent)
