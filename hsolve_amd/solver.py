"""hs::solve for HEP / GHEP (B Hermitian positive definite) on the GPU engine.

Restates driver::solve_with_helpers (src/driver/solver.cpp:523-694) for the definite path:
  * hermitian_input: finite values, relative Hermitian defect <= hermitian_tol, then the exact
    Hermitian part (solver.cpp:70-114);
  * norm estimates ||A||_2, ||B||_2 (30-step Lanczos, seed 2026; B = I -> 1) and scale
    (solver.cpp:558-562);
  * problem type: B absent -> standard; B diagonal with positive entries -> generalized definite;
    a general B is tested by the Sylvester count of B itself on the GPU (nu(B) = 0 <=> B HPD,
    solver.cpp:231-265); indefinite B (GHIEP) is unsupported here;
  * analysis, then solve_definite: nearest -> nearest.solve_nearest, interval / smallest /
    largest -> windows.WindowSolver (definite.cpp:2277-2335);
  * single-precision working factors with the FP64 fallback (solver.cpp:615-675): an attempt whose
    refinement fails, whose pairs miss tol or whose certificate fails restarts with double
    factors on the same analysis (report.precision.fallback);
  * report: SolveReport with the reference's keys (serialization.cpp:283-413), exclusive phase
    times that add up to time_total_s.
"""
from __future__ import annotations

import math
import resource
import threading
import time
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp
import torch

from . import autotune
from . import config as cfgmod
from . import subspace
from .driver_core import Ctx, PhaseClock, estimate_norm2
from .errors import InvalidInputError, NotHermitianError, UnsupportedError
from .factor import GpuFactorization, Pencil
from .ops import DeviceOps, InexactFactorFailure


def _nonzero_imag(M) -> bool:
    return M is not None and np.iscomplexobj(M.data) and bool(np.any(sp.csr_matrix(M).data.imag != 0))


def hermitian_input(M, name: str, tol: float, complex_problem: bool | None = None):
    """Checks and Hermitian part (solver.cpp:70-114): returns (csr, defect). The values are complex
    when the pencil is a complex problem (default: when M has a nonzero imaginary part), else real
    (python/src/module.cpp:279-291: complex dtypes with all-zero imaginary parts are solved as
    real)."""
    M = sp.csr_matrix(M)
    if M.shape[0] != M.shape[1]:
        raise InvalidInputError(f"{name} must be square, got {M.shape}")
    if M.nnz and not np.all(np.isfinite(M.data)):
        raise InvalidInputError(f"{name} has non-finite entries")
    if complex_problem is None:
        complex_problem = _nonzero_imag(M)
    M = M.astype(np.complex128) if complex_problem else (M.real if np.iscomplexobj(M.data) else M).astype(np.float64)
    fro = sp.linalg.norm(M) if M.nnz else 0.0
    if fro == 0.0:
        raise InvalidInputError(f"{name} is zero")
    D = (M - M.conj().T).tocsr()
    D.eliminate_zeros()
    defect = (sp.linalg.norm(D) / fro) if D.nnz else 0.0
    if defect > tol:
        raise NotHermitianError(f"{name} is not Hermitian: ||M - M^H||_F / ||M||_F = {defect:.3g} > {tol:.3g}")
    H = ((M + M.conj().T) * 0.5).tocsr()
    H.sum_duplicates()
    H.eliminate_zeros()
    H.sort_indices()
    return H, float(defect)


def _b_kind(B):
    """'identity' / 'diagonal' / 'general'."""
    if B is None:
        return "identity"
    Bc = sp.csr_matrix(B)
    n = Bc.shape[0]
    if Bc.nnz == n and np.array_equal(Bc.indices, np.arange(n)) and np.array_equal(Bc.indptr, np.arange(n + 1)):
        return "diagonal"
    off = Bc - sp.diags(Bc.diagonal())
    off.eliminate_zeros()
    return "diagonal" if off.nnz == 0 else "general"


@dataclass
class Prepared:
    """Checked input of one pencil with its device copies (see prepare())."""
    A: sp.csr_matrix
    B: sp.csr_matrix | None
    b_kind: str
    defect_a: float
    defect_b: float
    pencil: Pencil
    notes: list = field(default_factory=list)   # input remarks, reported as warnings "input: ..."


def prepare(A, B=None, config: dict | None = None) -> Prepared:
    """Input stage of a solve (solver.cpp:70-114 and the pencil set-up of pencil.cpp:67-170): the
    exact Hermitian parts of A and B with their defects, the B kind, the union pattern and the
    device copies of the values. A Prepared pencil can be solved many times (solve(prepared=...))."""
    cfg = cfgmod.parse_config(config or {})
    # a complex problem when A or B has a nonzero imaginary part; complex A with real B promotes B
    # (and vice versa), complex dtypes with all-zero imaginary parts are solved as real
    cplx = _nonzero_imag(A) or _nonzero_imag(B)
    notes = [f"{name}: real: complex input has all-zero imaginary parts; solved as real"
             for name, M in (("A", A), ("B", B)) if M is not None and not cplx and np.iscomplexobj(M.data)]
    A, def_a = hermitian_input(A, "A", cfg["hermitian_tol"], cplx)
    bk = _b_kind(B)
    def_b = 0.0
    if B is not None:
        B, def_b = hermitian_input(B, "B", cfg["hermitian_tol"], cplx)
        if B.shape != A.shape:
            raise InvalidInputError(f"A and B differ in shape: {A.shape} vs {B.shape}")
        if bk == "diagonal":
            B = sp.diags(B.diagonal(), format="csr")
    if B is not None and bk == "diagonal" and np.any(B.diagonal().real <= 0):
        raise UnsupportedError("B is diagonal but not positive definite (GHIEP): not supported by this engine")
    pencil = Pencil(A, B)
    pencil.to_device()
    return Prepared(A, B, bk, def_a, def_b, pencil, notes)


def slice_streams(fac: GpuFactorization, requested: int | None) -> int:
    """Streams for concurrent spectrum slices: `requested`, or by default up to 8, as many as the
    free device memory holds extra factorization objects for (panels, contribution blocks,
    panel workspace and the solve phase's two inverse-panel copies each)."""
    if requested is not None:
        return max(1, int(requested))
    per = 4.0 * fac.info.panel_bytes_double + fac.info.cb_bytes_double
    try:
        free, _ = torch.cuda.mem_get_info()
    except Exception:
        return 1
    return int(max(1, min(8, 1 + (0.5 * free) // max(per, 1.0))))


def solve(A=None, B=None, config: dict | None = None, *, prepared: Prepared | None = None, vectors: str = "host",
          streams: int | None = None):
    """Returns (eigenvalues ndarray, eigenvectors, report dict, config dict). The eigenvectors are an
    (n, m) Fortran-order ndarray (vectors="host") or an (m, n) tensor on the device (float64, complex128
    for a complex pencil) whose
    row j is eigenvector j (vectors="device"). `prepared` (from prepare() on the same matrices)
    skips the input stage and the uploads. `streams`: concurrent streams for the spectrum slices
    of a window target (default: automatic, see slice_streams; 1 = one slice after the other)."""
    if vectors not in ("host", "device"):
        raise InvalidInputError(f"vectors = {vectors!r} is not 'host' or 'device'")
    with _one_blas_thread():
        return _solve(A, B, config, prepared, vectors, streams)


def _one_blas_thread():
    """The host-side dense work of a solve is small (projected problems of order <= a few hundred,
    p x p Gram matrices): multithreaded BLAS only adds synchronization there, so a solve runs it
    on one thread (restored afterwards)."""
    try:
        from threadpoolctl import threadpool_limits
        return threadpool_limits(limits=1, user_api="blas")
    except Exception:  # threadpoolctl missing: leave the BLAS as configured
        import contextlib
        return contextlib.nullcontext()


def _solve(A, B, config, prepared, vectors, streams, fac: GpuFactorization | None = None, counted=()):
    cfg = cfgmod.parse_config(config or {})
    rep = cfgmod.new_report()
    clock = PhaseClock()
    clock.push("setup")
    if prepared is None:
        prepared = prepare(A, B, cfg)
    A, B, bk, pencil = prepared.A, prepared.B, prepared.b_kind, prepared.pencil
    rep["warnings"].extend("input: " + note for note in prepared.notes)
    rep["hermitian_defect_a"] = prepared.defect_a
    if B is not None:
        rep["hermitian_defect_b"] = prepared.defect_b
    n = A.shape[0]
    rep["n"] = n
    rep["nnz_a"] = int(A.nnz)
    rep["nnz_b"] = 0 if B is None else int(B.nnz)
    rep["sigma_requested"] = cfg["target"]["sigma"]
    if cfg["problem"] == "generalized_indefinite":
        raise UnsupportedError("indefinite B (GHIEP) is not supported by this engine")
    # the symbolic analysis (host only, the library releases the GIL) runs on a thread of its own
    # while the device estimates the norms; the factor phase then waits for it
    analysis = {}

    def analyze():
        try:
            plan = autotune.options_for(pencil)
            analysis["plan"] = plan
            analysis["fac"] = GpuFactorization(pencil, **plan)
        except BaseException as e:  # re-raised by the solve
            analysis["error"] = e

    worker = None
    if fac is None:
        worker = threading.Thread(target=analyze, name="hsolve-analysis")
        worker.start()
    try:
        ops = DeviceOps(pencil)
        n2a = estimate_norm2(ops, ops.apply_a, 30, cfg["seed"])
        n2b = 1.0 if B is None else estimate_norm2(ops, ops.apply_b, 30, cfg["seed"])
    except BaseException:
        if worker is not None:  # no factorization object outlives a failed solve
            worker.join()
            if "fac" in analysis:
                analysis["fac"].close()
        raise
    rep["norm2_a"], rep["norm2_b"] = n2a, n2b
    scale = n2a / n2b if n2b > 0 else n2a
    rep["scale"] = scale
    rep["problem"] = "standard" if B is None else "generalized_definite"
    rep["method"] = "shift_invert_lanczos"
    rep["complex_scalars"] = bool(pencil.is_complex)
    clock.pop()
    own_fac = fac is None
    with clock.phase("factor"):
        if worker is not None:
            worker.join()
            if "error" in analysis:
                raise analysis["error"]
            fac, plan = analysis["fac"], analysis["plan"]
            if plan:
                rep["parallel"]["notes"].append(f"launch plan from the autotuner: {plan}")
        elif fac.pencil is not pencil:
            raise InvalidInputError("solve: the given factorization object belongs to another pencil")
    if B is not None and bk == "general":
        with clock.phase("setup"):
            fac.set_null_pivot_detection(True)
            ok = fac.factorize(0.0, 1.0, precision="double")
            inn = fac.inertia()
            if not ok or inn.neg > 0:
                raise UnsupportedError("B is not positive definite (GHIEP): not supported by this engine")
    ctx = Ctx(cfg=cfg, rep=rep, clock=clock, pencil=pencil, ops=ops, fac=fac, n2a=n2a, n2b=n2b, scale=scale,
              counted=list(counted))
    ctx.streams = slice_streams(fac, streams)
    requested = cfg["factor_precision"]
    rep["precision"]["requested"] = requested
    ctx.precision, rep["precision"]["decision"] = choose_precision(fac, cfg, ctx.streams)
    if requested == "auto" and ctx.precision == "single":
        ctx.warn("factor_precision auto: single-precision factors (" + rep["precision"]["decision"] + ")")
    mode = cfg["target"]["mode"]
    result = None
    if ctx.precision == "single":
        t_attempt = time.perf_counter()
        try:
            result = _solve_definite(ctx, mode)
            if cfg["inexact_fault_reject"] and result[0].size:
                result[1][0].backward_error = math.inf
                result = (result[0], result[1], False)
            if not result[2]:
                raise InexactFactorFailure("pairs not accepted or certificate failed")
        except InexactFactorFailure as e:
            rep["precision"]["fallback"] = True
            rep["precision"]["fallback_reason"] = str(e)
            rep["precision"]["attempt_s"] = time.perf_counter() - t_attempt
            ctx.warn(f"single-precision attempt abandoned ({e}); solving again with double factors")
            keep = {k: rep[k] for k in ("warnings", "precision", "hermitian_defect_a", "hermitian_defect_b", "n",
                                        "nnz_a", "nnz_b", "norm2_a", "norm2_b", "scale", "problem", "method",
                                        "sigma_requested", "complex_scalars")}
            rep.clear()
            rep.update(cfgmod.new_report())
            rep.update(keep)
            ctx.precision = "double"
            ctx.held = type(ctx.held)()
            result = None
    if result is None:
        result = _solve_definite(ctx, mode)
    sel, metrics, ok = result
    rep["precision"]["used"] = ctx.precision
    # report
    with clock.phase("verify"):
        lam = np.asarray(sel.lam, dtype=np.float64)
        if vectors == "device":
            vecs = sel.x.detach() if sel.size else torch.zeros((0, n), dtype=ops.dtype, device=ops.dev)
        else:
            vecs = (sel.x.detach().cpu().numpy().T.copy(order="F") if sel.size else
                    np.zeros((n, 0), dtype=ops.np_dtype, order="F"))
    all_ok = ok
    for j, m in enumerate(metrics):
        acc = m.backward_error <= cfg["tol"]
        all_ok = all_ok and acc
        rep["pairs"].append({"lambda_re": float(lam[j]), "lambda_im": 0.0, "abs_residual": m.abs_residual,
                             "rel_residual": m.rel_residual, "backward_error": m.backward_error, "accepted": acc,
                             "subspace_sin_bound": None})
    rep["success"] = bool(all_ok)
    with clock.phase("verify"):
        subspace.attach(ctx, rep, sel, cfg)
    fs = fac.stats
    rep["factor_options"]["precision"] = ctx.precision
    rep["factor_options"]["null_pivot_tol"] = fac.NULL_PIVOT_TOL[0 if ctx.precision == "double" else 1]
    est = fac.estimate()
    rep["factor_stats"].update({
        "analyze_s": fs.analyze_s, "factorize_s": fs.factorize_s, "solve_s": fs.solve_s,
        "factorizations": fs.factorizations, "solves": fs.solves, "rhs_solved": fs.rhs_solved,
        "factor_entries": int(fs.factor_entries), "null_pivots": fs.null_pivots,
        "factor_mb_used": est["incore_mb"],
    })
    rep["factor_stats"]["estimate"].update({"factor_entries": est["factor_entries"], "incore_mb": est["incore_mb"],
                                            "flops": est["flops"]})
    if own_fac:
        fac.close()  # the device storage is freed here, when the solve's work is done
    total = clock.total()
    rep["time_total_s"] = total
    for ph in clock.PHASES:
        rep[f"time_{ph}_s"] = clock.t[ph]
    rep["peak_rss_bytes"] = int(resource.getrusage(resource.RUSAGE_SELF).ru_maxrss) * 1024
    try:
        free, tot = torch.cuda.mem_get_info()
        rep["memory"].update({"budget_gb": tot / 2 ** 30, "source": "device", "limit_gb": tot / 2 ** 30,
                              "limit_source": "hipMemGetInfo"})
    except Exception:
        pass
    rep["_events"] = ctx.events
    return lam, vecs, rep, cfg


def memory_budget_mb(cfg) -> tuple[float, str]:
    """The device memory a solve may hold (MB) and where the number comes from: memory_budget_gb
    when set, else the device memory free now (hipMemGetInfo)."""
    mb = cfg["memory_budget_gb"]
    if mb != "auto" and mb > 0:
        return float(mb) * 1024.0, "memory_budget_gb"
    import torch
    free, _ = torch.cuda.mem_get_info()
    return free / 2 ** 20, "device free memory"


def choose_precision(fac: GpuFactorization, cfg, slots: int = 1) -> tuple[str, str]:
    """Arithmetic of the working factorizations and the reason (report precision.decision).

    "single" and "double" are taken as requested. "auto" follows the reference's rule
    (choose_precision, src/driver/steps.cpp:149-180): double factors when the in-core estimate of
    the concurrently held factorizations (`slots`, one per worker stream) fits the memory budget,
    single when only the single estimate fits, double otherwise (then the allocation reports the
    shortage). Certificate counts are double in every case."""
    requested = cfg["factor_precision"]
    if requested != "auto":
        return requested, f"{requested} requested"
    est_d = fac.estimate("double")["incore_mb"] * max(1, slots)
    est_s = fac.estimate("single")["incore_mb"] * max(1, slots)
    cap, source = memory_budget_mb(cfg)
    if est_d <= cap:
        return "double", (f"double: in-core estimate {est_d:.0f} MB for {slots} factorization(s) fits the "
                          f"{source} ({cap:.0f} MB)")
    if est_s <= cap:
        return "single", (f"single: in-core estimates double {est_d:.0f} MB, single {est_s:.0f} MB, "
                          f"{source} {cap:.0f} MB")
    return "double", (f"double: in-core estimates double {est_d:.0f} MB, single {est_s:.0f} MB, both above the "
                      f"{source} ({cap:.0f} MB)")


def _solve_definite(ctx: Ctx, mode: str):
    if mode == "nearest":
        from .nearest import solve_nearest
        return solve_nearest(ctx)
    from .windows import solve_window
    return solve_window(ctx, ctx.streams)
