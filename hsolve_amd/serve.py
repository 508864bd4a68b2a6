"""Serving loop: solves of different pencils, n, k and sigma submitted as they arrive, run
concurrently on one device with admission by predicted device memory.

The reference serves one solve per process call (its Python binding serializes solves with a
process-wide mutex, python/src/module.cpp:14-16, and keeps no state between calls;
results/p39_measurements.md:7). What this module adds, on top of the ordinary solve
(hsolve_amd.solver, the same certificate and report per request):

  * resident state: symbolic analyses are kept in an LRU cache keyed by a hash of the factorized
    pattern and value maps, so a request whose pattern was seen before skips ordering and
    symbolic analysis; a request's numeric factorization is a clone of the cached analysis (its
    own factor storage). Krylov bases and vector blocks come from torch's caching allocator on
    the worker's stream, whose freed blocks serve the next request on that stream; factor blocks
    come from the library's block cache. Pencils stay on the device while their request runs.
  * admission (continuous batching): `slots` worker threads, each with a stream of its own, take
    requests in arrival order. A request is admitted when a slot is free and its predicted device
    memory fits what the admitted requests leave of the budget (predict_bytes: factor panels and
    contribution blocks, the solve phase's two inverse-panel copies and work arrays, the Krylov
    basis); requests that arrive while others run are admitted as slots and memory free up, and
    their kernels overlap on the device. A later request that fits may go before an earlier one
    that does not, unless the earlier one has waited longer than `max_wait_s` (then nothing is
    admitted ahead of it). A request larger than the whole budget fails with
    OutOfMemoryBudgetError.
  * every result is that of the same request solved alone: each request runs with its own
    factorization, Krylov scratch and report, one spectrum slice at a time (streams = 1), and the
    kernels' reductions have fixed orders, so concurrency changes only the timing.

Latency of a request = completion time - arrival time (queueing, preparation and solve).
"""
from __future__ import annotations

import hashlib
import math
import threading
import time
from collections import OrderedDict, deque
from concurrent.futures import Future
from dataclasses import dataclass

import numpy as np
import torch

from . import autotune
from . import config as cfgmod
from . import streams as stream_pool
from .errors import OutOfMemoryBudgetError
from .factor import GpuFactorization, Pencil
from .solver import _one_blas_thread, _solve, prepare


def pattern_key(p: Pencil) -> str:
    """Hash of what one analysis serves: order, scalar kind, factorized pattern and value maps."""
    h = hashlib.blake2b(digest_size=16)
    h.update(np.array([p.f_n, int(p.is_complex)], dtype=np.int64).tobytes())
    for a in (p.f_rowptr, p.f_col, p.f_a_index, p.f_b_index):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def predict_bytes(fac: GpuFactorization, cfg: dict, n: int, is_complex: bool) -> int:
    """Device memory one request holds at its peak (one factorization at a time, streams = 1):
    panels, contribution blocks and panel workspace of the factorization, the solve phase's two
    inverse-panel copies, and the Krylov basis of (d + ncv + p) columns plus a few blocks."""
    w = 8 if cfg["factor_precision"] != "single" else 4
    info = fac.info
    factor = (info.panel_bytes_double * 4.0 + info.cb_bytes_double) * w / 8.0
    t = cfg["target"]
    k = int(t["k"]) if t["mode"] != "interval" else 96  # an interval's slices hold about 96 pairs each
    cols = 3 * k + 64
    basis = float(n) * cols * (16 if is_complex else 8)
    return int(factor + 2.0 * basis)


def pick_admissible(waiting, free: int, now: float, max_wait_s: float) -> int | None:
    """Index of the request to admit next among `waiting` ((predicted bytes, arrival) in arrival
    order) with `free` bytes of the budget left: the earliest that fits, except that nothing goes
    ahead of the earliest once it has waited longer than max_wait_s. None: admit nothing now."""
    for i, (need, arrival) in enumerate(waiting):
        if need <= free:
            return i
        if i == 0 and now - arrival > max_wait_s:
            return None
    return None


@dataclass
class ServeResult:
    rid: int
    eigenvalues: np.ndarray
    eigenvectors: object
    report: dict
    config: dict
    arrival: float        # seconds on the server clock
    admitted: float
    finished: float
    predicted_bytes: int
    analysis_cached: bool

    @property
    def latency_s(self) -> float:
        return self.finished - self.arrival


@dataclass
class _Entry:
    rid: int
    A: object
    B: object
    config: dict
    arrival: float
    future: Future
    prepared: object = None
    head: GpuFactorization | None = None
    key: str = ""
    cached: bool = False
    need: int = 0


class Server:
    """A running serving loop on the current device (use as a context manager, or call close())."""

    def __init__(self, slots: int = 4, memory_budget_gb: float | None = None, vectors: str = "host",
                 analysis_cache: int = 32, max_wait_s: float = 2.0, prepare_workers: int = 2):
        self.slots = max(1, int(slots))
        if memory_budget_gb is None:
            free, _ = torch.cuda.mem_get_info()
            self.budget = int(0.8 * free)
        else:
            self.budget = int(memory_budget_gb * 2 ** 30)
        self.vectors = vectors
        self.max_wait_s = float(max_wait_s)
        self._cache_cap = max(1, int(analysis_cache))
        self._analyses: OrderedDict[str, GpuFactorization] = OrderedDict()
        self._analysis_locks: dict[str, threading.Lock] = {}
        self._pins: dict[str, int] = {}  # requests holding a cached analysis that is not cloned yet
        self._lock = threading.Lock()
        self._cv = threading.Condition(self._lock)
        self._incoming: deque[_Entry] = deque()
        self._waiting: list[_Entry] = []       # prepared, in arrival order
        self._in_use = 0
        self._running = 0
        self._closed = False
        self._next_id = 0
        self.t0 = time.perf_counter()
        self.stats = {"admitted": 0, "analyses": 0, "analysis_hits": 0, "peak_predicted_bytes": 0, "peak_running": 0}
        self._streams = stream_pool.checkout(self.slots)
        self._device = torch.cuda.current_device()
        self._threads = [threading.Thread(target=self._prepare_loop, name=f"hsolve-serve-prep-{i}", daemon=True)
                         for i in range(max(1, prepare_workers))]
        self._threads += [threading.Thread(target=self._worker, args=(w,), name=f"hsolve-serve-{w}", daemon=True)
                          for w in range(self.slots)]
        for t in self._threads:
            t.start()

    # ------------------------------------------------------------------ client side
    def now(self) -> float:
        return time.perf_counter() - self.t0

    def submit(self, A, B=None, config: dict | None = None) -> Future:
        """Queue a solve of (A, B) with `config` (as solver.solve); the future resolves to a
        ServeResult, or raises the solve's error."""
        fut: Future = Future()
        with self._cv:
            if self._closed:
                raise RuntimeError("submit on a closed server")
            e = _Entry(self._next_id, A, B, dict(config or {}), self.now(), fut)
            self._next_id += 1
            self._incoming.append(e)
            self._cv.notify_all()
        return fut

    def close(self) -> None:
        with self._cv:
            self._closed = True
            self._cv.notify_all()
        for t in self._threads:
            t.join()
        stream_pool.checkin(self._streams)
        for f in self._analyses.values():
            f.close()
        self._analyses.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ------------------------------------------------------------------ preparation
    def _analysis_for(self, pencil: Pencil) -> tuple[str, GpuFactorization, bool]:
        """The cached analysis of the pencil's pattern (computed on a miss), pinned against
        eviction until the request has cloned it (_unpin)."""
        key = pattern_key(pencil)
        with self._lock:
            lk = self._analysis_locks.setdefault(key, threading.Lock())
        with lk:  # one analysis per pattern even when its requests are prepared concurrently
            with self._lock:
                head = self._analyses.get(key)
                if head is not None:
                    self._analyses.move_to_end(key)
                    self._pins[key] = self._pins.get(key, 0) + 1
                    self.stats["analysis_hits"] += 1
                    return key, head, True
            head = GpuFactorization(pencil, **autotune.options_for(pencil))
            with self._lock:
                self._analyses[key] = head
                self._pins[key] = self._pins.get(key, 0) + 1
                self.stats["analyses"] += 1
                self._evict()
            return key, head, False

    def _evict(self) -> None:
        """Drops least recently used analyses beyond the cache size (lock held); pinned ones stay.
        Clones made from a dropped analysis keep its shared symbolic data alive."""
        for key in list(self._analyses):
            if len(self._analyses) <= self._cache_cap:
                break
            if self._pins.get(key, 0) == 0:
                self._analyses.pop(key).close()

    def _unpin(self, key: str) -> None:
        with self._lock:
            self._pins[key] -= 1
            if self._pins[key] == 0:
                del self._pins[key]
            self._evict()

    def _prepare_loop(self) -> None:
        torch.cuda.set_device(self._device)
        while True:
            with self._cv:
                while not self._incoming and not self._closed:
                    self._cv.wait()
                if not self._incoming:
                    return
                e = self._incoming.popleft()
            try:
                cfg = cfgmod.parse_config(e.config)
                e.config = cfg
                e.prepared = prepare(e.A, e.B, cfg)
                e.key, e.head, e.cached = self._analysis_for(e.prepared.pencil)
                e.need = predict_bytes(e.head, cfg, e.prepared.pencil.n, e.prepared.pencil.is_complex)
                if e.need > self.budget:
                    raise OutOfMemoryBudgetError(
                        f"request {e.rid}: predicted {e.need / 2 ** 30:.2f} GiB of device memory exceeds the "
                        f"serving budget of {self.budget / 2 ** 30:.2f} GiB")
            except BaseException as err:  # reported through the future
                if e.key:
                    self._unpin(e.key)
                e.future.set_exception(err)
                continue
            with self._cv:
                self._waiting.append(e)
                self._waiting.sort(key=lambda x: x.rid)  # arrival order whatever the preparation order
                self._cv.notify_all()

    # ------------------------------------------------------------------ admission and execution
    def _pick(self) -> _Entry | None:
        """The next admissible request (lock held), see pick_admissible."""
        i = pick_admissible([(e.need, e.arrival) for e in self._waiting], self.budget - self._in_use, self.now(),
                            self.max_wait_s)
        return None if i is None else self._waiting.pop(i)

    def _worker(self, w: int) -> None:
        torch.cuda.set_device(self._device)
        stream = self._streams[w]
        while True:
            with self._cv:
                while True:
                    e = self._pick()
                    if e is not None:
                        break
                    if self._closed and not self._incoming and not self._waiting and self._preparing_done():
                        return
                    self._cv.wait(timeout=0.05)
                self._in_use += e.need
                self._running += 1
                self.stats["admitted"] += 1
                self.stats["peak_predicted_bytes"] = max(self.stats["peak_predicted_bytes"], self._in_use)
                self.stats["peak_running"] = max(self.stats["peak_running"], self._running)
            admitted = self.now()
            fac = None
            try:
                with torch.cuda.stream(stream), _one_blas_thread():
                    try:
                        fac = e.head.clone_for(e.prepared.pencil)
                    finally:
                        self._unpin(e.key)
                    lam, vecs, rep, cfg_out = _solve(None, None, e.config, e.prepared, self.vectors, 1, fac)
                    stream.synchronize()
                rep.pop("_events", None)
                rep["parallel"]["notes"].append(
                    f"served: slot {w} of {self.slots}, analysis {'reused' if e.cached else 'computed'}, "
                    f"predicted device memory {e.need / 2 ** 20:.0f} MiB")
                res = ServeResult(e.rid, lam, vecs, rep, cfg_out, e.arrival, admitted, self.now(), e.need, e.cached)
                e.future.set_result(res)
            except BaseException as err:
                e.future.set_exception(err)
            finally:
                if fac is not None:
                    fac.close()
                e.prepared = None
                with self._cv:
                    self._in_use -= e.need
                    self._running -= 1
                    self._cv.notify_all()

    def _preparing_done(self) -> bool:
        return all(not t.is_alive() for t in self._threads if t.name.startswith("hsolve-serve-prep"))


def run_stream(requests, slots: int = 4, memory_budget_gb: float | None = None, vectors: str = "host",
               time_scale: float = 1.0) -> tuple[list, dict]:
    """Replays `requests` (dicts with A, B, config and arrival in seconds, sorted by arrival) against a
    Server: each is submitted at its arrival time (scaled by time_scale). Returns the results in
    request order (ServeResult, or the exception) and a summary: p50 / p99 latency, throughput in
    requests and pairs per second over the makespan, and the server's admission statistics."""
    with Server(slots=slots, memory_budget_gb=memory_budget_gb, vectors=vectors) as srv:
        futs = []
        start = srv.now()
        for r in requests:
            delay = start + time_scale * float(r["arrival"]) - srv.now()
            if delay > 0:
                time.sleep(delay)
            futs.append(srv.submit(r["A"], r.get("B"), r["config"]))
        out = []
        for f in futs:
            try:
                out.append(f.result())
            except BaseException as err:  # noqa: BLE001 (reported per request)
                out.append(err)
        end = srv.now()
        stats = dict(srv.stats)
        budget = srv.budget
    ok = [r for r in out if isinstance(r, ServeResult)]
    lat = sorted(r.latency_s for r in ok)

    def pct(q):
        if not lat:
            return math.nan
        k = (len(lat) - 1) * q / 100.0
        f, c = math.floor(k), math.ceil(k)
        return lat[f] + (lat[c] - lat[f]) * (k - f)
    span = max(1e-9, end - start)
    summary = {"requests": len(requests), "completed": len(ok), "failed": len(out) - len(ok),
               "makespan_s": span, "latency_p50_ms": 1e3 * pct(50), "latency_p99_ms": 1e3 * pct(99),
               "throughput_req_s": len(ok) / span, "pairs": int(sum(r.eigenvalues.size for r in ok)),
               "pairs_per_s": sum(r.eigenvalues.size for r in ok) / span, "slots": slots,
               "budget_gb": budget / 2 ** 30, **stats}
    return out, summary
