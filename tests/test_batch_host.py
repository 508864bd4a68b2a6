"""Device memory of a batch on the host side (hsolve_amd.batch.solve_batch): a pencil's numeric
factorization is a clone made when a worker takes the pencil and closed as soon as its solve ends,
so at most `workers` factorizations are held at a time whatever the number of pencils, and none
is left behind when a solve fails (the error is re-raised in pencil order). The GPU pieces are
replaced by stand-ins that count the live factorization objects; the batch logic runs as is."""
from __future__ import annotations

import contextlib
import threading
import time
from types import SimpleNamespace

import numpy as np
import pytest

from hsolve_amd import batch


class _Counter:
    def __init__(self):
        self.lock = threading.Lock()
        self.live = 0
        self.peak = 0
        self.made = 0
        self.closed_heads = 0


class _FakeFac:
    counter: _Counter = None

    def __init__(self, pencil, head=True, **opts):
        self.pencil = pencil
        self.head = head
        self.open = True
        if not head:
            with self.counter.lock:
                self.counter.live += 1
                self.counter.made += 1
                self.counter.peak = max(self.counter.peak, self.counter.live)

    def clone_for(self, pencil):
        return _FakeFac(pencil, head=False)

    def close(self):
        if not self.open:
            return
        self.open = False
        with self.counter.lock:
            if self.head:
                self.counter.closed_heads += 1
            else:
                self.counter.live -= 1


class _FakeStream:
    def synchronize(self):
        pass


@pytest.fixture
def stubs(monkeypatch):
    counter = _Counter()
    _FakeFac.counter = counter
    monkeypatch.setattr(batch, "GpuFactorization", _FakeFac)
    monkeypatch.setattr(batch, "same_pattern", lambda a, b: True)
    monkeypatch.setattr(batch.autotune, "options_for", lambda p: {})
    monkeypatch.setattr(batch.stream_pool, "checkout", lambda k: [_FakeStream() for _ in range(k)])
    monkeypatch.setattr(batch.stream_pool, "checkin", lambda s: None)
    monkeypatch.setattr(batch.torch.cuda, "stream", lambda s: contextlib.nullcontext())
    monkeypatch.setattr(batch.torch.cuda, "synchronize", lambda: None)
    return counter


def _pencils(n):
    return [SimpleNamespace(pencil=SimpleNamespace(name=f"p{i}")) for i in range(n)]


def _fake_solve(fail_at=None, counter=None, workers=None):
    def solve(A, B, cfg, prepared, vectors, streams, fac, counted=()):
        assert fac is not None and not fac.head and fac.open
        with counter.lock:
            assert counter.live <= workers
        time.sleep(0.002)  # let the workers overlap
        if fail_at is not None and prepared.pencil.name == f"p{fail_at}":
            raise RuntimeError("planted failure")
        return (np.zeros(0), None, {"parallel": {"notes": []}}, cfg)
    return solve


@pytest.mark.parametrize("workers", [1, 3, 8])
def test_at_most_workers_factorizations_held(stubs, monkeypatch, workers):
    monkeypatch.setattr(batch, "_solve", _fake_solve(counter=stubs, workers=workers))
    out = batch.solve_batch(_pencils(20), {"target": {"mode": "nearest", "sigma": 1.0, "k": 2}}, workers=workers)
    assert len(out) == 20
    assert stubs.made == 20 and stubs.live == 0
    assert 1 <= stubs.peak <= workers
    assert stubs.closed_heads == 1  # the one shared analysis is closed too
    assert f"at most {stubs.peak} factorizations held at once" in out[0][2]["parallel"]["notes"][-1]


def test_failed_solve_releases_everything_and_reraises(stubs, monkeypatch):
    monkeypatch.setattr(batch, "_solve", _fake_solve(fail_at=5, counter=stubs, workers=4))
    with pytest.raises(RuntimeError, match="planted failure"):
        batch.solve_batch(_pencils(20), {"target": {"mode": "nearest", "sigma": 1.0, "k": 2}}, workers=4)
    assert stubs.live == 0
    assert stubs.closed_heads == 1
