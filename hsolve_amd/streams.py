"""Worker streams kept for the life of the process.

Concurrent spectrum slices, counts and batched pencils each run on a stream of their own. torch's
caching allocator keeps freed device blocks per stream, so a solve that took fresh streams every
time could not reuse the blocks of the previous solve (Krylov bases, vector blocks) and would map
new device memory instead, several GB per solve. Solves therefore check streams out of this pool
and return them when they end: from the second solve on, their allocations are served from the
cache. A stream is held by one solve at a time (concurrent solves get distinct streams); every
stream has the default priority.
"""
from __future__ import annotations

import threading

import torch

_lock = threading.Lock()
_idle: dict[int, list] = {}  # device index -> streams not held by a solve


def checkout(k: int) -> list:
    """k streams of the current device, reused from earlier solves where possible."""
    dev = torch.cuda.current_device()
    with _lock:
        pool = _idle.setdefault(dev, [])
        out = [pool.pop() for _ in range(min(k, len(pool)))]
    while len(out) < k:
        out.append(torch.cuda.Stream())
    return out


def checkin(streams) -> None:
    """Returns streams to the pool (after the work queued on them, or with it still queued: the
    next holder's work is ordered after it on the same stream)."""
    if not streams:
        return
    dev = streams[0].device.index
    with _lock:
        _idle.setdefault(dev, []).extend(reversed(list(streams)))
