"""Interval planning of the window solver on the CPU (src/driver/definite.cpp:1192-1229): plan()
against an exact counting function, and the speculative counts that run ahead of it. The counts
come from a synthetic spectrum shaped like box48's (a null space of 37,835 zeros, then the Weyl
law N(s) = c s^{3/2} of a 3D cavity with 822 eigenvalues below 40): the density grows by a
factor 1.6 over the window [40, b] of 500 eigenvalues."""
import math
import types

import numpy as np
import pytest

from hsolve_amd.windows import CountPoint, Slice, WindowSolver


class _Ctx:
    def __init__(self):
        self.events = []

    def event(self, name, **kw):
        self.events.append((name, kw))


def _solver(spectrum):
    """A WindowSolver whose counts are exact counts of `spectrum` (no device)."""
    w = WindowSolver.__new__(WindowSolver)
    w._main = _Ctx()
    w._tls = types.SimpleNamespace()
    import threading
    w._lock = threading.Lock()
    w.streams = 1
    w._slot_list = []
    w.n = len(spectrum)
    w.scale = float(np.max(np.abs(spectrum)))
    w.known = []
    lam = np.sort(np.asarray(spectrum, dtype=float))
    w.calls = []

    def count(self, s, outward):
        for p in self.known:
            if p.s == s:
                return p
        self.calls.append(s)
        p = CountPoint(s, int(np.searchsorted(lam, s, side="left")))
        self.remember(p)
        return p

    w.count = types.MethodType(count, w)
    return w


def _cavity_spectrum(null, below_40, count, seed=0):
    """A null space of `null` zeros and N(s) = c s^1.5 eigenvalues below s (the Weyl law of a 3D
    cavity) with `below_40` of them below 40, jittered to be distinct, `count` of them above 40."""
    c = below_40 / 40.0 ** 1.5
    k = np.arange(1, below_40 + count + 1)
    lam = (k / c) ** (2.0 / 3.0)
    rng = np.random.default_rng(seed)
    return np.concatenate([np.zeros(null), lam + 1e-9 * rng.uniform(-1, 1, lam.size)])


def test_plan_slices_respect_the_size_limits():
    lam = _cavity_spectrum(37835, 822, 3000)
    w = _solver(lam)
    a, b = w.count(40.0, -1.0), w.count(54.0, +1.0)
    sl = w.plan(Slice(a, b), 96, 120)
    assert sl[0].lo.s == 40.0 and sl[-1].hi.s == 54.0
    assert sum(s.count() for s in sl) == b.nu - a.nu
    assert all(0 < s.count() <= 120 for s in sl)
    assert all(s.count() >= 48 for s in sl[:-1])  # plan's accepted range: [step/2, c_max]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_speculative_counts_leave_plan_without_serial_counts(seed):
    # the interpolated cuts undershoot where the density grows; with the tail midpoint among the
    # speculative points, plan() finds every cut among the points counted ahead
    lam = _cavity_spectrum(37835, 822, 600, seed)
    w = _solver(lam)
    a = w.count(40.0, -1.0)
    hi = float(np.sort(lam)[a.nu + 500 - 1]) + 1e-7  # 500 eigenvalues in the window, as box48
    b = w.count(hi, +1.0)
    win = Slice(a, b)
    spec = w.speculative_cuts(win, 96, 120)
    for s in spec:
        w.count(s, +1.0)
    n_before = len(w.calls)
    sl = w.plan(win, 96, 120)
    assert len(w.calls) == n_before, "plan() needed counts beyond the speculative ones"
    assert sum(s.count() for s in sl) == 500


def test_speculative_tail_point_is_the_midpoint_after_the_last_cut():
    w = _solver(np.linspace(1.0, 2.0, 10))
    a, b = CountPoint(40.0, 1000), CountPoint(54.0, 1500)
    spec = w.speculative_cuts(Slice(a, b), 96, 120)
    # interpolated cuts at 96, 192, 288, 384 eigenvalues, 116 left: then its midpoint
    assert len(spec) == 5
    assert spec[3] == pytest.approx(40.0 + 14.0 * 384 / 500)
    assert spec[4] == pytest.approx(0.5 * (spec[3] + 54.0))
    # 424 in the window: 96, 192, 288, then half of the 136 left (356), and the midpoint after it
    spec = w.speculative_cuts(Slice(a, CountPoint(54.0, 1424)), 96, 120)
    assert len(spec) == 5 and spec[3] == pytest.approx(40.0 + 14.0 * 356 / 424)
    # a window within c_max needs no cut at all
    assert w.speculative_cuts(Slice(a, CountPoint(54.0, 1100)), 96, 120) == []


def _dry_aware(w):
    """The exact-count solver of _solver with the dry-run rule of WindowSolver.count: an uncounted
    shift raises during plan_from_known."""
    from hsolve_amd.windows import _NeedsCount
    inner = w.count

    def count(self, s, outward):
        if getattr(self, "_dry", False) and all(p.s != s for p in self.known):
            raise _NeedsCount(s)
        return inner(s, outward)

    w.count = types.MethodType(count, w)
    return w


@pytest.mark.parametrize("streams,k", [(1, 0), (2, 0), (5, 3), (8, 5), (16, 5)])
def test_blind_points_are_equally_spaced_inside_the_window(streams, k):
    w = _solver(np.linspace(1.0, 2.0, 10))
    w.streams = streams
    pts = w.blind_points(40.0, 54.0)
    assert len(pts) == k
    assert np.allclose(pts, 40.0 + 14.0 * np.arange(1, k + 1) / (k + 1))
    assert w.blind_points(40.0, math.inf) == [] and w.blind_points(54.0, 40.0) == []


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_blind_points_supply_every_cut_of_a_box48_shaped_window(seed):
    # 500 eigenvalues over a window where the density grows 1.6-fold: five equally spaced points
    # leave intervals of 77-89 eigenvalues, all within plan()'s limits [48, 120]
    lam = _cavity_spectrum(37835, 822, 600, seed)
    w = _dry_aware(_solver(lam))
    w.streams = 8
    a = w.count(40.0, -1.0)
    hi = float(np.sort(lam)[a.nu + 500 - 1]) + 1e-7
    b = w.count(hi, +1.0)
    for s in w.blind_points(40.0, hi):
        w.count(s, +1.0)
    known = [(p.s, p.nu) for p in w.known]
    win = Slice(a, b)
    assert w.plan_from_known(win, 96, 120)
    assert [(p.s, p.nu) for p in w.known] == known  # the dry run leaves the known points alone
    n_before = len(w.calls)
    sl = w.plan(win, 96, 120)
    assert len(w.calls) == n_before
    assert sum(s.count() for s in sl) == 500 and all(48 <= s.count() <= 120 for s in sl)


def test_dry_run_reports_a_window_the_blind_points_do_not_cover():
    # about 2,400 eigenvalues in [40, 100]: intervals of ~350 between the blind points, so plan()
    # needs more counts; the dry run says so, counts nothing and leaves the known points alone
    lam = _cavity_spectrum(37835, 822, 3000)
    w = _dry_aware(_solver(lam))
    w.streams = 8
    a, b = w.count(40.0, -1.0), w.count(100.0, +1.0)
    assert b.nu - a.nu > 2000
    for s in w.blind_points(40.0, 100.0):
        w.count(s, +1.0)
    known = [(p.s, p.nu) for p in w.known]
    n_before = len(w.calls)
    assert not w.plan_from_known(Slice(a, b), 96, 120)
    assert len(w.calls) == n_before and [(p.s, p.nu) for p in w.known] == known
    assert not w._dry
