"""Host logic of the serving loop (hsolve_amd.serve): admission order and memory accounting, the
pattern key of the analysis cache, the memory prediction. No device work."""
from types import SimpleNamespace

import numpy as np
import scipy.sparse as sp

from hsolve_amd import config as cfgmod
from hsolve_amd.factor import Pencil
from hsolve_amd.serve import pattern_key, pick_admissible, predict_bytes


def test_pick_admits_earliest_that_fits():
    w = [(100, 0.0), (10, 0.1), (5, 0.2)]
    assert pick_admissible(w, 200, 0.3, 2.0) == 0
    assert pick_admissible(w, 50, 0.3, 2.0) == 1          # the first does not fit: a later one may go
    assert pick_admissible(w, 4, 0.3, 2.0) is None
    assert pick_admissible(w, 50, 5.0, 2.0) is None       # the first waited too long: nobody overtakes it
    assert pick_admissible([], 50, 0.0, 2.0) is None


def test_admission_never_exceeds_the_budget():
    # replay: requests of random sizes arrive; admitted ones run for a while; the sum of admitted
    # predictions never exceeds the budget and every request is eventually admitted in a
    # starvation-free order
    rng = np.random.default_rng(3)
    budget = 100
    sizes = rng.integers(1, 90, 60)
    arrivals = np.cumsum(rng.exponential(0.2, 60))
    waiting, running, done = [], [], []
    t, nxt = 0.0, 0
    while len(done) < 60:
        while nxt < 60 and arrivals[nxt] <= t:
            waiting.append((int(sizes[nxt]), float(arrivals[nxt]), nxt))
            nxt += 1
        running = [(end, need, rid) for end, need, rid in running if end > t]
        in_use = sum(need for _, need, _ in running)
        while True:
            i = pick_admissible([(need, arr) for need, arr, _ in waiting], budget - in_use, t, 2.0)
            if i is None:
                break
            need, arr, rid = waiting.pop(i)
            in_use += need
            assert in_use <= budget
            running.append((t + 0.5, need, rid))
            done.append(rid)
        t += 0.05
    assert sorted(done) == list(range(60))


def test_pattern_key_separates_patterns_and_ignores_values():
    A = sp.random(50, 50, density=0.1, random_state=1, format="csr")
    A = (A + A.T + sp.identity(50)).tocsr()
    k1 = pattern_key(Pencil(A, None))
    k2 = pattern_key(Pencil(A * 3.0, None))
    B = A.copy()
    B[0, 49] = 1.0
    B[49, 0] = 1.0
    k3 = pattern_key(Pencil(B.tocsr(), None))
    assert k1 == k2 and k1 != k3


def test_predicted_bytes_scale_with_precision_and_k():
    info = SimpleNamespace(panel_bytes_double=8e8, cb_bytes_double=4e9)
    fac = SimpleNamespace(info=info)
    cfg = cfgmod.parse_config({"target": {"mode": "nearest", "sigma": 1.0, "k": 10}})
    d = predict_bytes(fac, cfg, 100000, False)
    cfg_s = cfgmod.parse_config({"target": {"mode": "nearest", "sigma": 1.0, "k": 10}, "factor_precision": "single"})
    s_ = predict_bytes(fac, cfg_s, 100000, False)
    cfg_k = cfgmod.parse_config({"target": {"mode": "nearest", "sigma": 1.0, "k": 40}})
    assert d > s_ > 0 and predict_bytes(fac, cfg_k, 100000, False) > d
    assert d >= 4 * 8e8 + 4e9
