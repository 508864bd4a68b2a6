"""The autotuner's cache (hsolve_amd/autotune.py): tuned options are stored per pattern, returned
for that pattern only, ignored under HSG_AUTOTUNE=0, and nothing is hashed when no cache exists.
Host only (the measurement itself runs in tests/test_autotune_gpu.py)."""
import numpy as np

from hsolve_amd import autotune
from hsolve_amd.factor import Pencil
from hsolve_amd.gen import yee


def test_cache_round_trip(tmp_path, monkeypatch):
    monkeypatch.setenv("HSG_TUNE_CACHE", str(tmp_path / "t.json"))
    monkeypatch.setattr(autotune, "_memory", {})
    A, _, _ = yee.cube_cavity(5)
    P, Q = Pencil(A, None), Pencil(yee.cube_cavity(6)[0], None)
    assert autotune.options_for(P) == {} and not hasattr(P, "_tune_key")  # no cache: no hashing
    autotune._store(autotune.pattern_key(P), {"options": {"leaf_size": 32, "panel_width": 32, "small_front_max": 64}})
    monkeypatch.setattr(autotune, "_memory", {})  # a new process reads the file
    assert autotune.options_for(Pencil(A * 2.0, None)) == {"leaf_size": 32, "panel_width": 32, "small_front_max": 64}
    assert autotune.options_for(Q) == {}
    monkeypatch.setenv("HSG_AUTOTUNE", "0")
    assert autotune.options_for(P) == {}
