"""Keyword rules of hsolve_amd.api.solve (the reference's contract,
python/hsolve/__init__.py:241-278, 347-419): which keywords override config.json, how the target
mode is inferred, which values are rejected and with which error class. Host only: every case
fails or is decided before any device work."""
import json

import numpy as np
import pytest
import scipy.sparse as sp

from hsolve_amd import api
from hsolve_amd.errors import HsolveIOError, InvalidInputError


def test_mode_inference():
    assert api.merge_config(None, frozenset({"sigma"}), sigma=3.0)["target"]["mode"] == "nearest"
    assert api.merge_config(None, frozenset({"lower", "upper"}), lower=1.0, upper=2.0)["target"]["mode"] == "interval"
    with pytest.raises(InvalidInputError):
        api.merge_config(None, frozenset({"sigma", "lower"}), sigma=1.0, lower=0.5)
    # an explicit mode settles the conflict
    cfg = api.merge_config(None, frozenset({"sigma", "lower", "upper", "mode"}), sigma=1.0, lower=0.5, upper=2.0,
                           mode="interval")
    assert cfg["target"]["mode"] == "interval" and cfg["target"]["sigma"] == 1.0
    with pytest.raises(InvalidInputError):
        api.merge_config(None, frozenset({"mode"}), mode="biggest")
    with pytest.raises(TypeError):
        api.merge_config(None, frozenset({"mode"}), mode=3)


def test_config_keeps_its_values_unless_a_keyword_is_passed():
    base = {"tol": 1e-8, "threads": 3, "target": {"mode": "nearest", "sigma": 2.0, "k": 9}}
    cfg = api.merge_config(base)
    assert cfg["tol"] == 1e-8 and cfg["threads"] == 3 and cfg["target"]["k"] == 9
    # passing the default value still counts as passing it
    cfg = api.merge_config(base, frozenset({"tol", "k"}), tol=1e-10, k=6)
    assert cfg["tol"] == 1e-10 and cfg["target"]["k"] == 6 and cfg["threads"] == 3
    assert base["target"]["k"] == 9  # the caller's dict is not modified


@pytest.mark.parametrize("bad", [True, np.bool_(False), "0.5", 1 + 2j, None, [1.0]])
def test_real_keywords_reject_non_reals(bad):
    with pytest.raises(TypeError):
        api.as_real(bad, "sigma")


@pytest.mark.parametrize("good,val", [(2, 2.0), (np.float32(1.5), 1.5), (np.array(2.5), 2.5), (np.int64(3), 3.0)])
def test_real_keywords_accept_numbers(good, val):
    assert api.as_real(good, "sigma") == val


@pytest.mark.parametrize("bad", [True, "2", 2.0, np.float64(2.0)])
def test_integer_keywords_reject_non_integers(bad):
    with pytest.raises(TypeError):
        api.as_int(bad, "k")


def test_integer_keywords_accept_index_types():
    assert api.as_int(np.int32(7), "k") == 7 and api.as_int(5, "k") == 5


def test_config_sources(tmp_path):
    assert api.load_config(None) == {}
    assert api.load_config({"tol": np.float64(1e-9)}) == {"tol": 1e-9}
    good = tmp_path / "c.json"
    good.write_text(json.dumps({"tol": 1e-9}))
    assert api.load_config(str(good)) == {"tol": 1e-9}
    assert api.load_config(good) == {"tol": 1e-9}
    with pytest.raises(HsolveIOError):
        api.load_config(str(tmp_path / "missing.json"))
    bad = tmp_path / "bad.json"
    bad.write_text("{not json")
    with pytest.raises(InvalidInputError):
        api.load_config(bad)
    arr = tmp_path / "arr.json"
    arr.write_text("[1, 2]")
    with pytest.raises(InvalidInputError):
        api.load_config(arr)
    with pytest.raises(TypeError):
        api.load_config(5)


def test_solve_rejects_bad_arguments_before_any_device_work():
    A = sp.identity(4, format="csr")
    with pytest.raises(InvalidInputError):
        api.solve(A, sigma=1.0, storage="diagonal")
    with pytest.raises(TypeError):
        api.solve(A, sigma=1.0, k=True)
    with pytest.raises(TypeError):
        api.solve(A, sigma="1.0")
    with pytest.raises(TypeError):
        api.solve(A, sigma=1.0, verbose=1.5)
    with pytest.raises(TypeError):
        class Op:
            shape = (4, 4)

            def matvec(self, x):
                return x
        api.solve(Op(), sigma=1.0)
