"""File input: hsolve_amd.io.load_matrix against the reference's own IO fixtures and their expected
results (tests/data/io/*, expected.json, copied as data into tests/golden/io): for every case the
source format, restored storage, note tags in order, order, nnz, dtype, the dense values and the
Hermitian defect, or the error class of a rejected file. Also the npy writers against the golden
vectors and matrices (byte-identical to np.save)."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

from hsolve_amd import io as hio
from hsolve_amd.errors import HsolveIOError, InvalidInputError, NotHermitianError, UnsupportedError

ROOT = Path(__file__).parent / "golden" / "io"
EXPECTED = json.loads((ROOT / "expected.json").read_text())
ERRORS = {"io": HsolveIOError, "invalid_input": InvalidInputError, "not_hermitian": NotHermitianError,
          "unsupported": UnsupportedError}


def _options(opt: dict) -> dict:
    out = {}
    for k, v in opt.items():
        if k in ("storage", "hermitian_tol", "check_hermitian", "drop_zeros"):
            out[k] = v
        else:
            raise AssertionError(f"unknown load option {k}")
    return out


@pytest.mark.parametrize("case", EXPECTED["load"], ids=[c["name"] for c in EXPECTED["load"]])
def test_load_matches_reference(case):
    path = ROOT / case["file"]
    opts = _options(case.get("options", {}))
    if case.get("error"):
        with pytest.raises(ERRORS[case["error"]]):
            hio.load_matrix(path, **opts)
        return
    got = hio.load_matrix(path, **opts)
    assert got.source_format == case["source_format"]
    assert got.storage == case["storage"]
    assert [n.split(":")[0] for n in got.notes] == case["notes"]
    M = got.matrix
    assert M.shape == (case["n"], case["n"])
    assert M.nnz == case["nnz"]
    assert str(M.dtype) == case["dtype"]
    dense = M.toarray()
    np.testing.assert_array_equal(dense.real, np.asarray(case["re"], dtype=float))
    im = case["im"]
    if im is None:
        assert not np.iscomplexobj(dense) or np.all(dense.imag == 0)
    else:
        np.testing.assert_array_equal(dense.imag, np.asarray(im, dtype=float))
    if case["defect"] is None:
        assert got.hermitian_defect == -1.0
    elif case["defect"] == 0.0:
        assert got.hermitian_defect == 0.0
    else:
        assert got.hermitian_defect == pytest.approx(case["defect"], rel=1e-10)


@pytest.mark.parametrize("case", [c for c in EXPECTED["errors"] if c["reader"] == "read_npz"],
                         ids=[c["file"] for c in EXPECTED["errors"] if c["reader"] == "read_npz"])
def test_damaged_archives_rejected(case):
    if not (ROOT / case["file"]).exists():
        pytest.skip("fixture not copied")
    with pytest.raises(ERRORS[case["error"]]):
        hio.read_npz(ROOT / case["file"])


@pytest.mark.parametrize("case", EXPECTED["golden"], ids=[c["file"] for c in EXPECTED["golden"]])
def test_npy_writers_byte_identical(case, tmp_path):
    want = (ROOT / case["file"]).read_bytes()
    dt = {"float64": np.float64, "complex128": np.complex128, "float32": np.float32, "int32": np.int32,
          "int64": np.int64}[case["type"]]
    if np.dtype(dt).kind in "iu":
        vals = np.asarray(case["re"], dtype=dt)          # exact integers (float64 would round)
    else:
        re = np.asarray(case["re"], dtype=np.float64)
        vals = re if case["im"] is None else re + 1j * np.asarray(case["im"], dtype=np.float64)
    out = tmp_path / "x.npy"
    if case["kind"] == "vector":
        hio.write_npy_vector(out, vals.astype(dt))
    else:
        m = vals.astype(dt).reshape(case["cols"], case["rows"]).T  # values listed column-major
        hio.write_npy_matrix(out, m)
    assert out.read_bytes() == want


def _check_raw(arr, e):
    assert arr.dtype.str == e["dtype"] or (arr.dtype.str[1:] == e["dtype"][1:] and e["dtype"][0] in "<>|=")
    assert list(arr.shape) == e["shape"]
    if arr.ndim > 1 and arr.size:
        assert bool(arr.flags.f_contiguous and not arr.flags.c_contiguous) == e["fortran_order"]
    if e.get("convert") is None:
        return
    flat = np.ravel(arr, order="F" if e["fortran_order"] else "C")   # payload order
    if np.iscomplexobj(flat):
        np.testing.assert_array_equal(flat.real, np.asarray(e["re"], dtype=np.float64))
        np.testing.assert_array_equal(flat.imag, np.asarray(e["im"], dtype=np.float64))
    elif flat.dtype.kind in "iub":
        np.testing.assert_array_equal(flat.astype(np.int64), np.asarray(e["re"], dtype=np.int64))
    else:
        np.testing.assert_array_equal(flat.astype(np.float64), np.asarray(e["re"], dtype=np.float64))


@pytest.mark.parametrize("case", EXPECTED["npy"], ids=[c["file"] for c in EXPECTED["npy"]])
def test_read_npy_raw(case):
    _check_raw(hio.read_npy(ROOT / case["file"]), case)


@pytest.mark.parametrize("case", EXPECTED["npz"], ids=[c["file"] for c in EXPECTED["npz"]])
def test_read_npz_members(case):
    got = hio.read_npz(ROOT / case["file"])
    assert sorted(got) == sorted(case["keys"])
    for k, e in case["arrays"].items():
        _check_raw(got[k], e)
