"""The reference's data layout (tools/case.py): cases, references and result directories written
and read back, the shipped examples read as cases, and the evaluator's command line judging a
result directory against a case (tools/eval/evaluate.py:805-840) - passing a correct CPU solve
and failing one with a dropped pair. Host only."""
import json
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp

from hsolve_amd import case as cio
from hsolve_amd.gen import yee
from oracle import eigs_cpu
from oracle import evaluate as ev
from oracle import yee_exact

GOLDEN = Path(__file__).parent / "golden" / "examples"


def test_round_trip_case_ref_result(tmp_path):
    A, B, g = yee.jittered_box(cells=(5, 4, 4), lengths=(2 * np.pi, np.pi, 1.5 * np.pi), seed=1)
    d = cio.save_case("box", A, B, meta={"family": "yee", "problem": "GHEP", "scenarios": ["multiplicity"]},
                      config={"target": {"mode": "nearest", "sigma": 9.0, "k": 4}, "tol": 1e-10}, root=tmp_path)
    c = cio.load_case("box", tmp_path)
    assert c.path == d and c.n == A.shape[0] and (c.A != sp.csr_matrix(A)).nnz == 0
    assert c.meta["problem"] == "GHEP" and c.meta["nnz_B"] == B.nnz and c.config["target"]["k"] == 4
    cio.save_ref("box", [3.0, 1.0, 2.0], kind="dense", complete=True, root=tmp_path)
    r = cio.load_ref("box", tmp_path)
    assert list(r.eigenvalues) == [1.0, 2.0, 3.0] and r.info["complete"] and r.info["count"] == 3
    X = np.asfortranarray(np.random.default_rng(0).standard_normal((A.shape[0], 2)))
    cio.save_result(tmp_path / "res", [1.0, 2.0], X, {"success": True})
    res = cio.load_result(tmp_path / "res")
    assert np.array_equal(res.eigenvectors, X) and res.report["schema_version"] == 1
    assert res.complex_eigenvalues is None
    with pytest.raises(ValueError):
        cio.save_case("bad", A, meta={"family": "yee", "problem": "GHEP", "eigenvalues": [1.0]}, root=tmp_path)
    with pytest.raises(ValueError):
        cio.save_case("bad", A, meta={"problem": "HEP"}, root=tmp_path)


@pytest.mark.parametrize("name", sorted(p.name for p in GOLDEN.iterdir() if (p / "A.npz").exists()))
def test_shipped_examples_read_as_cases(name):
    c = cio.load_case(GOLDEN / name)
    ref = cio.load_ref(GOLDEN / name)
    assert c.n == c.A.shape[0] and "target" in c.config
    assert ref is not None and ref.eigenvalues.size == ref.info.get("count", ref.eigenvalues.size)


def test_evaluator_command_line_judges_a_result_directory(tmp_path):
    A, _, g = yee.cube_cavity(6)
    w = yee_exact.pec_box_spectrum(g.cells, g.lengths)
    cio.save_case("cav", A, meta={"family": "yee", "problem": "HEP"},
                  config={"target": {"mode": "nearest", "sigma": 9.0, "k": 8}, "tol": 1e-10}, root=tmp_path)
    cio.save_ref("cav", w, kind="discrete-exact", complete=True, root=tmp_path)
    r = eigs_cpu.solve_nearest(A, None, 9.0, 8)
    cio.save_result(tmp_path / "good", r["eigenvalues"], r["eigenvectors"], {"success": True})
    assert ev.main(["cav", str(tmp_path / "good"), "--root", str(tmp_path)]) == 0
    doc = json.loads((tmp_path / "good" / "evaluation.json").read_text())
    assert doc["verdict"] == "pass" and doc["summary"]["n_pairs"] == 8
    # a result missing one of the required pairs fails, with the reason given
    cio.save_result(tmp_path / "bad", r["eigenvalues"][:7], r["eigenvectors"][:, :7], {"success": True})
    assert ev.main(["cav", str(tmp_path / "bad"), "--root", str(tmp_path)]) == 1
    doc = json.loads((tmp_path / "bad" / "evaluation.json").read_text())
    assert doc["verdict"] == "fail" and any("missing" in x for x in doc["reasons"])
