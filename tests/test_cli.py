"""Command line (apps/cli.cpp): argument merging, exit codes and the error channel on the CPU;
a full run on the GPU writes eigenvalues.npy, eigenvectors.npy and report.json for a file input
whose pairs the evaluator accepts."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp

from hsolve_amd import cli

GOLDEN_IO = Path(__file__).parent / "golden" / "io"


def test_overrides_merge_and_revalidate(tmp_path):
    cfg_file = tmp_path / "config.json"
    cfg_file.write_text(json.dumps({"tol": 1e-9, "target": {"mode": "nearest", "sigma": 1.0, "k": 3}}))
    args = cli._parser().parse_args(["--A", "a.npz", "--config", str(cfg_file), "--interval", "2", "5",
                                     "--slice-size", "40", "--precision", "double"])
    c = cli.merged_config(args)
    assert c["target"]["mode"] == "interval" and c["target"]["lower"] == 2.0 and c["target"]["upper"] == 5.0
    assert c["tol"] == 1e-9 and c["slice_size"] == 40 and c["factor_precision"] == "double"
    args = cli._parser().parse_args(["--A", "a.npz", "--sigma", "7.5", "--k", "4"])
    c = cli.merged_config(args)
    assert (c["target"]["mode"], c["target"]["sigma"], c["target"]["k"]) == ("nearest", 7.5, 4)


def test_exit_codes_and_errors(tmp_path, capsys):
    assert cli.main(["--version"]) == 0
    assert cli.main([]) == 2                                   # --A missing
    assert "--A is required" in capsys.readouterr().err
    assert cli.main(["--interval", "1", "2", "--smallest", "3", "--A", "x.npz"]) == 2   # exclusive options
    rc = cli.main(["--A", str(GOLDEN_IO / "truncated.mtx"), "--sigma", "0", "--out", str(tmp_path / "o")])
    assert rc == 2 and "hsolve: error [io]" in capsys.readouterr().err
    rc = cli.main(["--A", str(GOLDEN_IO / "non_hermitian_f8.npz"), "--sigma", "0", "--out", str(tmp_path / "o")])
    assert rc == 2 and "hsolve: error [not_hermitian]" in capsys.readouterr().err
    bad = tmp_path / "bad.json"
    bad.write_text('{"tol": "small"}')
    rc = cli.main(["--A", str(GOLDEN_IO / "csr_matrix_f8.npz"), "--config", str(bad)])
    assert rc == 2 and "hsolve: error [invalid_input]" in capsys.readouterr().err


@pytest.mark.gpu
def test_cli_solves_file_input(gpu, tmp_path):
    from hsolve_amd.gen import yee
    from oracle import evaluate as ev
    from oracle import yee_exact
    A, _, g = yee.cube_cavity(6)
    ref = yee_exact.pec_box_spectrum(g.cells, g.lengths)
    a_path = tmp_path / "A.npz"
    sp.save_npz(a_path, sp.triu(A).tocsr())                    # one stored triangle: restored on load
    out = tmp_path / "out"
    rc = cli.main(["--A", str(a_path), "--sigma", "9.0", "--k", "10", "--out", str(out), "-q"])
    assert rc == 0
    lam = np.load(out / "eigenvalues.npy")
    X = np.load(out / "eigenvectors.npy")
    rep = json.loads((out / "report.json").read_text())
    assert lam.shape == (10,) and X.shape == (A.shape[0], 10) and X.flags["F_CONTIGUOUS"]
    assert rep["success"] and rep["certificate"]["passed"]
    assert any(w.startswith("input: A: storage") for w in rep["warnings"])
    assert not (out / "complex_eigenvalues.npy").exists()
    res = ev.evaluate(A, None, lam, X, {"mode": "nearest", "sigma": 9.0, "k": 10}, ref)
    assert res["passed"], res


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["pec-cavity-multiplicity", "interval-slicing"])
def test_cli_solves_a_case_directory_and_the_evaluator_passes_it(gpu, tmp_path, name):
    # the reference's data layout end to end: a shipped case (matrices, configuration, reference
    # spectrum) solved through the CLI's --case, the result directory judged by the evaluator's
    # command line
    from oracle import evaluate as ev
    case_dir = Path(__file__).parent / "golden" / "examples" / name
    out = tmp_path / "result"
    assert cli.main(["--case", str(case_dir), "--out", str(out), "-q"]) == 0
    assert ev.main([str(case_dir), str(out)]) == 0
    doc = json.loads((out / "evaluation.json").read_text())
    print(doc["case"], doc["verdict"], doc["summary"], flush=True)
    assert doc["verdict"] == "pass" and doc["reference"]["complete"]
