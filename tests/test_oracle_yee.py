"""Pins of the oracle's closed-form spectra and of the workload generator (CPU).

The closed forms (oracle/yee_exact.py, restating tools/gen/yee.py:231-260) are checked against the
dense eigenvalues of the assembled pencils (the reference's own check, tests/py/test_gen_yee.py:30-77,
1e-13 * scale), and the generator's orders against the workloads' stated sizes.
"""
import numpy as np
import pytest
import scipy.linalg as la

from hsolve_amd.gen import yee
from oracle import inertia as oinertia
from oracle import yee_exact


@pytest.mark.parametrize("cells", [(3, 3, 3), (4, 3, 5), (5, 4, 4)])
def test_pec_closed_form_matches_dense(cells):
    lengths = (2.0, 1.0, 1.5)
    g = yee.YeeGrid(cells, lengths, "pec")
    A, B = yee.assemble(g)
    w = la.eigvalsh(A.toarray(), B.toarray())
    ref = yee_exact.pec_box_spectrum(cells, lengths)
    assert np.max(np.abs(np.sort(w) - ref)) <= 1e-12 * ref.max()


@pytest.mark.parametrize("phases", [(0.0, 0.0, 0.0), (np.pi / 3, 0.0, np.pi), (0.7, 1.1, 2.0)])
def test_bloch_closed_form_matches_dense(phases):
    cells, lengths = (3, 4, 3), (1.0, 1.2, 0.9)
    g = yee.YeeGrid(cells, lengths, "bloch", phases)
    A, B = yee.assemble(g)
    w = la.eigvalsh(A.toarray(), B.toarray())
    ref = yee_exact.bloch_box_spectrum(cells, lengths, phases)
    assert np.max(np.abs(np.sort(w) - ref)) <= 1e-12 * ref.max()


def test_closed_form_catches_a_wrong_multiplicity():
    # a plausible mistake (TE/TM counted once) must fail the dense comparison
    cells, lengths = (3, 3, 3), (1.0, 1.0, 1.0)
    g = yee.YeeGrid(cells, lengths, "pec")
    A, B = yee.assemble(g)
    w = np.sort(la.eigvalsh(A.toarray(), B.toarray()))
    ref = yee_exact.pec_box_spectrum(cells, lengths)
    wrong = np.unique(ref)
    assert wrong.size != w.size or np.max(np.abs(w - np.pad(wrong, (0, w.size - wrong.size)))) > 1e-6


def test_workload_orders():
    A, B, g = yee.cube_cavity(20)
    assert A.shape[0] == 21660 and B is None
    A2, B2, g2 = yee.jittered_box()
    assert A2.shape[0] == 117036
    d = B2.diagonal()
    h = np.pi / 24
    assert np.all(np.abs(d / h - 1.0) <= 0.05 + 1e-12)
    A3, B3, g3 = yee.bloch_cell(16, (np.pi / 3, 0.0, np.pi))
    assert A3.shape[0] == 12288 and np.iscomplexobj(A3.data)
    assert np.allclose(np.unique(np.round(B3.diagonal() / (1.0 / 16), 12)), [1.0, 12.0])


def test_cube20_count_at_sigma10():
    A, _, g = yee.cube_cavity(20)
    w = yee_exact.pec_box_spectrum(g.cells, g.lengths)
    assert yee_exact.count_below(w, 10.0) == 6891
    # the 21 nearest: levels x6, x6, x6, x2 and one copy of a 3-fold level (a tie cut by k)
    d = np.abs(w - 10.0)
    order = np.argsort(d, kind="stable")
    kth = d[order[20]]
    assert np.sum(d < kth - 1e-9) == 20 and np.sum(np.abs(d - kth) <= 1e-9) == 3


def test_sparse_inertia_matches_dense_and_closed_form():
    A, _, g = yee.cube_cavity(4)
    w = yee_exact.pec_box_spectrum(g.cells, g.lengths)
    for s in yee_exact.gap_points(w, 0.5, w.max())[::3]:
        K = oinertia.shifted(A, None, s)
        assert oinertia.sparse_inertia(K) == oinertia.dense_inertia(K)
        assert oinertia.sparse_inertia(K)[0] == yee_exact.count_below(w, s)
