"""The reference's plain-text instance/answer interchange format
(optimisation_solver_amd/lpio.py): exact round trips, infinities, a
hand-written file, and the answer layout."""
import math

import numpy as np

import problems
from optimisation_solver_amd import box_lp
from optimisation_solver_amd.lpio import read_answer, read_lp_txt, write_answer, write_lp_txt


def _same(a, b):
    for f in ("rp", "ci", "rv", "cp", "ri", "cv", "c", "lo", "hi", "rlo", "rhi"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
    assert a.offset == b.offset and (a.m, a.n) == (b.m, b.n)


def test_round_trip_is_exact(tmp_path):
    for lp in (box_lp(300, 750, 5, seed=3), problems.closed_form()["equality_row"][0], problems.unbounded()):
        p = tmp_path / "x.lp.txt"
        write_lp_txt(lp, str(p))
        _same(lp, read_lp_txt(str(p)))


def test_hand_written_file(tmp_path):
    p = tmp_path / "h.lp.txt"
    p.write_text("2 3 4\n1.5\n1 -2 0 \n0 -inf 0 \ninf 4 1 \n-inf 3 \n2 inf \n0 0 1\n0 2 -1\n1 1 2.5\n1 0 1\n")
    lp = read_lp_txt(str(p))
    assert (lp.m, lp.n, lp.nnz, lp.offset) == (2, 3, 4, 1.5)
    assert lp.lo[1] == -math.inf and lp.hi[0] == math.inf and lp.rlo[0] == -math.inf and lp.rhi[1] == math.inf
    A = np.zeros((2, 3))
    for i in range(2):
        for k in range(lp.rp[i], lp.rp[i + 1]):
            A[i, lp.ci[k]] = lp.rv[k]
    assert np.array_equal(A, [[1, 0, -1], [1, 2.5, 0]])


def test_answer_layout(tmp_path):
    p = tmp_path / "a.answer.txt"
    x = np.array([0.1, -2.0, 1e-300])
    write_answer(str(p), "optimal", -1.25, 1e-9, 2e-9, 3e-9, 0.5, x)
    first = p.read_text().splitlines()[0].split()
    assert first[0] == "optimal" and len(first) == 6
    a = read_answer(str(p))
    assert a.status == "optimal" and a.primal_objective == -1.25 and np.array_equal(a.primal, x)
