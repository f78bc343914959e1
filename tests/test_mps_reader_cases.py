"""The LP MPS reader against the reference reader's own test cases.

The files under tests/golden/mps/reader_cases/ are the reference's reader
fixtures (tests/mps/test_cases/*.mps); each expectation below is the one the
reference's tests state for that file (tests/mps/mps_reader_tests.cpp:9-460),
read through this repository's LP-only reader.  Files that carry integer
records (BV/LI/UI bounds, MARKER lines) are rejected as LP-only, and their LP
content is checked with those records taken out.
"""
from __future__ import annotations

import math
from pathlib import Path

import numpy as np
import pytest

from optimisation_solver_amd.mps import MpsError, read_mps

CASES = Path(__file__).resolve().parent / "golden" / "mps" / "reader_cases"
INF = math.inf


def read(name):
    return read_mps(str(CASES / f"{name}.mps"))


def row(m, rname):
    i = m.row_names.index(rname)
    return float(m.lp.rlo[i]), float(m.lp.rhi[i])


def col(m, cname):
    j = m.col_names.index(cname)
    return float(m.lp.lo[j]), float(m.lp.hi[j])


def cost(m, cname):
    return float(m.lp.c[m.col_names.index(cname)])


def terms(m, rname):
    """Nonzero coefficients of a row: {column name: value}."""
    i = m.row_names.index(rname)
    a, b = m.lp.rp[i], m.lp.rp[i + 1]
    return {m.col_names[j]: float(v) for j, v in zip(m.lp.ci[a:b], m.lp.rv[a:b])}


def test_basic_lp_row_senses():
    m = read("01_basic_lp")
    assert (m.lp.n, m.lp.m) == (2, 3)
    assert np.count_nonzero(m.lp.c) == 2
    assert row(m, "C_LESS") == (-INF, 10.0)
    assert row(m, "C_GREATER") == (5.0, INF)
    assert row(m, "C_EQUAL") == (7.0, 7.0)
    assert not m.maximize


def test_multi_column_coefficients():
    m = read("02_multi_column_coeffs")
    assert (m.lp.n, m.lp.m) == (2, 2)
    assert terms(m, "ROW1") == {"VAR1": 2.5}


def test_all_bound_types():
    m = read("03_all_bound_types")
    assert col(m, "X_LO") == (2.5, INF)
    assert col(m, "X_UP") == (0.0, 10.0)
    assert col(m, "X_FX") == (7.0, 7.0)
    assert col(m, "X_FR") == (-INF, INF)


def test_variable_only_in_bounds():
    m = read("04_bounds_only_variable")
    assert m.lp.n == 2
    assert col(m, "X_UNSEEN") == (3.0, 8.0)


def test_negatives_comments_and_blank_lines():
    m = read("05_negatives_comments_blanks")
    assert m.name == "NEG_COMMENTS"
    assert np.count_nonzero(m.lp.c) == 2
    assert cost(m, "X1") == pytest.approx(-5.25, abs=1e-12)
    assert row(m, "CON1")[1] == -15.0


def test_missing_sections_give_an_empty_model():
    m = read("06_malformed_missing_section")
    assert m.name == "MALFORMED"
    assert m.lp.m == 0
    assert any("ROW1" in w for w in m.warnings)   # the RHS names a row that does not exist


def test_ranges_on_every_row_kind():
    m = read("07_ranges")
    assert row(m, "C1") == (6.0, 10.0)
    assert row(m, "C2") == (5.0, 9.0)
    assert row(m, "C3") == (7.0, 10.0)
    assert row(m, "C4") == (4.0, 7.0)


def test_extended_bounds_lp_part():
    text = (CASES / "09_extended_bounds.mps").read_text()
    with pytest.raises(MpsError, match="LP only"):
        read_mps(text, is_text=True)
    lp_only = "\n".join(ln for ln in text.splitlines() if ln.split()[:1] not in (["BV"], ["LI"], ["UI"]))
    m = read_mps(lp_only, is_text=True)
    assert col(m, "X_MI") == (-INF, 4.0)
    assert col(m, "X_PL")[1] == INF
    assert col(m, "X_NEGUP") == (-INF, -3.0)   # a negative UP opens the implicit lower bound
    assert any("X_NEGUP" in w for w in m.warnings)
    assert all(lo <= hi for lo, hi in zip(m.lp.lo, m.lp.hi))


def test_objective_sense():
    assert read("10_objsense_max").maximize
    assert read("11_objsense_inline").maximize
    assert not read("01_basic_lp").maximize
    m = read("10_objsense_max")
    assert cost(m, "X") == -3.0 and m.objective(-30.0) == 30.0   # returned as the minimisation


def test_objective_constant_is_the_negated_rhs():
    m = read("12_objective_constant")
    assert m.lp.offset == pytest.approx(25.0, abs=1e-12)
    assert m.lp.m == 1 and row(m, "C1")[0] == 4.0


def test_unknown_and_sos_sections_are_rejected():
    with pytest.raises(MpsError, match="NOTASECTION"):
        read("15_unknown_section")
    with pytest.raises(MpsError, match="SOS"):
        read("16_sos_section")


def test_first_rhs_vector_wins():
    m = read("17_multiple_rhs")
    assert row(m, "C1")[1] == 10.0      # a later vector does not overwrite
    assert row(m, "C2")[0] == 5.0       # the first vector's continuation counts
    assert row(m, "C3")[1] == 0.0       # a later vector does not fill an absent entry
    assert m.lp.offset == 25.0          # nor the objective constant
    assert m.warnings


def test_first_ranges_vector_wins():
    m = read("18_multiple_ranges")
    assert row(m, "C1") == (6.0, 10.0)
    assert row(m, "C2") == (5.0, 8.0)
    assert row(m, "C3") == (-INF, 8.0)
    assert any("RANGES vector 'RANGE_B'" in w for w in m.warnings)
    assert row(read("07_ranges"), "C1")[0] == 6.0   # the selection starts afresh per file


def test_rhs_and_ranges_without_a_vector_name():
    m = read("17_rhs_no_set_name")
    assert row(m, "C1") == (-INF, 10.0)
    assert row(m, "C2") == (5.0, 9.0)


def test_rows_and_columns_named_marker_are_data():
    text = (CASES / "19_marker_names.mps").read_text()
    with pytest.raises(MpsError, match="LP only"):
        read_mps(text, is_text=True)   # it does hold two real marker records
    markers = {"START", "STOP"}
    lp_only = "\n".join(ln for ln in text.splitlines() if ln.split()[:1] not in ([w] for w in markers))
    m = read_mps(lp_only, is_text=True)
    assert m.lp.n == 5
    assert len(terms(m, "C1")) == 4
    assert terms(m, "MARKER") == {"X": 5.0}
    assert row(m, "MARKER")[1] == 30.0
