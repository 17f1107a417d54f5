"""The library's MPS reader (sk_lp_read_mps; src/io/mps_reader.cpp, mps_bounds.cpp semantics)
against the Python reader on every MPS fixture, and rule by rule on small files."""
from pathlib import Path

import numpy as np
import pytest

import sih26_amd as sk
from tests.mps_ref import read_mps

GOLDEN = Path(__file__).resolve().parent / "golden"
FIXTURES = sorted(GOLDEN.glob("netlib/*.mps")) + sorted(GOLDEN.glob("checker/*.mps"))
ARRAYS = ("col_starts", "row_indices", "values", "cost", "col_lower", "col_upper", "row_lower",
          "row_upper")
INF = np.inf


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: p.stem)
def test_library_reader_matches_the_python_reader(path):
    lp = sk.Lp.read_mps(path)
    got, want = lp.export(), read_mps(path)
    for k in ARRAYS:
        np.testing.assert_array_equal(got[k], want[k], err_msg=k)
    assert (got["sense"], got["offset"]) == (want["sense"], want["offset"])
    name, rows, cols = lp.names()
    assert rows == want["row_names"] and cols == want["col_names"]
    assert lp.integer_columns == 0


def read(tmp_path, text, name="t.mps"):
    p = tmp_path / name
    p.write_text(text)
    return sk.Lp.read_mps(p)


def fails(tmp_path, text, *words):
    p = tmp_path / "bad.mps"
    p.write_text(text)
    with pytest.raises(sk.SkError) as e:
        sk.Lp.read_mps(p)
    assert e.value.status == 10
    msg = sk.lib().sk_gpu_last_error().decode()
    for w in words:
        assert w in msg, msg
    return msg


BASE = """NAME          TINY
ROWS
 N  COST
 L  LIM1
 G  LIM2
 E  MYEQN
COLUMNS
    X1        COST         1.0   LIM1         1.0
    X1        LIM2         1.0
    X2        COST         2.0   LIM1         1.0
    X2        MYEQN       -1.0
    X3        COST        -1.0   MYEQN        1.0
RHS
    RHS       COST        -3.5
    RHS       LIM1         4.0   LIM2         1.0
    RHS       MYEQN        7.0
    OTHER     LIM1        99.0
BOUNDS
 UP BND       X1           4.0
 LO BND       X2          -1.0
 UP BND       X2           1.0
 UP BND       X3          -2.0
ENDATA
"""


def test_semantics_of_a_small_file(tmp_path):
    lp = read(tmp_path, BASE)
    d = lp.export()
    assert lp.names()[0] == "TINY" and lp.dims == (3, 3, 5)
    assert d["offset"] == 3.5  # the objective row's RHS is the negated constant
    np.testing.assert_array_equal(d["row_lower"], [-INF, 1.0, 7.0])
    np.testing.assert_array_equal(d["row_upper"], [4.0, INF, 7.0])  # OTHER vector ignored
    # UP < 0 with no explicit lower bound: the lower bound becomes -inf (X3); not for X2 (LO)
    np.testing.assert_array_equal(d["col_lower"], [0.0, -1.0, -INF])
    np.testing.assert_array_equal(d["col_upper"], [4.0, 1.0, -2.0])
    np.testing.assert_array_equal(d["cost"], [1.0, 2.0, -1.0])


def test_ranges_and_objsense(tmp_path):
    text = """NAME R
OBJSENSE
    MAXIMIZE
ROWS
 N obj
 E e1
 E e2
 L l1
 G g1
COLUMNS
 x obj 1 e1 1
 x e2 1 l1 1
 x g1 1
RHS
 rhs e1 10 e2 10
 rhs l1 10 g1 10
RANGES
 rng e1 4 e2 -4
 rng l1 3 g1 -3
ENDATA
"""
    d = read(tmp_path, text).export()
    assert d["sense"] == -1
    np.testing.assert_array_equal(d["row_lower"], [10.0, 6.0, 7.0, 10.0])
    np.testing.assert_array_equal(d["row_upper"], [14.0, 10.0, 10.0, 13.0])
    # OBJSENSE with its value on the header line
    d2 = read(tmp_path, text.replace("OBJSENSE\n    MAXIMIZE\n", "OBJSENSE MAX\n"), "t2.mps").export()
    assert d2["sense"] == -1


def test_bound_types_markers_comments_and_free_rows(tmp_path):
    text = """NAME B
ROWS
 N obj
 N spare
 L c1
COLUMNS
 MARKER 'MARKER' 'INTORG'
 i1 obj 1 c1 1
 MARKER 'MARKER' 'INTEND'
 a obj 1 c1 1D0 $ a Fortran exponent and a comment
 a spare 5
 b obj 1 c1 2
 c obj 1 c1 3
 d obj 1 c1 4
 e obj 1 c1 1e-12
RHS
 rhs c1 1e30
BOUNDS
 FR bnd a
 MI bnd b
 PL bnd b
 BV bnd c
 UI bnd d 7
 LI bnd d 2
 FX bnd i1 3
ENDATA
"""
    lp = read(tmp_path, text)
    d = lp.export()
    assert lp.integer_columns == 3  # i1 by marker; c (BV) and d (UI/LI) by bound type
    assert lp.dims[0] == 1  # the second N row is dropped with its entries
    np.testing.assert_array_equal(d["col_lower"], [3.0, -INF, -INF, 0.0, 2.0, 0.0])
    np.testing.assert_array_equal(d["col_upper"], [3.0, INF, INF, 1.0, 7.0, INF])
    np.testing.assert_array_equal(d["row_upper"], [INF])  # 1e30 is infinity
    # the 1e-12 entry is below the drop tolerance (SparseMatrix::finalize)
    np.testing.assert_array_equal(d["values"], [1.0, 1.0, 2.0, 3.0, 4.0])


def fixed_line(*fields):
    """A fixed-format data line: fields 1-6 start in columns 2, 5, 15, 25, 40 and 50."""
    line = ""
    for start, text in zip((1, 4, 14, 24, 39, 49), fields):
        line = line.ljust(start) + text
    return line + "\n"


def test_fixed_format_with_blanks_in_names(tmp_path):
    text = ("NAME          FIXED\n" "ROWS\n" + fixed_line("N", "OBJ ROW") + fixed_line("L", "ROW ONE") +
            "COLUMNS\n" + fixed_line("", "COL A", "OBJ ROW", "1.0", "ROW ONE", "2.0") +
            "RHS\n" + fixed_line("", "RHS", "ROW ONE", "8.0") + "ENDATA\n")
    lp = read(tmp_path, text)
    d = lp.export()
    _, rows, cols = lp.names()
    assert rows == ["ROW ONE"] and cols == ["COL A"]
    np.testing.assert_array_equal(d["values"], [2.0])
    np.testing.assert_array_equal(d["cost"], [1.0])
    np.testing.assert_array_equal(d["row_upper"], [8.0])


@pytest.mark.parametrize("text, words", [
    (BASE.replace("    X1        LIM2         1.0\n", "    X1        NOPE         1.0\n"),
     ("bad.mps:9:", "unknown row 'NOPE'")),
    (BASE.replace("    X1        LIM2         1.0\n", "    X1        LIM1         2.0\n"),
     ("duplicate entry", "no accumulate semantics")),
    (BASE.replace("ENDATA\n", ""), ("without an ENDATA",)),
    (BASE.replace("4.0   LIM2", "four   LIM2"), ("'four' is not a number",)),
    (BASE.replace("4.0   LIM2", "nan   LIM2"), ("'nan' is not a number",)),
    (BASE.replace("1.0   LIM1         1.0\n    X1", "1.0   LIM1       1e400\n    X1"),
     ("must be finite",)),
    (BASE.replace("BOUNDS\n", "RANGES\n    RNG       COST         1.0\nBOUNDS\n"),
     ("RANGES entry on the objective row",)),
    (BASE.replace(" UP BND       X1           4.0\n", " SC BND       X1           4.0\n"),
     ("semi-continuous",)),
    (BASE.replace(" UP BND       X1           4.0\n", " UP BND       X9           4.0\n"),
     ("no COLUMNS entry",)),
    (BASE.replace("ROWS\n", "QUADOBJ\n"), ("quadratic",)),
    (BASE.replace(" L  LIM1\n", " L  LIM1\n L  LIM1\n"), ("duplicate row name",)),
])
def test_malformed_files_are_refused_with_their_line(tmp_path, text, words):
    fails(tmp_path, text, *words)


def test_missing_file_and_null_arguments(tmp_path):
    with pytest.raises(sk.SkError):
        sk.Lp.read_mps(tmp_path / "nowhere.mps")
    assert "cannot open" in sk.lib().sk_gpu_last_error().decode()
    assert sk.lib().sk_lp_read_mps(None, None, None) == 10
