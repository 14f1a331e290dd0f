"""Matrix Market reader: the CRS the reference's benchmarks build from a .mtx
file (examples/matrix_structs.hpp:52-180, 338-388).  Expected arrays are
worked out by hand from that procedure: mirror symmetric off-diagonal
entries right after the original, stable-sort by row, keep file order
within a row."""
from __future__ import annotations

import numpy as np
import pytest

from smaxkernels_amd import matgen
from tests.helpers import assert_bitwise


def write(tmp_path, text, name="m.mtx"):
    p = tmp_path / name
    p.write_text(text)
    return p


def test_symmetric_real_mirrors_and_keeps_file_order(tmp_path):
    p = write(tmp_path, """%%MatrixMarket matrix coordinate real symmetric
% a comment

3 3 4
1 1 4.0
2 1 -1.5
3 2 2e-1
3 3 7
""")
    A = matgen.read_mtx(p)
    assert (A.n_rows, A.n_cols) == (3, 3)
    assert_bitwise(A.row_ptr, np.array([0, 2, 4, 6], np.int32))
    assert_bitwise(A.col, np.array([0, 1, 0, 2, 1, 2], np.int32))
    assert_bitwise(A.val, np.array([4.0, -1.5, -1.5, 0.2, 0.2, 7.0]))


def test_general_rows_keep_unsorted_columns(tmp_path):
    """Columns within a row stay in file order (the reference sorts by row
    only), and rows are stable-sorted."""
    p = write(tmp_path, """%%MatrixMarket matrix coordinate real general
2 4 5
2 4 1.0
1 3 2.0
1 1 3.0
2 1 4.0
1 4 5.0
""")
    A = matgen.read_mtx(p)
    assert_bitwise(A.row_ptr, np.array([0, 3, 5], np.int32))
    assert_bitwise(A.col, np.array([2, 0, 3, 3, 0], np.int32))
    assert_bitwise(A.val, np.array([2.0, 3.0, 5.0, 1.0, 4.0]))


def test_pattern_and_integer_fields(tmp_path):
    p = write(tmp_path, """%%MatrixMarket matrix coordinate pattern general
2 2 2
1 2
2 1
""")
    A = matgen.read_mtx(p)
    assert_bitwise(A.val, np.array([0.01, 0.01]))  # the reference's pattern value
    q = write(tmp_path, """%%MatrixMarket MATRIX Coordinate INTEGER General
1 2 1
1 2 -3
""", "i.mtx")
    B = matgen.read_mtx(q)
    assert_bitwise(B.val, np.array([-3.0]))
    assert_bitwise(B.col, np.array([1], np.int32))


def test_empty_rows_and_round_trip(tmp_path):
    """A generated matrix written 1-based in CRS order reads back unchanged."""
    A = matgen.poisson2d(9)
    lines = ["%%MatrixMarket matrix coordinate real general",
             f"{A.n_rows} {A.n_cols} {A.nnz}"]
    for r in range(A.n_rows):
        for e in range(A.row_ptr[r], A.row_ptr[r + 1]):
            lines.append(f"{r + 1} {A.col[e] + 1} {float(A.val[e])!r}")
    p = write(tmp_path, "\n".join(lines) + "\n")
    B = matgen.read_mtx(p)
    assert_bitwise(B.row_ptr, A.row_ptr)
    assert_bitwise(B.col, A.col)
    assert_bitwise(B.val, A.val)
    q = write(tmp_path, """%%MatrixMarket matrix coordinate real general
4 4 1
3 3 1.0
""", "e.mtx")
    E = matgen.read_mtx(q)
    assert_bitwise(E.row_ptr, np.array([0, 0, 0, 1, 1], np.int32))


@pytest.mark.parametrize("text,msg", [
    ("%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1 0\n", "unsupported"),
    ("%%MatrixMarket matrix array real general\n1 1\n1\n", "unsupported"),
    ("%%MatrixMarket matrix coordinate real skew-symmetric\n2 2 1\n2 1 1\n", "unsupported"),
    ("%%MatrixMarket matrix coordinate real hermitian\n2 2 1\n2 1 1\n", "unsupported"),
    ("1 1 1\n1 1 1\n", "banner"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n", "out-of-range"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1.0\n", "out-of-range"),
    ("%%MatrixMarket matrix coordinate real general\n2 2\n", "size"),
])
def test_rejections(tmp_path, text, msg):
    p = write(tmp_path, text)
    with pytest.raises(ValueError, match=msg):
        matgen.read_mtx(p)


def test_missing_file(tmp_path):
    with pytest.raises(ValueError, match="cannot open"):
        matgen.read_mtx(tmp_path / "nope.mtx")
