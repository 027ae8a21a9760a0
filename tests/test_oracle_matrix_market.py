"""Pins the oracle's Matrix Market reader (oracle/mtj_oracle_mm.c), the CPU
restatement of CompRowMatrix(MatrixVectorReader) and DenseVector(
MatrixVectorReader), without retyping its rules:

* the reference's own write-then-read test (T/io/MatrixVectorIoTest.java:
  36-59): the 3x3 matrix written in MatrixVectorWriter's layout reads back
  as the same CSR (fixture in tests/golden/mtj_fixtures.json);
* scipy.io.mmread (an independent reader) on seeded general, symmetric,
  skew-symmetric and integer files without duplicate entries, compared as
  CSR arrays, values bitwise (17 significant digits round-trip exactly);
* known answers fixed by the Java semantics the reader inherits: last
  duplicate wins (CompRowMatrix.set), mirrored values of a symmetric file
  assigned after all direct ones, Double.valueOf's grammar (suffixes, hex
  floats; "nan"/"infinity" rejected after lower-casing), Double.intValue
  truncation of indices, comments and line breaks anywhere between tokens,
  and the exception taxonomy of each failure."""
import io
import json
import os

import numpy as np
import pytest
import scipy.io as sio
import scipy.sparse as sp

import oracle

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "mtj_fixtures.json")))


def _csr(res):
    code, r, c, rp, ci, va, sym = res
    assert code == oracle.OK, code
    return sp.csr_matrix((va, ci, rp), shape=(r, c)), sym


def test_reference_write_read_fixture():
    fx = GOLD["io_3x3"]
    text = GOLD["io_3x3_mm_text"]["text"].encode()
    code, r, c, rp, ci, va, sym = oracle.mm_matrix(text)
    assert code == oracle.OK and (r, c) == (fx["n_rows"], fx["n_cols"]) and sym == 0
    assert rp.tolist() == fx["row_pointer"]
    assert ci.tolist() == fx["column_index"]
    assert va.tolist() == fx["data"]


def _random_file(seed, n_rows, n_cols, density, symmetry="general", field="real"):
    rng = np.random.default_rng(seed)
    cells = set()
    entries = []
    while len(entries) < int(density * n_rows * n_cols):
        i, j = int(rng.integers(n_rows)), int(rng.integers(n_cols))
        if symmetry != "general":
            if i < j:
                i, j = j, i
            if symmetry == "skew-symmetric" and i == j:
                continue
        if (i, j) in cells:
            continue
        cells.add((i, j))
        v = int(rng.integers(-50, 50)) if field == "integer" else float(rng.normal())
        entries.append((i, j, v))
    lines = ["%%%%MatrixMarket matrix coordinate %s %s" % (field, symmetry), "% seeded test file",
             "%d %d %d" % (n_rows, n_cols, len(entries))]
    for i, j, v in entries:
        lines.append("%d %d %s" % (i + 1, j + 1, repr(v)))
    return ("\n".join(lines) + "\n").encode()


@pytest.mark.parametrize("case", [
    (1, 40, 40, 0.1, "general", "real"),
    (2, 37, 53, 0.08, "general", "real"),
    (3, 45, 45, 0.1, "symmetric", "real"),
    (4, 45, 45, 0.1, "skew-symmetric", "real"),
    (5, 30, 30, 0.2, "general", "integer"),
    (6, 60, 60, 0.05, "symmetric", "integer"),
])
def test_matches_scipy_mmread(case):
    text = _random_file(*case)
    A, _ = _csr(oracle.mm_matrix(text))
    B = sio.mmread(io.BytesIO(text)).tocsr()
    B.sort_indices()
    assert A.shape == B.shape
    np.testing.assert_array_equal(A.indptr, B.indptr)
    np.testing.assert_array_equal(A.indices, B.indices)
    np.testing.assert_array_equal(A.data, B.data.astype(np.float64))


def test_duplicates_last_wins_and_explicit_zeros_kept():
    text = b"%%MatrixMarket matrix coordinate real general\n2 2 4\n1 1 5\n1 1 7\n2 2 0\n1 1 9\n"
    A, _ = _csr(oracle.mm_matrix(text))
    assert A.indptr.tolist() == [0, 1, 2] and A.indices.tolist() == [0, 1]
    assert A.data.tolist() == [9.0, 0.0]


def test_symmetric_mirror_assigned_after_direct_entries():
    # both (2,1) and (1,2) given: the mirror pass overwrites each with the
    # other's value (file order), so the two end up swapped
    text = b"%%MatrixMarket matrix coordinate real symmetric\n2 2 3\n2 1 1.5\n1 2 2.5\n1 1 4\n"
    A, sym = _csr(oracle.mm_matrix(text))
    assert sym == 1
    assert A.toarray().tolist() == [[4.0, 1.5], [2.5, 0.0]]


def test_skew_symmetric_negates_mirror():
    text = b"%%MatrixMarket matrix coordinate real skew-symmetric\n3 3 2\n2 1 1.5\n3 2 -2\n"
    A, sym = _csr(oracle.mm_matrix(text))
    assert sym == 2
    assert A.toarray().tolist() == [[0, -1.5, 0], [1.5, 0, 2.0], [0, -2.0, 0]]


def test_no_header_comments_and_line_breaks_anywhere():
    text = (b"% leading comment, no header line\r\n3 3\n  2 %size split\n"
            b"1\n1.0 5%value then comment\r\n3\t2\x0b0x1.8p1\n")
    A, _ = _csr(oracle.mm_matrix(text))
    assert A.toarray().tolist() == [[5.0, 0, 0], [0, 0, 0], [0, 3.0, 0]]


@pytest.mark.parametrize("tok,val", [
    ("1.5d", 1.5), ("-.5e1f", -5.0), ("0x1.8p1", 3.0), ("1.", 1.0), ("+2", 2.0),
    ("1E2", 100.0), ("0X1P-2D", 0.25), ("1e400", float("inf")), ("4.9e-324", 5e-324),
])
def test_java_number_forms(tok, val):
    text = ("%%%%MatrixMarket matrix coordinate real general\n1 1 1\n1 1 %s\n" % tok).encode()
    A, _ = _csr(oracle.mm_matrix(text))
    assert A.data.tolist() == [val]


@pytest.mark.parametrize("tok", ["nan", "NaN", "Infinity", "1e", "0x1.8", ".", "1,5", "--1", "1df", "0x1.8p"])
def test_java_number_rejects(tok):
    text = ("%%%%MatrixMarket matrix coordinate real general\n1 1 1\n1 1 %s\n" % tok).encode()
    assert oracle.mm_matrix(text)[0] == oracle.BAD_ARG


def test_index_truncation():
    # indices pass through Double.valueOf(...).intValue()
    text = b"%%MatrixMarket matrix coordinate real general\n3.9 3 2\n2.7 3e0 1\n1 1.99 2\n"
    A, _ = _csr(oracle.mm_matrix(text))
    assert A.shape == (3, 3)
    assert A.toarray().tolist() == [[2.0, 0, 0], [0, 0, 1.0], [0, 0, 0]]


@pytest.mark.parametrize("text,code", [
    (b"", oracle.IO),
    (b"%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1\n", oracle.IO),
    (b"%%MatrixMarket matrix coordinate real\n2 2 0\n", oracle.IO),
    (b"%%MatrixMarket matrix\tcoordinate real general x\n", oracle.IO),
    (b"%%MatrixMarket vector coordinate real general\n1 1 0\n", oracle.IO),
    (b"%%MatrixMarket matrix sparse real general\n1 1 0\n", oracle.IO),
    (b"%%MatrixMarket matrix coordinate double general\n1 1 0\n", oracle.IO),
    (b"%%MatrixMarket matrix coordinate real upper\n1 1 0\n", oracle.IO),
    (b"%%MatrixMarket matrix array real general\n1 1\n1\n", oracle.UNSUPPORTED),
    (b"%%MatrixMarket matrix coordinate pattern general\n1 1 1\n1 1\n", oracle.UNSUPPORTED),
    (b"%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1 0\n", oracle.UNSUPPORTED),
    (b"%%MatrixMarket matrix coordinate real hermitian\n1 1 0\n", oracle.BAD_ARG),
    (b"%%MatrixMarket matrix array pattern general\n1 1\n", oracle.BAD_ARG),
    (b"%%MatrixMarket matrix coordinate pattern skew-symmetric\n1 1 0\n", oracle.BAD_ARG),
    (b"%%MatrixMarket matrix coordinate real general\n2 2 1\n0 1 1\n", oracle.INDEX),
    (b"%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n", oracle.INDEX),
    (b"%%MatrixMarket matrix coordinate real general\n2 2 1\n1 3 1\n", oracle.BAD_ARG),
    (b"%%MatrixMarket matrix coordinate real symmetric\n2 3 1\n1 3 1\n", oracle.INDEX),
    (b"%%MatrixMarket matrix coordinate real general\n2 -2 0\n", oracle.BAD_ARG),
])
def test_errors(text, code):
    assert oracle.mm_matrix(text)[0] == code


def test_header_is_case_insensitive_and_extra_data_ignored():
    # runs of spaces between header words are one separator (split(" +"))
    text = b"%%matrixmarket MATRIX  Coordinate REAL General  \n1 2 1\n1 2 -3\n1 1 99 trailing words\n"
    A, _ = _csr(oracle.mm_matrix(text))
    assert A.shape == (1, 2) and A.toarray().tolist() == [[0, -3.0]]


def test_vectors():
    code, x = oracle.mm_vector(b"%%MatrixMarket vector coordinate real\n4 3\n2 1.5\n4 -2\n2 7\n")
    assert code == oracle.OK and x.tolist() == [0, 7.0, 0, -2.0]
    code, x = oracle.mm_vector(b"%%MatrixMarket vector array integer\n3\n1\n2 3e0\n")
    assert code == oracle.OK and x.tolist() == [1.0, 2.0, 3.0]
    code, x = oracle.mm_vector(b"2 1\n2 5\n")  # no header: coordinate real
    assert code == oracle.OK and x.tolist() == [0, 5.0]
    assert oracle.mm_vector(b"%%MatrixMarket vector coordinate pattern\n2 1\n1\n")[0] == oracle.UNSUPPORTED
    assert oracle.mm_vector(b"%%MatrixMarket vector array complex\n1\n1 1\n")[0] == oracle.UNSUPPORTED
    assert oracle.mm_vector(b"%%MatrixMarket vector array pattern\n1\n")[0] == oracle.BAD_ARG
    assert oracle.mm_vector(b"%%MatrixMarket vector coordinate real\n2 1\n3 1\n")[0] == oracle.INDEX
    assert oracle.mm_vector(b"%%MatrixMarket vector array real\n3\n1 2\n")[0] == oracle.IO
    assert oracle.mm_vector(b"%%MatrixMarket matrix array real\n3\n1 2 3\n")[0] == oracle.IO
