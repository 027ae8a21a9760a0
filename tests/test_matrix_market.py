"""The library's Matrix Market readers (mtjg_mm_*, host code) against the
oracle's restatement of CompRowMatrix(MatrixVectorReader) and
DenseVector(MatrixVectorReader): same CSR arrays (values bitwise) or the same
exception class, on the oracle's pinned cases and on files large enough to
take the multi-threaded path (newline-aligned chunks, global token numbering,
parallel row assembly).  Parsing needs no GPU; the device upload is checked in
the gpu-marked test at the end."""
import numpy as np
import pytest

import mtjgpu as M
import oracle
from mtjgpu import matrix_market as MM
from mtjgpu import problems as P

from test_oracle_matrix_market import _random_file

EXC = {oracle.IO: M.IOException, oracle.UNSUPPORTED: M.UnsupportedOperationException,
       oracle.BAD_ARG: M.IllegalArgumentException, oracle.INDEX: M.IndexOutOfBoundsException}

SMALL = [
    b"%%MatrixMarket matrix coordinate real general\n2 2 4\n1 1 5\n1 1 7\n2 2 0\n1 1 9\n",
    b"%%MatrixMarket matrix coordinate real symmetric\n2 2 3\n2 1 1.5\n1 2 2.5\n1 1 4\n",
    b"%%MatrixMarket matrix coordinate real skew-symmetric\n3 3 2\n2 1 1.5\n3 2 -2\n",
    b"% no header\r\n3 3\n  2 %c\n1\n1.0 5%c\r\n3\t2\x0b0x1.8p1\n",
    b"%%MatrixMarket matrix coordinate real general\n3.9 3 2\n2.7 3e0 1\n1 1.99 2\n",
    b"%%matrixmarket MATRIX  Coordinate REAL General  \n1 2 1\n1 2 -3\n1 1 99 trailing words\n",
    b"%%MatrixMarket matrix coordinate real general\n0 0 0\n",
    b"\n1 1 1\n1 1 2\n",
    b"\r\n\r\n1 1 1\r\n1 1 2\r\n",
    b"%%MatrixMarket matrix coordinate real general\r1 1 1\r1 1 -0.0\r",
    b"%%MatrixMarket matrix coordinate real general\n1 1 1\n1 1 1.5d\n",
    b"%%MatrixMarket matrix coordinate real general\n1 1 1\n1 1 -.5e1f\n",
    b"%%MatrixMarket matrix coordinate real general\n1 1 1\n1 1 0X1P-2D\n",
    b"%%MatrixMarket matrix coordinate real general\n1 1 1\n1 1 1e400\n",
    b"%%MatrixMarket matrix coordinate real general\n1 1 1\n1 1 4.9e-324\n",
    # errors
    b"",
    b"%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1\n",
    b"%%MatrixMarket matrix coordinate real\n2 2 0\n",
    b"%%MatrixMarket matrix\tcoordinate real general x\n",
    b"%%MatrixMarket vector coordinate real general\n1 1 0\n",
    b"%%MatrixMarket matrix sparse real general\n1 1 0\n",
    b"%%MatrixMarket matrix array real general\n1 1\n1\n",
    b"%%MatrixMarket matrix coordinate pattern general\n1 1 1\n1 1\n",
    b"%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1 0\n",
    b"%%MatrixMarket matrix coordinate real hermitian\n1 1 0\n",
    b"%%MatrixMarket matrix coordinate pattern skew-symmetric\n1 1 0\n",
    b"%%MatrixMarket matrix coordinate real general\n2 2 1\n0 1 1\n",
    b"%%MatrixMarket matrix coordinate real general\n2 2 1\n1 3 1\n",
    b"%%MatrixMarket matrix coordinate real symmetric\n2 3 1\n1 3 1\n",
    b"%%MatrixMarket matrix coordinate real general\n2 -2 0\n",
] + [("%%%%MatrixMarket matrix coordinate real general\n1 1 1\n1 1 %s\n" % t).encode()
     for t in ["nan", "NaN", "Infinity", "1e", "0x1.8", ".", "1,5", "--1", "1df", "0x1.8p", "1.", "+2"]]


def _same(text):
    code, r, c, rp, ci, va, sym = oracle.mm_matrix(text)
    if code != oracle.OK:
        with pytest.raises(EXC[code]):
            MM.read_matrix(text)
        return
    r2, c2, rp2, ci2, va2, sym2 = MM.read_matrix(text)
    assert (r2, c2, sym2) == (r, c, sym)
    np.testing.assert_array_equal(rp2, rp)
    np.testing.assert_array_equal(ci2, ci)
    assert va2.tobytes() == va.tobytes()


@pytest.mark.parametrize("k", range(len(SMALL)))
def test_small_cases_match_oracle(k):
    _same(SMALL[k])


@pytest.mark.parametrize("case", [
    (1, 40, 40, 0.1, "general", "real"),
    (3, 45, 45, 0.1, "symmetric", "real"),
    (4, 45, 45, 0.1, "skew-symmetric", "real"),
    (5, 30, 30, 0.2, "general", "integer"),
])
def test_random_files_match_oracle(case):
    _same(_random_file(*case))


def _stencil_text(nx, ny, nz, style, seed=0):
    """A 3-D Laplacian written in one of several layouts (MTJ's writer,
    several entries per line with comments, CRLF, symmetric lower half with
    duplicates) -- all large enough for the threaded path."""
    n, _, rp, ci, va = P.poisson3d(nx, ny, nz)
    rng = np.random.default_rng(seed)
    va = va + rng.normal(size=va.size) * 1e-3
    rows = np.repeat(np.arange(n), np.diff(rp))
    if style == "writer":
        head = "%%%%MatrixMarket matrix coordinate real general\n%10d %10d %19d\n" % (n, n, va.size)
        body = "".join("%10d %10d % .12e\n" % (i + 1, j + 1, v) for i, j, v in zip(rows, ci, va))
        return (head + body).encode()
    if style == "messy":
        toks = []
        for t, (i, j, v) in enumerate(zip(rows, ci, va)):
            sep = " " if t % 3 else "\n% interleaved comment line\n"
            toks.append("%d %d %r%s" % (i + 1, j + 1, float(v), sep))
        return ("%%%%MatrixMarket matrix coordinate real general\r\n%d %d %d\r\n" % (n, n, va.size)
                + "".join(toks).replace("\n", "\r\n")).encode()
    if style == "symmetric_dups":
        low = rows >= ci
        r, c, v = rows[low], ci[low], va[low]
        extra = rng.integers(0, r.size, size=r.size // 10)
        r = np.concatenate([r, r[extra]])
        c = np.concatenate([c, c[extra]])
        v = np.concatenate([v, v[extra] * 2.0])
        head = "%%%%MatrixMarket matrix coordinate real symmetric\n%d %d %d\n" % (n, n, v.size)
        return (head + "".join("%d %d %r\n" % (i + 1, j + 1, float(x)) for i, j, x in zip(r, c, v))).encode()
    raise ValueError(style)


@pytest.mark.parametrize("style", ["writer", "messy", "symmetric_dups"])
def test_large_files_take_threaded_path_and_match_oracle(style):
    text = _stencil_text(24, 20, 22, style)
    assert len(text) > (1 << 20)
    _same(text)


def test_read_from_path_equals_parse(tmp_path):
    text = _stencil_text(16, 16, 16, "writer")
    f = tmp_path / "A.mtx"
    f.write_bytes(text)
    a = MM.read_matrix(str(f))
    b = MM.read_matrix(text)
    for x, y in zip(a[2:5], b[2:5]):
        np.testing.assert_array_equal(x, y)
    with pytest.raises(M.IOException):
        MM.read_matrix(str(tmp_path / "missing.mtx"))


def test_late_format_error_reported_in_order():
    # an error far into a large file is found by a worker thread; an EOF
    # (too few entries) is reported only when no token before it is bad
    text = _stencil_text(24, 20, 22, "writer")
    lines = text.split(b"\n")
    bad = list(lines)
    bad[-100] = bad[-100].replace(b"e", b"q", 1)
    _same(b"\n".join(bad))
    short = b"\n".join(lines[:-1000]) + b"\n"
    _same(short)


def test_vectors_match_oracle():
    cases = [b"%%MatrixMarket vector coordinate real\n4 3\n2 1.5\n4 -2\n2 7\n",
             b"%%MatrixMarket vector array integer\n3\n1\n2 3e0\n",
             b"2 1\n2 5\n",
             b"%%MatrixMarket vector coordinate pattern\n2 1\n1\n",
             b"%%MatrixMarket vector array complex\n1\n1 1\n",
             b"%%MatrixMarket vector array pattern\n1\n",
             b"%%MatrixMarket vector coordinate real\n2 1\n3 1\n",
             b"%%MatrixMarket vector array real\n3\n1 2\n",
             b"%%MatrixMarket matrix array real\n3\n1 2 3\n",
             ("%%MatrixMarket vector array real\n200000\n" + "".join("%r\n" % float(v) for v in
              np.random.default_rng(2).normal(size=200000))).encode()]
    for text in cases:
        code, x = oracle.mm_vector(text)
        if code != oracle.OK:
            with pytest.raises(EXC[code]):
                MM.read_vector(text)
            continue
        y = MM.read_vector(text)
        assert y.tobytes() == x.tobytes()


@pytest.mark.gpu
def test_device_matrix_from_file_multiplies_like_oracle(ctx, tmp_path):
    text = _stencil_text(20, 18, 16, "symmetric_dups")
    f = tmp_path / "S.mtx"
    f.write_bytes(text)
    code, r, c, rp, ci, va, sym = oracle.mm_matrix(text)
    A = MM.load_matrix(str(f), ctx)
    assert A.shape() == (r, c, len(ci))
    rp2, ci2, va2 = A.arrays()
    np.testing.assert_array_equal(rp2, rp)
    np.testing.assert_array_equal(ci2, ci)
    assert va2.tobytes() == va.tobytes()
    x = P.uniform(c, 9)
    O = oracle.Csr(r, c, rp, ci, va)
    y = M.DenseVector(np.zeros(r), ctx)
    A.mult(M.DenseVector(x, ctx), y)
    assert y.getData().tobytes() == O.mult(x).tobytes()
    vtext = ("%%MatrixMarket vector array real\n" + "%d\n" % c + "".join("%r\n" % float(v) for v in x)).encode()
    vf = tmp_path / "x.mtx"
    vf.write_bytes(vtext)
    xv = MM.load_vector(str(vf), ctx)
    assert xv.getData().tobytes() == x.tobytes()


def test_decimal_conversion_is_correctly_rounded():
    # the reader's exact fast path and its strtod fallback against Python's
    # float() (correctly rounded, like Java's parseDouble) on literals of
    # every digit count and exponent range
    rng = np.random.default_rng(17)
    vals = np.concatenate([rng.normal(size=4000) * 10.0 ** rng.integers(-30, 30, size=4000),
                           rng.integers(-10**6, 10**6, size=1000).astype(np.float64),
                           [0.0, -0.0, 1e22, 1e23, 9007199254740993.0, 5e-324, 1.7976931348623157e308]])
    toks = []
    for k, v in enumerate(vals):
        p = k % 19
        toks.append(["%.{}e".format(p), "%.{}f".format(min(p, 12)), "%r", "%.{}g".format(p + 1)][k % 4] % v
                    if k % 4 != 2 else repr(float(v)))
    toks += ["123456789012345678901234", "0.000000000000000000000000001", "-00012.5000", ".5", "5.",
             "9007199254740993", "4.94065645841246544e-324", "2.2250738585072011e-308"]
    text = ("%%%%MatrixMarket vector array real\n%d\n" % len(toks) + "\n".join(toks) + "\n").encode()
    x = MM.read_vector(text)
    want = np.array([float(t) for t in toks])
    assert x.tobytes() == want.tobytes()
    code, xo = oracle.mm_vector(text)
    assert code == oracle.OK and xo.tobytes() == want.tobytes()
