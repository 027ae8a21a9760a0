"""GPU parity of the CompRowMatrix products (S/CompRowMatrix.java:294-388)
through the C ABI against the oracle, for each device format: the automatic
choice, the row-block CSR kernels only (MTJG_SPMV_FORMAT=csr), the SELL-64
copy for every matrix (MTJG_SPMV_FORMAT=sell) and the XCD column blocks in
front of the row-block kernels for every matrix (MTJG_SPMV_FORMAT=xcd).

Bar: bitwise equality wherever the kernel sums a row sequentially (every row
of the SELL-64 kernels; rows in row-blocks of >= 128 rows for the CSR
kernels, so every stencil matrix); otherwise per row
|y_gpu - y_ref| <= 1e-14 * sum_j |a_ij x_j| (only the summation order
differs)."""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

import mtjgpu as M
import oracle
from mtjgpu import problems as P

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
TOL = 1e-14


@pytest.fixture(params=["auto", "csr", "sell", "xcd"], autouse=True)
def spmv_format(request, monkeypatch):
    # read when a matrix is created
    monkeypatch.setenv("MTJG_SPMV_FORMAT", request.param)
    return request.param


def _sell_forced():
    return os.environ.get("MTJG_SPMV_FORMAT") == "sell"


def _abs_rowsum(rp, ci, va, x):
    return np.add.reduceat(np.abs(va * x[ci]), rp[:-1]) * (np.diff(rp) > 0) if len(va) else np.zeros(len(rp) - 1)


def _check(y, ref, rp, ci, va, x, exact):
    if exact or _sell_forced():
        np.testing.assert_array_equal(y, ref)
    else:
        bound = TOL * _abs_rowsum(rp, ci, va, x) + 1e-300
        err = np.abs(y - ref)
        assert np.all(err <= bound), (err / bound).max()


def _pair(ctx, sysm):
    n, m, rp, ci, va = sysm
    return M.CompRowMatrix(n, m, rp, ci, va, ctx=ctx), oracle.Csr(n, m, rp, ci, va)


def test_io_fixture_3x3(ctx):
    fx = json.load(open(os.path.join(GOLD, "mtj_fixtures.json")))["io_3x3"]
    A = M.CompRowMatrix(3, 3, fx["row_pointer"], fx["column_index"], fx["data"], ctx=ctx)
    x = M.DenseVector(np.array([1.0, 2.0, 3.0]), ctx)
    y = M.DenseVector(3, ctx)
    np.testing.assert_array_equal(A.mult(x, y).getData(), [8.0, 15.0, 10.0])
    np.testing.assert_array_equal(A.transMult(x, y).getData(), [6.0, 16.0, 10.0])
    y.upload([1.0, 1.0, 1.0])
    np.testing.assert_array_equal(A.multAdd(2.0, x, y).getData(), [17.0, 31.0, 21.0])
    assert A.get(0, 1) == 1.0 and A.get(2, 2) == 0.0
    with pytest.raises(M.IndexOutOfBoundsException):
        A.set(2, 2, 1.0)
    A.set(2, 1, 7.0)
    assert A.get(2, 1) == 7.0


@pytest.mark.parametrize("name", ["p2d_256", "p3d_40", "varcoef3d_24", "convdiff_24"])
def test_stencils_bitwise(ctx, name):
    sysm = {"p2d_256": lambda: P.poisson2d(256), "p3d_40": lambda: P.poisson3d(40),
            "varcoef3d_24": lambda: P.varcoef_poisson(3, 24, seed=2),
            "convdiff_24": lambda: P.convdiff3d(24, seed=3)}[name]()
    A, O = _pair(ctx, sysm)
    n = sysm[0]
    xh = P.uniform(n, 11)
    yh = P.uniform(n, 12)
    x = M.DenseVector(xh, ctx)
    y = M.DenseVector(n, ctx)
    np.testing.assert_array_equal(A.mult(x, y).getData(), O.mult(xh))
    y.upload(yh)
    np.testing.assert_array_equal(A.multAdd(-1.0, x, y).getData(), O.mult_add(-1.0, xh, yh))
    y.upload(yh)
    np.testing.assert_array_equal(A.transMult(x, y).getData(), O.trans_mult(xh))
    y.upload(yh)
    np.testing.assert_array_equal(A.transMultAdd(0.75, x, y).getData(), O.trans_mult_add(0.75, xh, yh))
    np.testing.assert_array_equal(A.mult(2.5, x, y).getData(), O.mult_add(2.5, xh, np.zeros(n)))


@pytest.mark.parametrize("n,seed", [(20000, 0), (3000, 1)])
def test_powerlaw_rows(ctx, n, seed):
    sysm = P.powerlaw(n, seed=seed)
    _, _, rp, ci, va = sysm
    A, O = _pair(ctx, sysm)
    xh = P.uniform(n, 5)
    x = M.DenseVector(xh, ctx)
    y = M.DenseVector(n, ctx)
    _check(A.mult(x, y).getData(), O.mult(xh), rp, ci, va, xh, exact=False)
    yh = P.uniform(n, 6)
    y.upload(yh)
    got = A.multAdd(1.5, x, y).getData()
    ref = O.mult_add(1.5, xh, yh)
    if _sell_forced():
        np.testing.assert_array_equal(got, ref)
    assert np.all(np.abs(got - ref) <= 1.5 * TOL * _abs_rowsum(rp, ci, va, xh) + 1e-15 * np.abs(ref) + 1e-300)
    # transpose: compare to the transposed oracle bound
    m = sp.csr_matrix((va, ci, rp), shape=(n, n))
    t = m.T.tocsr()
    t.sort_indices()
    got = A.transMult(x, y).getData()
    _check(got, O.trans_mult(xh), t.indptr, t.indices, t.data, xh, exact=False)


def test_long_rows_and_empty_rows(ctx):
    rng = np.random.default_rng(3)
    n = 6000
    lens = np.zeros(n, int)
    lens[::7] = rng.integers(1, 40, size=len(lens[::7]))
    lens[5] = 5000   # longer than one staged row-block
    lens[77] = 2048  # exactly one row-block
    lens[78] = 2049
    rows = []
    for i, L in enumerate(lens):
        rows.append(np.sort(rng.choice(n, size=L, replace=False)))
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    ci = np.concatenate(rows).astype(np.int32)
    va = rng.standard_normal(len(ci))
    A, O = _pair(ctx, (n, n, rp, ci, va))
    xh = rng.standard_normal(n)
    x = M.DenseVector(xh, ctx)
    y = M.DenseVector(np.full(n, 9.0), ctx)
    got = A.mult(x, y).getData()
    _check(got, O.mult(xh), rp, ci, va, xh, exact=False)
    assert np.all(got[lens == 0] == 0.0)


def test_rectangular_and_tiny(ctx):
    for (nr, nc) in [(1, 1), (1, 5), (7, 3), (3, 700)]:
        rng = np.random.default_rng(nr * 100 + nc)
        d = rng.standard_normal((nr, nc)) * (rng.random((nr, nc)) < 0.5)
        m = sp.csr_matrix(d)
        m.sort_indices()
        A, O = _pair(ctx, (nr, nc, m.indptr, m.indices, m.data))
        xh = rng.standard_normal(nc)
        # fewer than 128 rows: rows are reduced by lane groups (tolerance bar)
        _check(A.mult(M.DenseVector(xh, ctx), M.DenseVector(nr, ctx)).getData(), O.mult(xh),
               m.indptr, m.indices, m.data, xh, exact=False)
        xt = rng.standard_normal(nr)
        t = m.T.tocsr()
        t.sort_indices()
        _check(A.transMult(M.DenseVector(xt, ctx), M.DenseVector(nc, ctx)).getData(), O.trans_mult(xt),
               t.indptr, t.indices, t.data, xt, exact=False)
    with pytest.raises(M.IndexOutOfBoundsException):
        A.mult(M.DenseVector(3, ctx), M.DenseVector(3, ctx))


def test_unsorted_input_is_sorted(ctx):
    rp = [0, 3, 5]
    ci = [2, 0, 1, 1, 0]
    va = [3.0, 1.0, 2.0, 5.0, 4.0]
    A = M.CompRowMatrix(2, 3, rp, ci, va, ctx=ctx)
    r, c, v = A.arrays()
    np.testing.assert_array_equal(c, [0, 1, 2, 0, 1])
    np.testing.assert_array_equal(v, [1.0, 2.0, 3.0, 4.0, 5.0])
    with pytest.raises(M.IllegalArgumentException):
        M.CompRowMatrix(2, 3, [0, 2, 2], [1, 1], [1.0, 2.0], ctx=ctx)
    with pytest.raises(M.IllegalArgumentException):
        M.CompRowMatrix(2, 3, [0, 1, 2], [0, 3], [1.0, 2.0], ctx=ctx)


def test_large_closed_form_3d_128(ctx):
    """Size-independent check at a large size: A*1 for the 3-D Laplacian is
    the count of missing neighbours, exactly."""
    n1 = 128
    sysm = P.poisson3d(n1)
    A, _ = _pair(ctx, sysm)
    n = sysm[0]
    y = A.mult(M.DenseVector(np.ones(n), ctx), M.DenseVector(n, ctx)).getData()
    idx = np.arange(n)
    missing = sum(((c == 0).astype(int) + (c == n1 - 1).astype(int))
                  for c in (idx % n1, (idx // n1) % n1, idx // (n1 * n1)))
    np.testing.assert_array_equal(y, missing.astype(float))


def test_run_to_run_deterministic(ctx):
    sysm = P.powerlaw(30000, seed=2)
    A, _ = _pair(ctx, sysm)
    xh = P.uniform(30000, 1)
    x = M.DenseVector(xh, ctx)
    a = A.mult(x, M.DenseVector(30000, ctx)).getData()
    for _ in range(3):
        np.testing.assert_array_equal(A.mult(x, M.DenseVector(30000, ctx)).getData(), a)


def test_value_updates_reach_every_copy(ctx):
    # setValues / set change the CSR values; the SELL-64 copy (and the FP32
    # copies) must follow before the next product
    sysm = P.varcoef_poisson(3, 14, 12, 10, seed=7)
    n, m, rp, ci, va = sysm
    A, _ = _pair(ctx, sysm)
    xh = P.uniform(m, 2)
    x = M.DenseVector(xh, ctx)
    y = M.DenseVector(n, ctx)
    A.mult(x, y)
    va2 = va * 1.5 + 0.25
    A.setValues(va2)
    _check(A.mult(x, y).getData(), oracle.Csr(n, m, rp, ci, va2).mult(xh), rp, ci, va2, xh,
           exact=A.format()[0] == "sell")
    A.set(5, int(ci[rp[5]]), -3.0)
    va3 = va2.copy()
    va3[rp[5]] = -3.0
    _check(A.mult(x, y).getData(), oracle.Csr(n, m, rp, ci, va3).mult(xh), rp, ci, va3, xh,
           exact=A.format()[0] == "sell")


@pytest.mark.parametrize("gen", ["powerlaw_20k", "long_rows", "rect_3x700"])
def test_xcd_blocks_bitwise_row_block(ctx, monkeypatch, gen):
    """The XCD column-block products form each product val[j] * x[col[j]] in a
    pass of their own and the row-block kernels sum them in the order they sum
    the products they form: every result is bitwise the row-block one's."""
    rng = np.random.default_rng(9)
    if gen == "powerlaw_20k":
        sysm = P.powerlaw(20000, seed=4)
    elif gen == "long_rows":
        n = 5000
        lens = rng.integers(0, 12, size=n)
        lens[17], lens[900] = 4100, 2049
        rows = [np.sort(rng.choice(n, size=L, replace=False)) for L in lens]
        rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
        sysm = (n, n, rp, np.concatenate(rows).astype(np.int32), rng.standard_normal(int(rp[-1])))
    else:
        m = sp.random(3, 700, density=0.4, random_state=5, format="csr")
        m.sort_indices()
        sysm = (3, 700, m.indptr, m.indices, m.data)
    nr, nc = sysm[0], sysm[1]
    mats = {}
    for fmt in ("csr", "xcd"):
        monkeypatch.setenv("MTJG_SPMV_FORMAT", fmt)
        mats[fmt] = M.CompRowMatrix(*sysm, ctx=ctx)
    assert mats["csr"].format()[0] == "csr" and mats["xcd"].format()[0] == "csr_xcd"
    xh, xt = rng.standard_normal(nc), rng.standard_normal(nr)
    yh, yt = rng.standard_normal(nr), rng.standard_normal(nc)
    out = {}
    for fmt, A in mats.items():
        x, y = M.DenseVector(xh, ctx), M.DenseVector(nr, ctx)
        r = [A.mult(x, y).getData()]
        y.upload(yh)
        r.append(A.multAdd(-0.5, x, y).getData())
        xT, yT = M.DenseVector(xt, ctx), M.DenseVector(yt, ctx)
        r.append(A.transMultAdd(1.25, xT, yT).getData())
        r.append(A.transMult(xT, yT).getData())
        out[fmt] = r
        for o in (x, y, xT, yT, A):
            o.close()
    for a, b in zip(out["csr"], out["xcd"]):
        np.testing.assert_array_equal(a, b)
