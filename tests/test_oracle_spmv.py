"""Pins the oracle's CompRowMatrix products (S/CompRowMatrix.java:294-388)
against facts that do not come from its own formula: the reference test's
known 3x3 matrix, closed forms of the Laplacian stencils, scipy's CSR product
(sequential per-row sums, so bitwise equal when compiled without FMA), and
dense brute force."""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from mtjgpu import problems as P

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _csr(m):
    m = sp.csr_matrix(m)
    m.sort_indices()
    return oracle.Csr(m.shape[0], m.shape[1], m.indptr, m.indices, m.data)


def test_mtj_io_fixture_3x3():
    # T/io/MatrixVectorIoTest.java:37-44: pattern {{1,2},{0,2},{1}}, values 1..5
    with open(os.path.join(GOLD, "mtj_fixtures.json")) as f:
        fx = json.load(f)["io_3x3"]
    A = oracle.Csr(3, 3, fx["row_pointer"], fx["column_index"], fx["data"])
    x = np.array([1.0, 2.0, 3.0])
    np.testing.assert_array_equal(A.mult(x), [8.0, 15.0, 10.0])
    np.testing.assert_array_equal(A.trans_mult(x), [6.0, 16.0, 10.0])
    assert A.get(0, 1) == 1.0 and A.get(1, 2) == 4.0 and A.get(2, 2) == 0.0
    np.testing.assert_array_equal(A.mult_add(2.0, x, [1.0, 1.0, 1.0]), [17.0, 31.0, 21.0])


@pytest.mark.parametrize("dim,n", [(2, 17), (3, 9)])
def test_laplacian_closed_forms(dim, n):
    gen = P.poisson2d if dim == 2 else P.poisson3d
    N, _, rp, ci, va = gen(n)
    A = oracle.Csr(N, N, rp, ci, va)
    # A*1 = number of missing neighbours (Dirichlet by omission)
    y = A.mult(np.ones(N))
    idx = np.arange(N)
    coords = [idx % n, (idx // n) % n] + ([idx // (n * n)] if dim == 3 else [])
    missing = sum((c == 0).astype(int) + (c == n - 1).astype(int) for c in coords)
    np.testing.assert_array_equal(y, missing.astype(float))
    # A e_j = column j
    j = N // 2 + 1
    e = np.zeros(N)
    e[j] = 1.0
    col = A.mult(e)
    dense = np.zeros((N, N))
    for i in range(N):
        dense[i, ci[rp[i]:rp[i + 1]]] = va[rp[i]:rp[i + 1]]
    np.testing.assert_array_equal(col, dense[:, j])
    # symmetric: A'x = Ax exactly (same products, same order per row)
    x = P.uniform(N, 3)
    np.testing.assert_array_equal(A.trans_mult(x), A.mult(x))


def test_nnz_counts_match_formula():
    N, _, rp, ci, va = P.poisson2d(256)
    assert N == 65536 and len(va) == 5 * 65536 - 4 * 256 == 326656
    N, _, rp, ci, va = P.poisson3d(32)
    assert len(va) == 7 * 32 ** 3 - 6 * 32 ** 2


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_against_scipy_bitwise(seed):
    rng = np.random.default_rng(seed)
    m = sp.random(300, 250, density=0.03, random_state=seed, format="csr")
    m.data = rng.standard_normal(m.nnz)
    m.sort_indices()
    A = _csr(m)
    x = rng.standard_normal(250)
    np.testing.assert_array_equal(A.mult(x), m @ x)
    xt = rng.standard_normal(300)
    # transpose scatter adds rows in ascending order: equals the sequential
    # sum over the CSC (ascending row) entries, i.e. scipy's m.T @ x
    np.testing.assert_allclose(A.trans_mult(xt), m.T @ xt, rtol=1e-13, atol=1e-13)


def test_brute_force_dense_small():
    rng = np.random.default_rng(7)
    for n in range(1, 20):
        d = rng.standard_normal((n, n)) * (rng.random((n, n)) < 0.4)
        A = _csr(d)
        x = rng.standard_normal(n)
        ref = np.array([sum(d[i, j] * x[j] for j in range(n) if d[i, j] != 0) for i in range(n)])
        np.testing.assert_allclose(A.mult(x), ref, rtol=1e-14, atol=1e-14)


def test_trans_mult_add_alpha_rules():
    m = sp.csr_matrix(np.array([[1.0, 2.0], [0.0, 3.0]]))
    A = _csr(m)
    x = np.array([1.0, 1.0])
    y = np.array([4.0, 8.0])
    # exact integers: alpha*((y/alpha) + A'x) = y + alpha*A'x
    np.testing.assert_array_equal(A.trans_mult_add(2.0, x, y), y + 2.0 * (m.T @ x))
    # DenseVector.scale has no alpha == 0 shortcut: y * (1/0) * 0 is NaN
    assert np.all(np.isnan(A.trans_mult_add(0.0, x, y)))


def test_mult_empty_rows_and_nan():
    m = sp.csr_matrix(np.array([[0.0, 0.0], [1.0, 0.0]]))
    A = _csr(m)
    np.testing.assert_array_equal(A.mult(np.array([np.nan, 1.0])), [0.0, np.nan])
