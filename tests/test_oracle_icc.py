"""The oracle's restatement of S/ICC.java, pinned without running MTJ:

* the reference's own ICCTest systems (tri- and penta-diagonal, n in 1..10,
  ICCTest.java / IncompleteFactorizationTestAbstract.java), where no fill is
  dropped, so the factor is the exact Cholesky factor (numpy) and apply
  solves the system (the reference's 1e-5 bar, here far tighter);
* full patterns, where the factor is again the exact Cholesky factor;
* patterns that drop fill, through a closed-form characterization of MTJ's
  row recurrence: row k of R is the up-looking Cholesky row computed against
  the already stored (fill-dropped) rows, i.e. with z solving
  R[:k,:k]' z = A[k,:k] (scipy's triangular solve), R[k,c] =
  (A[k,c] - z.R[:k,c]) / sqrt(A[k,k] - z.R[:k,k]) on row k's pattern.  The
  textbook IC(0), which drops the fill inside the row as well, differs from
  this by 5e-2 on the 2-D Poisson matrix, so the check tells the two apart;
* apply against scipy's triangular solves with the stored factor;
* MTJ's error and NaN behaviour, and CG with ICC on the reference's random
  SPD systems (CGICCTest), solution to 1e-4.
"""
import numpy as np
import pytest
import scipy.linalg as sl
import scipy.sparse as sp

import oracle
from mtjgpu import problems as P


def _ocsr(m):
    m = sp.csr_matrix(m)
    m.sort_indices()
    n = m.shape[0]
    return oracle.Csr(n, n, m.indptr.astype(np.int32), m.indices.astype(np.int32), m.data)


def _factor_dense(O, Mo):
    """The stored upper factor R as a dense matrix (zeros off its pattern)."""
    v = Mo.ilu_values()
    n = O.nr
    R = np.zeros((n, n))
    for i in range(n):
        for e in range(O.rp[i], O.rp[i + 1]):
            if O.ci[e] >= i:
                R[i, O.ci[e]] = v[e]
    return R


def _band(n, diag, offs):
    d = np.diag(np.full(n, float(diag)))
    for k, v in enumerate(offs, start=1):
        if k < n:
            d += np.diag(np.full(n - k, float(v)), k) + np.diag(np.full(n - k, float(v)), -k)
    return d


@pytest.mark.parametrize("diag,offs", [(2, [-1]), (4, [-1, -1])])
@pytest.mark.parametrize("n", range(1, 11))
def test_reference_band_systems_are_exact_cholesky(n, diag, offs):
    d = _band(n, diag, offs)
    O = _ocsr(d)
    Mo = oracle.Precond(oracle.PREC_ICC, O)
    assert Mo.code == oracle.OK
    R = _factor_dense(O, Mo)
    np.testing.assert_allclose(R, np.linalg.cholesky(d).T, rtol=0, atol=1e-14 * diag)
    x = np.ones(n)
    b = d @ x
    xs = Mo.apply(b)
    assert np.linalg.norm(d @ xs - b) < 1e-12  # ICCTest: 1e-5
    np.testing.assert_array_equal(Mo.trans_apply(b), xs)


@pytest.mark.parametrize("seed", range(6))
def test_full_pattern_is_exact_cholesky(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 30))
    g = rng.standard_normal((n, n))
    d = g @ g.T + n * np.eye(n)
    O = oracle.Csr(n, n, np.arange(0, n * n + 1, n, dtype=np.int32),
                   np.tile(np.arange(n, dtype=np.int32), n), d.ravel())
    Mo = oracle.Precond(oracle.PREC_ICC, O)
    assert Mo.code == oracle.OK
    R = _factor_dense(O, Mo)
    L = np.linalg.cholesky(d)
    np.testing.assert_allclose(R, L.T, rtol=0, atol=1e-12 * np.abs(L).max())


def _dropping_matrices():
    yield "poisson2d_8", P.poisson2d(8)
    yield "varcoef3d_5", P.varcoef_poisson(3, 5, seed=3)
    rng = np.random.default_rng(11)
    n = 60
    d = sp.random(n, n, density=0.06, random_state=11, format="csr")
    d = d + d.T + sp.diags(rng.random(n) + 8.0)
    d = sp.csr_matrix(d)
    d.sort_indices()
    yield "random_spd_60", (n, n, d.indptr, d.indices, d.data)


@pytest.mark.parametrize("name,sysm", list(_dropping_matrices()))
def test_row_recurrence_characterization(name, sysm):
    n, _, rp, ci, va = sysm
    O = oracle.Csr(n, n, rp, ci, va)
    Mo = oracle.Precond(oracle.PREC_ICC, O)
    assert Mo.code == oracle.OK
    A = sp.csr_matrix((va, ci, rp), shape=(n, n)).toarray()
    R = _factor_dense(O, Mo)
    worst = 0.0
    for k in range(n):
        z = sl.solve_triangular(R[:k, :k], A[k, :k], trans="T", lower=False) if k else np.zeros(0)
        rk = A[k, :] - z @ R[:k, :]
        assert rk[k] > 0
        row = rk / np.sqrt(rk[k])
        cols = [c for c in ci[rp[k]:rp[k + 1]] if c >= k]
        worst = max(worst, np.abs(row[cols] - R[k, cols]).max() / max(1.0, np.abs(row[cols]).max()))
    assert worst < 1e-13, worst
    # apply is R' y = b then R x = y with the stored factor
    b = P.uniform(n, 3)
    y = sl.solve_triangular(R, b, trans="T", lower=False)
    x = sl.solve_triangular(R, y, lower=False)
    np.testing.assert_allclose(Mo.apply(b), x, rtol=1e-12, atol=1e-12 * np.abs(x).max())


def test_textbook_ic0_differs_where_fill_is_dropped():
    n, _, rp, ci, va = P.poisson2d(8)
    O = oracle.Csr(n, n, rp, ci, va)
    R = _factor_dense(O, oracle.Precond(oracle.PREC_ICC, O))
    A = sp.csr_matrix((va, ci, rp), shape=(n, n)).toarray()
    T = np.zeros((n, n))
    for k in range(n):
        for c in range(k, n):
            if A[k, c] != 0:
                s = A[k, c] - T[:k, k] @ T[:k, c]
                T[k, c] = np.sqrt(s) if c == k else s / T[k, k]
    assert np.abs(T - R).max() > 1e-2


def test_errors_and_nan():
    # missing diagonal (S/ICC.java:141-155)
    m = sp.csr_matrix(np.array([[1.0, 1.0], [1.0, 0.0]]))
    m.eliminate_zeros()
    Mo = oracle.Precond(oracle.PREC_ICC, _ocsr(m))
    assert Mo.code == oracle.MISSING_DIAGONAL and Mo.bad.value == 1
    # Rk[k] == 0 before the square root (S/ICC.java:127-130)
    O = oracle.Csr(2, 2, np.array([0, 2, 4], np.int32), np.array([0, 1, 0, 1], np.int32),
                   np.array([1.0, 1.0, 1.0, 1.0]))
    Mo = oracle.Precond(oracle.PREC_ICC, O)
    assert Mo.code == oracle.ZERO_DIAGONAL and Mo.bad.value == 1
    O = oracle.Csr(2, 2, np.array([0, 2, 4], np.int32), np.array([0, 1, 0, 1], np.int32),
                   np.array([0.0, 1.0, 1.0, 1.0]))
    Mo = oracle.Precond(oracle.PREC_ICC, O)
    assert Mo.code == oracle.ZERO_DIAGONAL and Mo.bad.value == 0
    # a negative pivot is not an error: sqrt gives NaN, as Math.sqrt does
    O = oracle.Csr(2, 2, np.array([0, 2, 4], np.int32), np.array([0, 1, 0, 1], np.int32),
                   np.array([1.0, 2.0, 2.0, 1.0]))
    Mo = oracle.Precond(oracle.PREC_ICC, O)
    assert Mo.code == oracle.OK
    v = Mo.ilu_values()
    assert v[0] == 1.0 and v[1] == 2.0 and np.isnan(v[3])
    assert np.isnan(Mo.apply(np.ones(2))).all()


def test_nan_pivot_reaches_every_later_row_through_its_upper_entries():
    """MTJ scans every earlier row: 0 / NaN is NaN, so a row with a NaN pivot
    subtracts NaN at each of its upper columns in every later row, also rows
    with no structural path to it."""
    n = 6
    d = np.eye(n) * 4.0
    d[0, 0] = 1.0
    d[0, 1] = d[1, 0] = 3.0   # row 1's pivot is 4 - 3 * 3 < 0: R_11 is NaN
    d[1, 5] = d[5, 1] = 1.0   # row 1's upper entry at column 5
    d[3, 5] = d[5, 3] = 0.5
    O = _ocsr(d)
    Mo = oracle.Precond(oracle.PREC_ICC, O)
    assert Mo.code == oracle.OK
    R = _factor_dense(O, Mo)
    assert np.isnan(R[1, 1])
    # rows 2..4 have no path to row 1, yet row 3 stores column 5: NaN
    assert R[2, 2] == 2.0 and R[4, 4] == 2.0
    assert np.isnan(R[3, 5])
    assert np.isnan(R[5, 5])


def _mtj_random_spd(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 50))
    k = int(rng.integers(0, min(10, n) + 1))
    d = np.zeros((n, n))
    for i in range(n):
        cols = rng.choice(n, size=k, replace=False) if k else []
        d[i, cols] = rng.random(len(cols))
    d = d + d.T
    while True:
        d[np.arange(n), np.arange(n)] += 100.0
        if np.linalg.eigvalsh(d)[0] > 0:
            break
    return d


@pytest.mark.parametrize("seed", range(8))
def test_cg_with_icc_reference_style(seed):
    d = _mtj_random_spd(seed)
    n = d.shape[0]
    rng = np.random.default_rng(300 + seed)
    xt = rng.random(n)
    b = d @ xt
    O = _ocsr(d)
    Mp = oracle.Precond(oracle.PREC_ICC, O)
    assert Mp.code == oracle.OK
    code, x, it, _ = oracle.solve("cg", O, b, rng.random(n), Mp, oracle.monitor(100000, 1e-5))
    assert code == oracle.OK
    np.testing.assert_allclose(x, xt, atol=1e-4, rtol=0)


def test_icc_cuts_cg_iterations_on_poisson():
    sysm = P.poisson2d(32)
    n = sysm[0]
    O = oracle.Csr(*sysm)
    b = O.mult(P.uniform(n, 1))
    _, _, it_j, _ = oracle.solve("cg", O, b, np.zeros(n), oracle.Precond(oracle.PREC_DIAGONAL, O),
                                 oracle.monitor(10000, 1e-8))
    code, _, it_c, _ = oracle.solve("cg", O, b, np.zeros(n), oracle.Precond(oracle.PREC_ICC, O),
                                    oracle.monitor(10000, 1e-8))
    assert code == oracle.OK
    assert it_c < 0.6 * it_j
