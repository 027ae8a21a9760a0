"""Pins the oracle's DenseVector norms, DiagonalPreconditioner, ILU(0) with
its triangular solves, GivensRotation, the dtrtrs back-substitution, the
DefaultIterationMonitor predicate and the CG / BiCGstab / GMRES / IR loops
against independent facts: the reference's own ILU known answers
(tri/penta-diagonal full-band ILU(0) is the exact LU), a dense
pattern-restricted ILU(0) written from Saad's Algorithm 10.4, scipy's
triangular and sparse solvers, closed forms and textbook invariants."""
import json
import os

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp

import oracle
from mtjgpu import problems as P

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _csr(m):
    m = sp.csr_matrix(m)
    m.sort_indices()
    return oracle.Csr(m.shape[0], m.shape[1], m.indptr, m.indices, m.data)


def _fixtures():
    with open(os.path.join(GOLD, "mtj_fixtures.json")) as f:
        return json.load(f)


# --------------------------------------------------------------------- vectors

def test_norms_closed_forms():
    x = np.array([3.0, -4.0])
    assert oracle.dv_norm(x, "two") == 5.0
    assert oracle.dv_norm(x, "one") == 7.0
    assert oracle.dv_norm(x, "inf") == 4.0
    assert oracle.dv_norm(x, "robust") == pytest.approx(5.0, rel=1e-15)
    big = np.full(4, 1e200)  # overflows the unscaled norm, not the robust one
    assert oracle.dv_norm(big, "two") == np.inf
    assert oracle.dv_norm(big, "robust") == pytest.approx(2e200, rel=1e-14)
    assert np.isnan(oracle.dv_norm(np.array([1.0, np.nan]), "inf"))
    rng = np.random.default_rng(0)
    v = rng.standard_normal(1000)
    assert oracle.dv_dot(v, v) == pytest.approx(np.dot(v, v), rel=1e-13)


# --------------------------------------------------------------- preconditioners

def test_jacobi_closed_form_and_errors():
    m = sp.csr_matrix(np.array([[4.0, 1.0, 0.0], [1.0, 2.0, 0.0], [0.0, 0.0, 8.0]]))
    A = _csr(m)
    M = oracle.Precond(oracle.PREC_DIAGONAL, A)
    assert M.code == oracle.OK
    np.testing.assert_array_equal(M.apply(np.array([4.0, 2.0, 8.0])), [1.0, 1.0, 1.0])
    z = _csr(np.array([[1.0, 0.0], [1.0, 0.0]]))  # row 1 diagonal not stored
    assert oracle.Precond(oracle.PREC_DIAGONAL, z).code == oracle.ZERO_DIAGONAL


def _band(n, diag, offs):
    d = np.diag(np.full(n, diag))
    for k, v in enumerate(offs, start=1):
        if k >= n:
            break
        d += np.diag(np.full(n - k, v), k) + np.diag(np.full(n - k, v), -k)
    return d


@pytest.mark.parametrize("name", ["ilu_tridiagonal", "ilu_pentadiagonal"])
@pytest.mark.parametrize("n", [1, 2, 3, 5, 9])
def test_ilu_full_band_is_exact_lu(name, n):
    # S/ILUTest.java:34-44 on the IncompleteFactorizationTestAbstract matrices
    fx = _fixtures()[name]
    d = _band(n, fx["diagonal"], fx["offdiagonal"])
    A = _csr(d)
    x = np.full(n, fx["x"])
    b = d @ x
    M = oracle.Precond(oracle.PREC_ILU, A)
    assert M.code == oracle.OK
    xs = M.apply(b)
    assert np.linalg.norm(b - d @ xs) <= fx["residual_tol"]
    np.testing.assert_allclose(xs, x, rtol=1e-12)


def _saad_ilu0(dense, mask):
    """Saad, Iterative Methods for Sparse Linear Systems (2nd ed.), Alg. 10.4."""
    a = dense.copy()
    n = a.shape[0]
    for i in range(1, n):
        for k in range(i):
            if not mask[i, k]:
                continue
            a[i, k] = a[i, k] / a[k, k]
            for j in range(k + 1, n):
                if mask[i, j]:
                    a[i, j] = a[i, j] - a[i, k] * a[k, j]
    return a


@pytest.mark.parametrize("seed", range(6))
def test_ilu_matches_dense_pattern_restricted(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(5, 40))
    mask = rng.random((n, n)) < 0.25
    mask = mask | mask.T | np.eye(n, dtype=bool)  # structurally symmetric
    d = np.where(mask, rng.standard_normal((n, n)), 0.0)
    d[np.arange(n), np.arange(n)] = 10.0 + np.abs(d).sum(1)
    A = _csr(d)
    M = oracle.Precond(oracle.PREC_ILU, A)
    assert M.code == oracle.OK
    ref = _saad_ilu0(d, mask)
    lu = M.ilu_values()
    got = np.zeros_like(d)
    for i in range(n):
        got[i, A.ci[A.rp[i]:A.rp[i + 1]]] = lu[A.rp[i]:A.rp[i + 1]]
    np.testing.assert_array_equal(got, np.where(mask, ref, 0.0))
    # apply = U^-1 L^-1 b with L unit lower, U upper (scipy triangular solves)
    Lm = np.tril(got, -1) + np.eye(n)
    Um = np.triu(got)
    b = rng.standard_normal(n)
    ref_x = sla.solve_triangular(Um, sla.solve_triangular(Lm, b, lower=True, unit_diagonal=True))
    np.testing.assert_allclose(M.apply(b), ref_x, rtol=1e-12, atol=1e-12)
    ref_t = sla.solve_triangular(Lm.T, sla.solve_triangular(Um.T, b, lower=True), unit_diagonal=True)
    np.testing.assert_allclose(M.trans_apply(b), ref_t, rtol=1e-12, atol=1e-12)


def test_ilu_errors():
    A = _csr(np.array([[1.0, 1.0], [1.0, 0.0]]) * np.array([[1, 1], [1, 0]]))
    assert oracle.Precond(oracle.PREC_ILU, A).code == oracle.MISSING_DIAGONAL
    # zero pivot: a00 = 0 used by row 1's elimination
    m = sp.csr_matrix(np.array([[1.0, 1.0], [1.0, 1.0]]))
    m.data[0] = 0.0
    A = oracle.Csr(2, 2, m.indptr, m.indices, m.data)
    assert oracle.Precond(oracle.PREC_ILU, A).code == oracle.ZERO_PIVOT


# ------------------------------------------------------------ GMRES helpers

@pytest.mark.parametrize("x,y", [(3.0, 4.0), (-3.0, 4.0), (1e-300, -2.0), (5.0, 0.0), (0.0, -7.0)])
def test_givens_rotation_zeroes_second_component(x, y):
    c, s = oracle.givens(x, y)
    assert c * c + s * s == pytest.approx(1.0, rel=1e-15)
    assert -s * x + c * y == pytest.approx(0.0, abs=1e-15 * (abs(x) + abs(y)))
    r = c * x + s * y
    assert abs(r) == pytest.approx(np.hypot(x, y), rel=1e-15)
    # sign of r follows the larger-magnitude component (M/GivensRotation.java:43,50-51)
    roe = x if abs(x) > abs(y) else y
    assert np.sign(r) == np.sign(roe)
    assert oracle.givens(0.0, 0.0) == (1.0, 0.0)


def test_triu_solve_matches_scipy_and_detects_singular():
    rng = np.random.default_rng(3)
    m = 31
    H = np.triu(rng.standard_normal((m, m - 1)) + 3 * np.eye(m, m - 1))
    for k in (0, 1, 5, 30):
        s = rng.standard_normal(m)
        code, got = oracle.triu_solve(H, k, s)
        assert code == oracle.OK
        if k:
            np.testing.assert_allclose(got[:k], sla.solve_triangular(H[:k, :k], s[:k]), rtol=1e-12)
        np.testing.assert_array_equal(got[k:], s[k:])
    H2 = H.copy()
    H2[2, 2] = 0.0
    assert oracle.triu_solve(H2, 5, np.ones(m))[0] == oracle.SINGULAR


# ----------------------------------------------------------------- monitor

def test_monitor_predicate_order():
    m = oracle.monitor(10, 1e-3, 1e-50, 10.0)
    lib = oracle.load()
    import ctypes as C
    assert lib.oc_monitor_converged(C.byref(m), 2.0) == 0  # captures initR = 2
    assert m.init_r == 2.0
    m.iter = 1
    assert lib.oc_monitor_converged(C.byref(m), 1e-3) == 1  # 1e-3 < 2e-3
    assert lib.oc_monitor_converged(C.byref(m), 25.0) == -oracle.DIVERGENCE
    m.iter = 10
    assert lib.oc_monitor_converged(C.byref(m), 1.0) == -oracle.ITERATIONS
    # NaN after the iteration limit reports Iterations (checks run in order)
    assert lib.oc_monitor_converged(C.byref(m), float("nan")) == -oracle.ITERATIONS
    m.iter = 3
    assert lib.oc_monitor_converged(C.byref(m), float("nan")) == -oracle.DIVERGENCE


# ------------------------------------------------------------------ solvers

def _spd(n, seed):
    rng = np.random.default_rng(seed)
    d = rng.standard_normal((n, n)) * (rng.random((n, n)) < 0.2)
    d = d + d.T
    d[np.arange(n), np.arange(n)] = np.abs(d).sum(1) + 1.0
    return d


@pytest.mark.parametrize("method,prec", [("cg", 0), ("cg", 1), ("cg", 2), ("bicgstab", 0),
                                         ("bicgstab", 1), ("bicgstab", 2), ("gmres", 0),
                                         ("gmres", 1), ("gmres", 2), ("ir", 1)])
def test_solvers_reach_solution(method, prec):
    d = _spd(40, 11)
    if method in ("bicgstab", "gmres"):
        d = d + np.triu(d, 1) * 0.3  # nonsymmetric, still diagonally dominant
    A = _csr(d)
    x_true = P.uniform(40, 5)
    b = d @ x_true
    M = oracle.Precond(prec, A) if prec else None
    code, x, it, res = oracle.solve(method, A, b, np.zeros(40), M, oracle.monitor(10000, 1e-10))
    assert code == oracle.OK
    np.testing.assert_allclose(x, np.linalg.solve(d, b), rtol=1e-7, atol=1e-8)


def test_cg_finite_termination_and_diagonal_one_step():
    d = _spd(8, 2)
    A = _csr(d)
    b = np.arange(1.0, 9.0)
    code, x, it, res = oracle.solve("cg", A, b, np.zeros(8), None, oracle.monitor(100, 1e-12))
    assert code == oracle.OK and it <= 8 + 2
    D = _csr(np.diag(np.arange(1.0, 6.0)))
    code, x, it, _ = oracle.solve("cg", D, np.ones(5), np.zeros(5),
                                  oracle.Precond(oracle.PREC_DIAGONAL, D), oracle.monitor(100, 1e-12))
    assert code == oracle.OK and it == 1


def test_gmres_restart_at_least_n_is_exact_within_n_steps():
    rng = np.random.default_rng(4)
    n = 12
    d = rng.standard_normal((n, n)) + n * np.eye(n)
    A = _csr(d)
    b = rng.standard_normal(n)
    code, x, it, res = oracle.solve("gmres", A, b, np.zeros(n), None, oracle.monitor(100, 1e-12), restart=n)
    assert code == oracle.OK and it <= n + 1
    np.testing.assert_allclose(x, np.linalg.solve(d, b), rtol=1e-9)


def test_forced_failures():
    d = _spd(30, 1)
    A = _csr(d)
    b = d @ np.ones(30)
    code, *_ = oracle.solve("cg", A, b, np.zeros(30), None, oracle.monitor(1, 1e-12))
    assert code == oracle.ITERATIONS
    code, *_ = oracle.solve("cg", A, b, np.full(30, 1e3), None, oracle.monitor(100, 1e-12, 1e-50, 1e-3))
    assert code in (oracle.DIVERGENCE, oracle.OK)
    bn = b.copy()
    bn[3] = np.nan
    code, *_ = oracle.solve("cg", A, bn, np.zeros(30), None, oracle.monitor(100, 1e-12))
    assert code == oracle.DIVERGENCE
    # BiCGstab breakdown: rtilde.r == 0 cannot happen at r0 != 0, but a zero
    # right-hand side with a zero start converges at once (0 < atol fails, initR = 0)
    code, x, it, _ = oracle.solve("bicgstab", A, np.zeros(30), np.zeros(30), None,
                                  oracle.monitor(100, 1e-12, 0.0))
    assert code == oracle.DIVERGENCE or code == oracle.BREAKDOWN


def test_against_scipy_on_poisson_config():
    n, _, rp, ci, va = P.poisson2d(64)
    A = oracle.Csr(n, n, rp, ci, va)
    m = sp.csr_matrix((va, ci, rp), shape=(n, n))
    x_true = P.uniform(n, 1)
    b = A.mult(x_true)
    for method in ("cg", "bicgstab", "gmres"):
        code, x, it, res = oracle.solve(method, A, b, np.zeros(n),
                                        oracle.Precond(oracle.PREC_DIAGONAL, A),
                                        oracle.monitor(100000, 1e-10))
        assert code == oracle.OK
        assert np.linalg.norm(b - m @ x) / np.linalg.norm(b) < 1e-8
        np.testing.assert_allclose(x, x_true, atol=1e-6)
