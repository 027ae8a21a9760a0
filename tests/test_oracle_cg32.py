"""Pins the oracle's FP32 inner-CG preconditioner (the restatement of the
library's MTJG_PREC_CG32) and IR with it (S/IR.java:57-70), without retyping
its loop:

* on a diagonal matrix Jacobi-CG converges in one iteration and apply(b) is
  b / diag to binary32 rounding;
* one apply tracks the FP64 oracle CG (S/CG.java:58-87, the same loop in
  binary64) on b / ||b|| to the same relative tolerance: the iteration
  counts agree within two and the solutions within binary32 accuracy;
* with a tight inner tolerance apply(b) is A^-1 b to binary32 accuracy
  (numpy's dense solve);
* IR with it reduces the FP64 residual by about the inner tolerance per
  refinement: to 1e-10 with inner tolerance 1e-4 takes three refinements,
  and the result is the FP64 solution (numpy) to FP64 accuracy."""
import numpy as np
import pytest

import oracle
from mtjgpu import problems as P


def _sys(gen):
    s = gen()
    O = oracle.Csr(*s)
    return s, O


def test_diagonal_matrix_one_iteration():
    n = 50
    d = 1.0 + np.arange(n) * 0.37
    rp = np.arange(n + 1, dtype=np.int32)
    O = oracle.Csr(n, n, rp, np.arange(n, dtype=np.int32), d)
    b = P.uniform(n, 3)
    M = oracle.Precond(oracle.PREC_CG32, O, inner_rtol=1e-6)
    assert M.code == oracle.OK
    x = M.apply(b)
    assert M.s.inner_total == 1
    np.testing.assert_allclose(x, b / d, rtol=2e-7)


@pytest.mark.parametrize("gen", [lambda: P.poisson3d(10), lambda: P.varcoef_poisson(3, 10, seed=2),
                                 lambda: P.powerlaw_spd(3000, seed=4)])
def test_apply_tracks_fp64_cg(gen):
    s, O = _sys(gen)
    n = s[0]
    b = O.mult(P.uniform(n, 1))
    M = oracle.Precond(oracle.PREC_CG32, O, inner_rtol=1e-4)
    x32 = M.apply(b)
    it32 = M.s.inner_total
    bn = b / np.linalg.norm(b)
    code, x64, it64, _ = oracle.solve("cg", O, bn, np.zeros(n), oracle.Precond(oracle.PREC_DIAGONAL, O),
                                      oracle.monitor(100000, 1e-4))
    assert code == oracle.OK
    assert abs(it32 - it64) <= 2, (it32, it64)
    x64 = x64 * np.linalg.norm(b)
    assert np.linalg.norm(x32 - x64) <= 1e-4 * np.linalg.norm(x64)


def test_tight_apply_is_the_inverse_to_binary32_accuracy():
    s, O = _sys(lambda: P.varcoef_poisson(2, 12, seed=7))
    n, _, rp, ci, va = s
    A = np.zeros((n, n))
    for i in range(n):
        A[i, ci[rp[i]:rp[i + 1]]] = va[rp[i]:rp[i + 1]]
    b = P.uniform(n, 5)
    x = oracle.Precond(oracle.PREC_CG32, O, inner_rtol=1e-6).apply(b)
    xe = np.linalg.solve(A, b)
    assert np.linalg.norm(x - xe) <= 1e-5 * np.linalg.norm(xe)


def test_zero_diagonal_and_zero_rhs():
    n = 3
    rp = np.array([0, 1, 2, 3], np.int32)
    O = oracle.Csr(n, n, rp, np.array([0, 1, 2], np.int32), np.array([1.0, 0.0, 2.0]))
    M = oracle.Precond(oracle.PREC_CG32, O)
    assert M.code == oracle.ZERO_DIAGONAL and M.bad.value == 1
    O = oracle.Csr(n, n, rp, np.array([0, 1, 2], np.int32), np.array([1.0, 3.0, 2.0]))
    M = oracle.Precond(oracle.PREC_CG32, O)
    np.testing.assert_array_equal(M.apply(np.zeros(n), np.ones(n)), np.zeros(n))
    assert np.all(np.isnan(M.apply(np.array([1.0, np.inf, 0.0]))))


def test_ir_refinement_count_and_fp64_solution():
    s, O = _sys(lambda: P.varcoef_poisson(3, 12, seed=9))
    n, _, rp, ci, va = s
    x_true = P.uniform(n, 1)
    b = O.mult(x_true)
    M = oracle.Precond(oracle.PREC_CG32, O, inner_rtol=1e-4)
    mon = oracle.monitor(100, 1e-10)
    code, x, it, res = oracle.solve("ir", O, b, np.zeros(n), M, mon)
    assert code == oracle.OK and it == 3
    assert res < 1e-10 * np.linalg.norm(b)
    A = np.zeros((n, n))
    for i in range(n):
        A[i, ci[rp[i]:rp[i + 1]]] = va[rp[i]:rp[i + 1]]
    xe = np.linalg.solve(A, b)
    assert np.linalg.norm(x - xe) <= 1e-8 * np.linalg.norm(xe)
