"""Pins the oracle's SSOR (S/SSOR.java:81-199) by the matrix splitting it
implements, A = L + D + U: from the incoming x0, the forward sweep solves
(D + wF L) xx = wF b - (wF U + (wF - 1) D) x0 and the reverse sweep
(D + wR U) x = wR b - (wR L + (wR - 1) D) xx, each a dense triangular solve
here (scipy.linalg.solve_triangular).  With w = 1 and x0 = 0 this is the
symmetric Gauss-Seidel step.  Also: setOmega's range check, the missing
diagonal error, and SSOR-preconditioned CG on the reference's CGSSORTest
systems (random SPD, solution to 1e-4).  The sweeps start from the
preconditioner output's incoming values, so SSOR inside a Krylov solver is
not a fixed operator: on harder systems (e.g. variable-coefficient Poisson)
MTJ's CG or GMRES + SSOR stagnates, and so does this restatement."""
import numpy as np
import pytest
import scipy.linalg as sl
import scipy.sparse as sp

import oracle
from mtjgpu import problems as P


def _split(A):
    A = A.toarray()
    return np.tril(A, -1), np.diag(np.diag(A)), np.triu(A, 1)


def _closed_form(A, b, x0, wf, wr, reverse):
    L, D, U = _split(A)
    xx = sl.solve_triangular(D + wf * L, wf * b - (wf * U + (wf - 1) * D) @ x0, lower=True)
    if not reverse:
        return xx
    return sl.solve_triangular(D + wr * U, wr * b - (wr * L + (wr - 1) * D) @ xx, lower=False)


def _matrices():
    n, _, rp, ci, va = P.varcoef_poisson(2, 9, 7, seed=3)
    yield sp.csr_matrix((va, ci, rp), shape=(n, n))
    n, _, rp, ci, va = P.convdiff3d(5, peclet=10.0, seed=2)
    yield sp.csr_matrix((va, ci, rp), shape=(n, n))


@pytest.mark.parametrize("wf,wr,reverse", [(1.0, 1.0, True), (1.3, 0.7, True), (0.8, 1.9, True), (1.5, 1.0, False)])
def test_sweeps_match_the_splitting(wf, wr, reverse):
    for A in _matrices():
        A.sort_indices()
        n = A.shape[0]
        O = oracle.Csr(n, n, A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data)
        M = oracle.Precond(oracle.PREC_SSOR, O, reverse=reverse, omega_f=wf, omega_r=wr)
        assert M.code == oracle.OK
        b = P.uniform(n, 5)
        for x0 in (np.zeros(n), P.uniform(n, 6)):
            x = M.apply(b, x0)
            want = _closed_form(A, b, x0, wf, wr, reverse)
            assert np.abs(x - want).max() <= 1e-12 * max(1.0, np.abs(want).max())
            np.testing.assert_array_equal(M.trans_apply(b, x0), x)


def test_symmetric_gauss_seidel_from_zero():
    n, _, rp, ci, va = P.poisson2d(6)
    A = sp.csr_matrix((va, ci, rp), shape=(n, n))
    L, D, U = _split(A)
    b = P.uniform(n, 3)
    xf = np.linalg.solve(D + L, b)
    want = np.linalg.solve(D + U, b - L @ xf)
    M = oracle.Precond(oracle.PREC_SSOR, oracle.Csr(n, n, rp, ci, va))
    assert np.abs(M.apply(b) - want).max() <= 1e-14


def test_errors():
    n, _, rp, ci, va = P.poisson2d(4)
    O = oracle.Csr(n, n, rp, ci, va)
    for wf, wr in ((2.5, 1.0), (1.0, -0.1)):
        assert oracle.Precond(oracle.PREC_SSOR, O, omega_f=wf, omega_r=wr).code == oracle.BAD_ARG
    # drop the diagonal of row 5
    A = sp.csr_matrix((va, ci, rp), shape=(n, n)).tolil()
    A[5, 5] = 0
    A = A.tocsr()
    A.eliminate_zeros()
    M = oracle.Precond(oracle.PREC_SSOR, oracle.Csr(n, n, A.indptr.astype(np.int32), A.indices.astype(np.int32),
                                                    A.data))
    assert M.code == oracle.MISSING_DIAGONAL and M.bad.value == 5


def _mtj_random_spd(seed):
    """symmetryPopulate-style SPD system as in the reference's solver tests:
    n in [1, 50), up to 10 entries per row, diagonal shifted by 100 until the
    smallest eigenvalue is positive."""
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
    m = sp.csr_matrix(d)
    m.sort_indices()
    return m


@pytest.mark.parametrize("seed", range(6))
def test_ssor_preconditioned_cg_reference_style(seed):
    # CGSSORTest's systems (random SPD, random x0, the default monitor,
    # solution to 1e-4).  That test never attaches its SSOR to the solver
    # (IterativeSolverTestAbstract.setUp only calls M.setMatrix), so it runs
    # plain CG; here SSOR is attached, with omega = 1 (symmetric Gauss-Seidel):
    # with omega != 1 the stateful sweeps make CG stagnate on some of these
    # systems, in MTJ as in this restatement.
    A = _mtj_random_spd(seed)
    n = A.shape[0]
    rng = np.random.default_rng(50 + seed)
    xt = rng.random(n)
    b = A @ xt
    O = oracle.Csr(n, n, A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data)
    Mp = oracle.Precond(oracle.PREC_SSOR, O, omega_f=1.0, omega_r=1.0)
    code, x, it, _ = oracle.solve("cg", O, b, rng.random(n), Mp, oracle.monitor(100000, 1e-5))
    assert code == oracle.OK
    np.testing.assert_allclose(x, xt, atol=1e-4, rtol=0)
