"""Pins the oracle's BiCG, CGS and QMR (S/BiCG.java:63-102, S/CGS.java:65-105,
S/QMR.java:117-213) without retyping their loops:

* SciPy's bicg / cgs / qmr implement the same Templates algorithms with the
  same stopping rule for x0 = 0 (||r_k|| < rtol ||b||): on seeded
  nonsymmetric systems, unpreconditioned and with a Jacobi preconditioner
  given as a LinearOperator (matvec and rmatvec both scale by 1/diag, which
  is what apply and transApply do), the iteration counts must be equal and
  the solutions agree to rounding;
* BiCG with the identity preconditioner on an SPD matrix produces CG's
  iterates (rtilde stays equal to r), so its iteration count equals the
  oracle CG's;
* exact-arithmetic breakdowns on 2x2 integer systems (worked by hand in the
  comments) and the monitor's iteration limit and NaN verdicts."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

import oracle
from mtjgpu import problems as P


def _nonsym(grid, seed, vary_diag=False):
    n, _, rp, ci, va = P.convdiff3d(grid, peclet=10.0, seed=seed)
    A = sp.csr_matrix((va, ci, rp), shape=(n, n))
    if vary_diag:
        rng = np.random.default_rng(seed)
        A = (A + sp.diags(rng.uniform(0.5, 3.0, size=n))).tocsr()
        A.sort_indices()
    return A


def _ocsr(A):
    A = A.tocsr()
    A.sort_indices()
    return oracle.Csr(A.shape[0], A.shape[1], A.indptr.astype(np.int32), A.indices.astype(np.int32),
                      A.data.astype(np.float64))


SCIPY = {"bicg": sla.bicg, "cgs": sla.cgs, "qmr": sla.qmr}


@pytest.mark.parametrize("method", ["bicg", "cgs", "qmr"])
@pytest.mark.parametrize("grid", [8, 12])
def test_unpreconditioned_matches_scipy(method, grid):
    A = _nonsym(grid, 1)
    n = A.shape[0]
    b = A @ P.uniform(n, 1)
    code, x, it, _ = oracle.solve(method, _ocsr(A), b, np.zeros(n), None, oracle.monitor(10000, 1e-10))
    assert code == oracle.OK
    cnt = [0]
    xs, info = SCIPY[method](A, b, rtol=1e-10, atol=0.0, maxiter=10000,
                             callback=lambda xk: cnt.__setitem__(0, cnt[0] + 1))
    assert info == 0
    assert it == cnt[0]
    assert np.linalg.norm(x - xs) <= 1e-9 * np.linalg.norm(xs)


@pytest.mark.parametrize("method", ["bicg", "cgs", "qmr"])
def test_jacobi_matches_scipy(method):
    A = _nonsym(10, 3, vary_diag=True)
    n = A.shape[0]
    b = A @ P.uniform(n, 2)
    O = _ocsr(A)
    code, x, it, _ = oracle.solve(method, O, b, np.zeros(n), oracle.Precond(oracle.PREC_DIAGONAL, O),
                                  oracle.monitor(10000, 1e-10))
    assert code == oracle.OK
    inv = 1.0 / A.diagonal()
    J = sla.LinearOperator((n, n), matvec=lambda v: inv * v, rmatvec=lambda v: inv * v, dtype=float)
    cnt = [0]
    # MTJ's QMR.setPreconditioner uses M on both sides (M1 = M2 = M)
    kw = {"M1": J, "M2": J} if method == "qmr" else {"M": J}
    xs, info = SCIPY[method](A, b, rtol=1e-10, atol=0.0, maxiter=10000,
                             callback=lambda xk: cnt.__setitem__(0, cnt[0] + 1), **kw)
    assert info == 0
    assert it == cnt[0]
    assert np.linalg.norm(x - xs) <= 1e-8 * np.linalg.norm(xs)


def test_bicg_on_spd_is_cg():
    n, _, rp, ci, va = P.poisson2d(24)
    O = oracle.Csr(n, n, rp, ci, va)
    b = sp.csr_matrix((va, ci, rp), shape=(n, n)) @ P.uniform(n, 4)
    c1, x1, it1, _ = oracle.solve("bicg", O, b, np.zeros(n), None, oracle.monitor(10000, 1e-10))
    c2, x2, it2, _ = oracle.solve("cg", O, b, np.zeros(n), None, oracle.monitor(10000, 1e-10))
    assert c1 == c2 == oracle.OK
    assert abs(it1 - it2) <= 1
    assert np.linalg.norm(x1 - x2) <= 1e-8 * np.linalg.norm(x2)


def _tiny(M, b):
    A = sp.csr_matrix(np.array(M, float))
    A.eliminate_zeros()
    return _ocsr(A), np.array(b, float)


def test_exact_breakdowns():
    # BiCG, A = [[-1,-1],[0,2]], b = [0,2]: r = rt = [0,2], rho = 4, q = [-2,4],
    # qt = A^T rt = [0,4], alpha = 4/8; then r = [1,0] and rt = [0,0], so the
    # next rho = z.rt is exactly 0 -> Breakdown at iteration 1.
    O, b = _tiny([[-1, -1], [0, 2]], [0, 2])
    code, x, it, _ = oracle.solve("bicg", O, b, np.zeros(2), None, oracle.monitor(50, 1e-10))
    assert (code, it) == (oracle.BREAKDOWN, 1)
    np.testing.assert_array_equal(x, [0.0, 1.0])
    # CGS on the same system: rho = rt.r = 4, vhat = A r = [-2,4], alpha = 1/2,
    # q = -alpha vhat + u = [1,0], uhat = u + q = [1,2], qhat = A uhat = [-3,4],
    # r = [0,2] - [-1.5,2] = [1.5,0]; rt.r = 0 at iteration 1.
    code, x, it, _ = oracle.solve("cgs", O, b, np.zeros(2), None, oracle.monitor(50, 1e-10))
    assert (code, it) == (oracle.BREAKDOWN, 1)
    np.testing.assert_array_equal(x, [0.5, 1.0])
    # QMR on the singular A = [[0,0],[1,2]], b = [-1,2]: a scalar of the
    # second step is exactly 0 (breakdown at iteration 1)
    O, b = _tiny([[0, 0], [1, 2]], [-1, 2])
    code, x, it, _ = oracle.solve("qmr", O, b, np.zeros(2), None, oracle.monitor(50, 1e-10))
    assert (code, it) == (oracle.BREAKDOWN, 1)


@pytest.mark.parametrize("method", ["bicg", "cgs", "qmr"])
def test_monitor_verdicts(method):
    A = _nonsym(8, 5)
    n = A.shape[0]
    b = A @ P.uniform(n, 3)
    O = _ocsr(A)
    code, _, it, _ = oracle.solve(method, O, b, np.zeros(n), None, oracle.monitor(5, 1e-12))
    assert (code, it) == (oracle.ITERATIONS, 5)
    bad = b.copy()
    bad[3] = np.nan
    code, _, it, _ = oracle.solve(method, O, bad, np.zeros(n), None, oracle.monitor(50, 1e-8))
    assert (code, it) == (oracle.DIVERGENCE, 0)
    code, x, it, _ = oracle.solve(method, O, np.zeros(n), np.zeros(n), None, oracle.monitor(50, 1e-8))
    assert (code, it) == (oracle.OK, 0) and not x.any()
