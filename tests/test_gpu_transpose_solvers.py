"""GPU BiCG, CGS and QMR (MTJG_BICG / MTJG_CGS / MTJG_QMR, through the C ABI)
against the oracle's restatement of S/BiCG.java, S/CGS.java and S/QMR.java.

Bars, as for the other solvers: same verdict as the oracle, iteration count
within 5% (at least +-1; dots and norms are summed in a different order),
solution agreeing with the oracle's to the solve tolerance, true relative
residual at the requested tolerance; exact breakdowns on the hand-worked 2x2
systems of the oracle tests give the same breakdown at the same iteration."""
import numpy as np
import pytest
import scipy.sparse as sp

import mtjgpu as M
import oracle
from mtjgpu import problems as P

pytestmark = pytest.mark.gpu

CLS = {"bicg": M.BiCG, "cgs": M.CGS, "qmr": M.QMR}


def _system(grid, seed, vary_diag):
    n, _, rp, ci, va = P.convdiff3d(grid, peclet=10.0, seed=seed)
    A = sp.csr_matrix((va, ci, rp), shape=(n, n))
    if vary_diag:
        rng = np.random.default_rng(seed)
        A = (A + sp.diags(rng.uniform(0.5, 3.0, size=n))).tocsr()
    A.sort_indices()
    return A


def _run(ctx, method, A, b, prec, max_iter=10000, rtol=1e-10):
    n = A.shape[0]
    Ad = M.CompRowMatrix(A, ctx=ctx)
    x = M.DenseVector(n, ctx)
    S = CLS[method](x)
    if prec == "diag":
        Mp = M.DiagonalPreconditioner(n, ctx)
        Mp.setMatrix(Ad)
        S.setPreconditioner(Mp)
    elif prec == "ilu":
        Mp = M.ILU(Ad)
        Mp.setMatrix(Ad)
        S.setPreconditioner(Mp)
    S.setIterationMonitor(M.DefaultIterationMonitor(max_iter, rtol))
    code = 0
    try:
        S.solve(Ad, M.DenseVector(b, ctx), x)
    except M.IterativeSolverNotConvergedException as e:
        code = e.reason.value
    return code, x.getData(), S.getIterationMonitor().iterations()


def _oracle(method, A, b, prec, max_iter=10000, rtol=1e-10):
    n = A.shape[0]
    O = oracle.Csr(n, n, A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data)
    pr = None
    if prec == "diag":
        pr = oracle.Precond(oracle.PREC_DIAGONAL, O)
    elif prec == "ilu":
        pr = oracle.Precond(oracle.PREC_ILU, O)
    code, x, it, _ = oracle.solve(method, O, b, np.zeros(n), pr, oracle.monitor(max_iter, rtol))
    return code, x, it


@pytest.mark.parametrize("method", ["bicg", "cgs", "qmr"])
@pytest.mark.parametrize("prec", ["none", "diag", "ilu"])
def test_matches_oracle(ctx, method, prec):
    A = _system(14, 7, vary_diag=True)
    n = A.shape[0]
    b = A @ P.uniform(n, 5)
    oc, ox, oit = _oracle(method, A, b, prec)
    gc, gx, git = _run(ctx, method, A, b, prec)
    assert gc == oc == 0
    assert abs(git - oit) <= max(1, int(0.05 * oit)), (git, oit)
    assert np.linalg.norm(gx - ox) <= 1e-7 * np.linalg.norm(ox)
    assert np.linalg.norm(b - A @ gx) <= 1e-9 * np.linalg.norm(b)


@pytest.mark.parametrize("method", ["bicg", "qmr"])
def test_transpose_path_on_larger_system(ctx, method):
    # 32^3 conv-diff: the transposed product runs on the CSC built at first use
    A = _system(32, 2, vary_diag=False)
    n = A.shape[0]
    b = A @ P.uniform(n, 6)
    oc, ox, oit = _oracle(method, A, b, "none", rtol=1e-8)
    gc, gx, git = _run(ctx, method, A, b, "none", rtol=1e-8)
    assert gc == oc == 0
    assert abs(git - oit) <= max(1, int(0.05 * oit)), (git, oit)
    assert np.linalg.norm(b - A @ gx) <= 1e-8 * np.linalg.norm(b) * 1.01


def test_exact_breakdowns(ctx):
    for method, Mx, b, x_expect in [("bicg", [[-1, -1], [0, 2]], [0, 2], [0.0, 1.0]),
                                    ("cgs", [[-1, -1], [0, 2]], [0, 2], [0.5, 1.0]),
                                    ("qmr", [[0, 0], [1, 2]], [-1, 2], None)]:
        A = sp.csr_matrix(np.array(Mx, float))
        A.eliminate_zeros()
        A.sort_indices()
        code, x, it = _run(ctx, method, A, np.array(b, float), "none", max_iter=50, rtol=1e-10)
        assert (code, it) == (M.Reason.Breakdown.value, 1), (method, code, it)
        if x_expect is not None:
            np.testing.assert_array_equal(x, x_expect)


@pytest.mark.parametrize("method", ["bicg", "cgs", "qmr"])
def test_monitor_verdicts(ctx, method):
    A = _system(8, 5, vary_diag=False)
    n = A.shape[0]
    b = A @ P.uniform(n, 3)
    code, _, it = _run(ctx, method, A, b, "none", max_iter=5, rtol=1e-12)
    assert (code, it) == (M.Reason.Iterations.value, 5)
    bad = b.copy()
    bad[3] = np.nan
    code, _, it = _run(ctx, method, A, bad, "none", max_iter=50, rtol=1e-8)
    assert (code, it) == (M.Reason.Divergence.value, 0)
