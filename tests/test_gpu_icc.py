"""GPU ICC (mtjg_prec_icc_*, through the C ABI) against the oracle's
restatement of S/ICC.java: the factor is MTJ's bit for bit and both
triangular solves keep MTJ's per-row order, so apply and transApply equal
the oracle's bitwise, on meshes, random patterns, a nonsymmetric matrix and a
matrix whose NaN pivot MTJ spreads into unrelated rows; errors; refactoring on
new values; the single-launch solves; and CG with ICC against the oracle's
iteration counts."""
import numpy as np
import pytest
import scipy.sparse as sp

import mtjgpu as M
import oracle
from mtjgpu import problems as P

pytestmark = pytest.mark.gpu


def _random_spd(n, density, seed):
    rng = np.random.default_rng(seed)
    d = sp.random(n, n, density=density, random_state=seed, format="csr")
    d = d + d.T + sp.diags(rng.random(n) + 8.0)
    d = sp.csr_matrix(d)
    d.sort_indices()
    return n, n, d.indptr.astype(np.int32), d.indices.astype(np.int32), d.data


def _nan_pivot():
    n = 6
    d = np.eye(n) * 4.0
    d[0, 0] = 1.0
    d[0, 1] = d[1, 0] = 3.0
    d[1, 5] = d[5, 1] = 1.0
    d[3, 5] = d[5, 3] = 0.5
    m = sp.csr_matrix(d)
    m.sort_indices()
    return n, n, m.indptr.astype(np.int32), m.indices.astype(np.int32), m.data


SYSTEMS = {
    "p2d_48": lambda: P.poisson2d(48),
    "varcoef3d_12": lambda: P.varcoef_poisson(3, 12, 10, 9, seed=2),
    "convdiff_10": lambda: P.convdiff3d(10, peclet=10.0, seed=3),
    "random_spd_500": lambda: _random_spd(500, 0.01, 5),
    "nan_pivot": _nan_pivot,
}


@pytest.mark.parametrize("name", sorted(SYSTEMS))
def test_apply_bitwise(ctx, name):
    sysm = SYSTEMS[name]()
    n = sysm[0]
    A = M.CompRowMatrix(*sysm, ctx=ctx)
    Mg = M.ICC(A)
    Mg.setMatrix(A)
    Mo = oracle.Precond(oracle.PREC_ICC, oracle.Csr(*sysm))
    assert Mo.code == oracle.OK
    bh = P.uniform(n, 4)
    b = M.DenseVector(bh, ctx)
    ref = Mo.apply(bh)
    np.testing.assert_array_equal(Mg.apply(b, M.DenseVector(n, ctx)).getData(), ref)
    np.testing.assert_array_equal(Mg.transApply(b, M.DenseVector(n, ctx)).getData(), ref)
    # b and x may be one vector (apply reads b into its own temporary first)
    bb = M.DenseVector(bh, ctx)
    np.testing.assert_array_equal(Mg.apply(bb, bb).getData(), ref)
    # setMatrix with new values on the same pattern refactors
    sysm2 = (sysm[0], sysm[1], sysm[2], sysm[3], np.asarray(sysm[4]) * 3.0)
    A2 = M.CompRowMatrix(*sysm2, ctx=ctx)
    Mg.setMatrix(A2)
    Mo2 = oracle.Precond(oracle.PREC_ICC, oracle.Csr(*sysm2))
    np.testing.assert_array_equal(Mg.apply(b, M.DenseVector(n, ctx)).getData(), Mo2.apply(bh))


def test_single_launch_solves_bitwise(ctx, monkeypatch):
    monkeypatch.setenv("MTJG_ILU_SYNCFREE", "1")
    sysm = P.varcoef_poisson(3, 12, 10, 9, seed=2)
    n = sysm[0]
    A = M.CompRowMatrix(*sysm, ctx=ctx)
    Mg = M.ICC(A)
    Mg.setMatrix(A)
    Mo = oracle.Precond(oracle.PREC_ICC, oracle.Csr(*sysm))
    bh = P.uniform(n, 6)
    b = M.DenseVector(bh, ctx)
    for _ in range(3):
        np.testing.assert_array_equal(Mg.apply(b, M.DenseVector(n, ctx)).getData(), Mo.apply(bh))


def test_reference_band_systems(ctx):
    """ICCTest's tri- and penta-diagonal systems: exact factorization, so
    apply(A x) returns x."""
    for diag, offs in ((2.0, [-1.0]), (4.0, [-1.0, -1.0])):
        for n in (1, 2, 3, 7, 10):
            d = np.diag(np.full(n, diag))
            for k, v in enumerate(offs, start=1):
                if k < n:
                    d += np.diag(np.full(n - k, v), k) + np.diag(np.full(n - k, v), -k)
            A = M.CompRowMatrix(sp.csr_matrix(d), ctx=ctx)
            icc = M.ICC(A)
            icc.setMatrix(A)
            x = icc.apply(M.DenseVector(d @ np.ones(n), ctx), M.DenseVector(n, ctx)).getData()
            r = d @ x - d @ np.ones(n)
            assert np.linalg.norm(r) < 1e-12


def test_errors(ctx):
    m = sp.csr_matrix(np.array([[1.0, 1.0], [1.0, 0.0]]))
    m.eliminate_zeros()
    A = M.CompRowMatrix(m, ctx=ctx)
    with pytest.raises(M.MissingDiagonalException) as e:
        M.ICC(A).setMatrix(A)
    assert e.value.row == 1
    A = M.CompRowMatrix(2, 2, [0, 2, 4], [0, 1, 0, 1], [1.0, 1.0, 1.0, 1.0], ctx=ctx)
    with pytest.raises(M.ZeroDiagonalException) as e:
        M.ICC(A).setMatrix(A)
    assert e.value.row == 1
    icc = M.ICC(A)
    with pytest.raises(M.IllegalArgumentException):
        icc.apply(M.DenseVector(2, ctx), M.DenseVector(2, ctx))  # setMatrix must come first
    R = M.CompRowMatrix(sp.random(4, 5, density=0.5, format="csr", random_state=1), ctx=ctx)
    with pytest.raises(M.IllegalArgumentException):
        M.ICC(R)


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
    m = sp.csr_matrix(d)
    m.sort_indices()
    return m


@pytest.mark.parametrize("seed", range(5))
def test_cg_with_icc_reference_style(ctx, seed):
    Am = _mtj_random_spd(seed)
    n = Am.shape[0]
    rng = np.random.default_rng(300 + seed)
    xt = rng.random(n)
    b = Am @ xt
    x0 = rng.random(n)
    O = oracle.Csr(n, n, Am.indptr.astype(np.int32), Am.indices.astype(np.int32), Am.data)
    oc, _, oit, _ = oracle.solve("cg", O, b, x0, oracle.Precond(oracle.PREC_ICC, O),
                                 oracle.monitor(100000, 1e-5))
    A = M.CompRowMatrix(Am, ctx=ctx)
    x = M.DenseVector(x0, ctx)
    S = M.CG(x)
    Pc = M.ICC(A)
    Pc.setMatrix(A)
    S.setPreconditioner(Pc)
    S.setIterationMonitor(M.DefaultIterationMonitor(100000, 1e-5))
    S.solve(A, M.DenseVector(b, ctx), x)
    assert oc == 0
    assert abs(S.getIterationMonitor().iterations() - oit) <= 1
    np.testing.assert_allclose(x.getData(), xt, atol=1e-4, rtol=0)


def test_cg_with_icc_poisson(ctx):
    sysm = P.poisson2d(64)
    n = sysm[0]
    O = oracle.Csr(*sysm)
    bh = O.mult(P.uniform(n, 1))
    oc, xo, oit, _ = oracle.solve("cg", O, bh, np.zeros(n), oracle.Precond(oracle.PREC_ICC, O),
                                  oracle.monitor(10000, 1e-8))
    A = M.CompRowMatrix(*sysm, ctx=ctx)
    x = M.DenseVector(n, ctx)
    S = M.CG(x)
    Pc = M.ICC(A)
    Pc.setMatrix(A)
    S.setPreconditioner(Pc)
    S.setIterationMonitor(M.DefaultIterationMonitor(10000, 1e-8))
    S.solve(A, M.DenseVector(bh, ctx), x)
    assert oc == 0
    it = S.getIterationMonitor().iterations()
    assert abs(it - oit) <= max(1, 0.05 * oit)
    r = bh - O.mult(x.getData())
    assert np.linalg.norm(r) / np.linalg.norm(bh) < 1e-7
