"""GPU parity of DenseVector operations (M/DenseVector.java), the
DiagonalPreconditioner and ILU(0) (S/DiagonalPreconditioner.java, S/ILU.java)
against the oracle.  Elementwise operations and the level-scheduled ILU
factor and triangular solves are bitwise; reductions are within 1e-14
relative (summation order only)."""
import numpy as np
import pytest
import scipy.sparse as sp

import mtjgpu as M
import oracle
from mtjgpu import problems as P

pytestmark = pytest.mark.gpu


def test_vector_ops(ctx):
    n = 100003
    a, b = P.uniform(n, 1), P.uniform(n, 2)
    x, y = M.DenseVector(a, ctx), M.DenseVector(b, ctx)
    np.testing.assert_array_equal(x.copy().scale(0.3).getData(), a * 0.3)
    np.testing.assert_array_equal(x.copy().add(-1.7, y).getData(), a + (-1.7) * b)
    np.testing.assert_array_equal(x.copy().add(y).getData(), a + b)
    np.testing.assert_array_equal(x.copy().set(2.0, y).getData(), 2.0 * b)
    # alpha == 0 shortcuts: set zeroes (NaN not copied), add is a no-op
    yn = b.copy()
    yn[7] = np.nan
    w = M.DenseVector(yn, ctx)
    np.testing.assert_array_equal(x.copy().set(0.0, w).getData(), np.zeros(n))
    np.testing.assert_array_equal(x.copy().add(0.0, w).getData(), a)
    assert x.dot(y) == pytest.approx(oracle.dv_dot(a, b), rel=1e-13)
    assert x.norm(M.Norm.Two) == pytest.approx(oracle.dv_norm(a, "two"), rel=1e-14)
    assert x.norm(M.Norm.One) == pytest.approx(oracle.dv_norm(a, "one"), rel=1e-14)
    assert x.norm(M.Norm.Infinity) == oracle.dv_norm(a, "inf")
    assert x.norm(M.Norm.TwoRobust) == pytest.approx(oracle.dv_norm(a, "robust"), rel=1e-13)
    assert np.isnan(w.norm(M.Norm.Infinity))
    with pytest.raises(M.IndexOutOfBoundsException):
        x.add(M.DenseVector(5, ctx))


@pytest.mark.parametrize("n", [1, 255, 64 * 1024 + 3, 2 ** 21 + 7])
def test_vector_reductions_grids(ctx, n):
    """Reductions over one workgroup, one group of workgroups and the full
    2048-workgroup grid (32 groups, both levels), with the NaN and the
    maximum placed in the last workgroup."""
    a = P.uniform(n, 3)
    a[-1] = 7.5
    x = M.DenseVector(a, ctx)
    # summation order only: the oracle's sequential sum of n same-sign terms
    # carries a relative rounding error of order sqrt(n) * 2^-53 (1.6e-13 at 2^21)
    tol = 1e-14 if n < 4096 else 1e-12
    assert x.dot(x) == pytest.approx(oracle.dv_dot(a, a), rel=tol)
    assert x.norm(M.Norm.Two) == pytest.approx(oracle.dv_norm(a, "two"), rel=tol)
    assert x.norm(M.Norm.One) == pytest.approx(oracle.dv_norm(a, "one"), rel=tol)
    assert x.norm(M.Norm.Infinity) == 7.5
    assert x.norm(M.Norm.TwoRobust) == pytest.approx(oracle.dv_norm(a, "robust"), rel=tol)
    a[-1] = np.nan
    assert np.isnan(M.DenseVector(a, ctx).norm(M.Norm.Infinity))


def test_jacobi_bitwise_and_errors(ctx):
    sysm = P.varcoef_poisson(3, 20, seed=1)
    n, _, rp, ci, va = sysm
    A = M.CompRowMatrix(*sysm, ctx=ctx)
    Mg = M.DiagonalPreconditioner(n, ctx)
    Mg.setMatrix(A)
    Mo = oracle.Precond(oracle.PREC_DIAGONAL, oracle.Csr(*sysm))
    bh = P.uniform(n, 3)
    got = Mg.apply(M.DenseVector(bh, ctx), M.DenseVector(n, ctx)).getData()
    np.testing.assert_array_equal(got, Mo.apply(bh))
    bad = sp.csr_matrix(np.array([[1.0, 2.0], [3.0, 0.0]]))
    bad.eliminate_zeros()
    B = M.CompRowMatrix(bad, ctx=ctx)
    Mb = M.DiagonalPreconditioner(2, ctx)
    with pytest.raises(M.ZeroDiagonalException) as e:
        Mb.setMatrix(B)
    assert e.value.row == 1


@pytest.mark.parametrize("name", ["convdiff_16", "p2d_64", "random_spd"])
def test_ilu_apply_bitwise(ctx, name):
    if name == "random_spd":
        rng = np.random.default_rng(5)
        n = 400
        d = sp.random(n, n, density=0.02, random_state=5, format="csr")
        d = d + d.T + sp.diags(np.full(n, 10.0))
        d = d.tocsr()
        d.sort_indices()
        sysm = (n, n, d.indptr, d.indices, d.data)
    else:
        sysm = P.convdiff3d(16, seed=2) if name == "convdiff_16" else P.poisson2d(64)
    n = sysm[0]
    A = M.CompRowMatrix(*sysm, ctx=ctx)
    Mg = M.ILU(A)
    Mg.setMatrix(A)
    Mo = oracle.Precond(oracle.PREC_ILU, oracle.Csr(*sysm))
    assert Mo.code == oracle.OK
    bh = P.uniform(n, 4)
    b = M.DenseVector(bh, ctx)
    np.testing.assert_array_equal(Mg.apply(b, M.DenseVector(n, ctx)).getData(), Mo.apply(bh))
    np.testing.assert_array_equal(Mg.transApply(b, M.DenseVector(n, ctx)).getData(), Mo.trans_apply(bh))
    # setMatrix again with new values on the same pattern refactors
    sysm2 = (sysm[0], sysm[1], sysm[2], sysm[3], np.asarray(sysm[4]) * 2.0)
    A2 = M.CompRowMatrix(*sysm2, ctx=ctx)
    Mg.setMatrix(A2)
    Mo2 = oracle.Precond(oracle.PREC_ILU, oracle.Csr(*sysm2))
    np.testing.assert_array_equal(Mg.apply(b, M.DenseVector(n, ctx)).getData(), Mo2.apply(bh))


@pytest.mark.parametrize("name", ["convdiff_16", "p2d_64"])
def test_ilu_single_launch_solves_bitwise(ctx, monkeypatch, name):
    """The opt-in single-launch triangular solves (dependency values polled
    in place) give the same bits as the per-level launches and the oracle."""
    monkeypatch.setenv("MTJG_ILU_SYNCFREE", "1")
    sysm = P.convdiff3d(16, seed=2) if name == "convdiff_16" else P.poisson2d(64)
    n = sysm[0]
    A = M.CompRowMatrix(*sysm, ctx=ctx)
    Mg = M.ILU(A)
    Mg.setMatrix(A)
    Mo = oracle.Precond(oracle.PREC_ILU, oracle.Csr(*sysm))
    bh = P.uniform(n, 4)
    b = M.DenseVector(bh, ctx)
    for _ in range(3):
        np.testing.assert_array_equal(Mg.apply(b, M.DenseVector(n, ctx)).getData(), Mo.apply(bh))
        np.testing.assert_array_equal(Mg.transApply(b, M.DenseVector(n, ctx)).getData(), Mo.trans_apply(bh))


def test_ilu_errors(ctx):
    m = sp.csr_matrix(np.array([[1.0, 1.0], [1.0, 0.0]]))
    m.eliminate_zeros()
    A = M.CompRowMatrix(m, ctx=ctx)
    with pytest.raises(M.MissingDiagonalException) as e:
        M.ILU(A).setMatrix(A)
    assert e.value.row == 1
    A = M.CompRowMatrix(2, 2, [0, 2, 4], [0, 1, 0, 1], [0.0, 1.0, 1.0, 1.0], ctx=ctx)
    with pytest.raises(M.ZeroPivotException) as e:
        M.ILU(A).setMatrix(A)
    assert e.value.row == 1


@pytest.mark.parametrize("name", ["ilu_tridiagonal", "ilu_pentadiagonal"])
def test_ilu_reference_known_answers(ctx, name):
    import json
    import os
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "mtj_fixtures.json")))[name]
    for n in (1, 2, 3, 6, 9):
        d = np.diag(np.full(n, fx["diagonal"]))
        for k, v in enumerate(fx["offdiagonal"], start=1):
            if k < n:
                d += np.diag(np.full(n - k, v), k) + np.diag(np.full(n - k, v), -k)
        A = M.CompRowMatrix(sp.csr_matrix(d), ctx=ctx)
        ilu = M.ILU(A)
        ilu.setMatrix(A)
        b = d @ np.full(n, fx["x"])
        x = ilu.apply(M.DenseVector(b, ctx), M.DenseVector(n, ctx)).getData()
        assert np.linalg.norm(b - d @ x) <= fx["residual_tol"]


def _wide_level_systems():
    # 40^3 conv-diff: levels up to ~1,200 rows wide (launched on their own),
    # rows of at most 3 strictly-lower / upper entries (slab levels)
    yield "convdiff_40", P.convdiff3d(40, peclet=10.0, seed=5)
    # random SPD, ~20 entries per row: wide levels of long rows (row-major)
    n = 20000
    rng = np.random.default_rng(9)
    d = sp.random(n, n, density=10.0 / n, random_state=9, format="csr")
    d = d + d.T + sp.diags(rng.random(n) + 30.0)
    d = sp.csr_matrix(d)
    d.sort_indices()
    yield "random_spd_20k", (n, n, d.indptr.astype(np.int32), d.indices.astype(np.int32), d.data)
    # 2-D Poisson 200^2: every level at most 200 rows, one fused run of 40,000
    # rows, cut into LDS chunks that import the previous chunk's last levels
    yield "poisson2d_200", P.poisson2d(200)
    # 2-D variable-coefficient Poisson 190 x 170: the same with unequal sides
    yield "varcoef2d_190x170", P.varcoef_poisson(2, 190, 170, seed=6)
    # 27-point stencil on 24^3, SPD with random weights: rows of 13 strictly
    # lower and 13 strictly upper entries, levels of up to 144 rows; in LDS
    # chunks rows spread over 2 (triangles) or 4 lanes (SSOR), SSOR's widest
    # levels over two 512-lane steps
    m = 24
    T = sp.diags([np.ones(m - 1), np.ones(m), np.ones(m - 1)], [-1, 0, 1])
    pat = sp.kron(sp.kron(T, T), T).tocsr()
    rng = np.random.default_rng(13)
    off = sp.triu(pat, 1).tocsr()
    off.data = -rng.random(off.nnz) - 0.1
    off = off + off.T
    d = sp.csr_matrix(off + sp.diags(np.asarray(abs(off).sum(axis=1)).ravel() + 1.0))
    d.sort_indices()
    n = d.shape[0]
    yield "stencil27_24", (n, n, d.indptr.astype(np.int32), d.indices.astype(np.int32), d.data)
    # tridiagonal with a dense last row and column, SPD: a chain of
    # one-row levels in LDS chunks, then a level whose row holds 1,499
    # entries (too long to group), launched on its own after the chunks
    n = 1500
    rng = np.random.default_rng(21)
    d = sp.diags([-np.ones(n - 1), np.full(n, 4.0), -np.ones(n - 1)], [-1, 0, 1]).tolil()
    tail = -rng.random(n - 1) * 0.01
    d[n - 1, : n - 1] = tail
    d[: n - 1, n - 1] = tail.reshape(-1, 1)
    d[n - 1, n - 1] = 4.0 + np.abs(tail).sum()
    d = sp.csr_matrix(d)
    d.sort_indices()
    yield "dense_last_row_1500", (n, n, d.indptr.astype(np.int32), d.indices.astype(np.int32), d.data)


@pytest.mark.parametrize("name,sysm", list(_wide_level_systems()))
def test_wide_levels_bitwise(ctx, name, sysm):
    """Levels wider than one fused workgroup run as launches of their own, in
    both packed layouts, and long fused runs are cut into several LDS chunks;
    ILU apply/transApply, ICC apply/transApply and SSOR stay bitwise."""
    n = sysm[0]
    A = M.CompRowMatrix(*sysm, ctx=ctx)
    O = oracle.Csr(*sysm)
    bh = P.uniform(n, 11)
    x0 = P.uniform(n, 12)
    b = M.DenseVector(bh, ctx)
    ilu = M.ILU(A)
    ilu.setMatrix(A)
    Oi = oracle.Precond(oracle.PREC_ILU, O)
    np.testing.assert_array_equal(ilu.apply(b, M.DenseVector(n, ctx)).getData(), Oi.apply(bh))
    np.testing.assert_array_equal(ilu.transApply(b, M.DenseVector(n, ctx)).getData(), Oi.trans_apply(bh))
    ssor = M.SSOR(A, omegaF=1.2, omegaR=0.8)
    ssor.setMatrix(A)
    Os = oracle.Precond(oracle.PREC_SSOR, O, omega_f=1.2, omega_r=0.8)
    np.testing.assert_array_equal(ssor.apply(b, M.DenseVector(x0, ctx)).getData(), Os.apply(bh, x0))
    ssor1 = M.SSOR(A, reverse=False, omegaF=1.3, omegaR=1.3)
    ssor1.setMatrix(A)
    Os1 = oracle.Precond(oracle.PREC_SSOR, O, reverse=False, omega_f=1.3, omega_r=1.3)
    np.testing.assert_array_equal(ssor1.apply(b, M.DenseVector(x0, ctx)).getData(), Os1.apply(bh, x0))
    if not name.startswith("convdiff"):
        icc = M.ICC(A)
        icc.setMatrix(A)
        Oc = oracle.Precond(oracle.PREC_ICC, O)
        np.testing.assert_array_equal(icc.apply(b, M.DenseVector(n, ctx)).getData(), Oc.apply(bh))
        np.testing.assert_array_equal(icc.transApply(b, M.DenseVector(n, ctx)).getData(), Oc.trans_apply(bh))
