"""GPU parity of the Krylov solvers (S/CG.java, S/BiCGstab.java,
S/GMRES.java, S/IR.java) with Jacobi and ILU(0), through the C ABI.

Bars: the reference's own solver tests (|x - x_true| <= 1e-4 per entry after
a solve and after repeated solves, TS/IterativeSolverTestAbstract.java:
140-168) on seeded MTJ-style random systems, and on the stencil workloads the
same final monitor verdict as the oracle, iteration counts within 5% of the
oracle's, and a true relative residual at the requested tolerance."""
import numpy as np
import pytest
import scipy.sparse as sp

import mtjgpu as M
import oracle
from mtjgpu import problems as P

pytestmark = pytest.mark.gpu


def _mtj_random(seed, spd):
    """Utilities.rowPopulate / symmetryPopulate-style system: n in [1,50),
    up to 10 entries per row, diagonal shifted by 100 until nonsingular
    (SPD: until the smallest eigenvalue is positive)."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 50))
    b = int(rng.integers(0, min(10, n) + 1))
    d = np.zeros((n, n))
    for i in range(n):
        cols = rng.choice(n, size=b, replace=False) if b else []
        d[i, cols] = rng.random(len(cols))
    if spd:
        d = d + d.T
    while True:
        d[np.arange(n), np.arange(n)] += 100.0
        if spd and np.linalg.eigvalsh(d)[0] > 0:
            break
        if not spd and abs(np.linalg.det(d)) > 1e-300:
            break
    m = sp.csr_matrix(d)
    m.sort_indices()
    return m


def _prec(kind, A, n, ctx):
    if kind == "diag":
        p = M.DiagonalPreconditioner(n, ctx)
    elif kind == "ilu":
        p = M.ILU(A)
    else:
        return None
    p.setMatrix(A)
    return p


@pytest.mark.parametrize("solver", ["CG", "BiCGstab", "GMRES", "IR"])
@pytest.mark.parametrize("prec", ["none", "diag", "ilu"])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_reference_style_random_systems(ctx, solver, prec, seed):
    if solver == "IR" and prec == "none":
        pytest.skip("Richardson without a preconditioner diverges for these shifts")
    m = _mtj_random(seed, spd=(solver == "CG"))
    n = m.shape[0]
    A = M.CompRowMatrix(m, ctx=ctx)
    rng = np.random.default_rng(100 + seed)
    x_true = rng.random(n)
    b = M.DenseVector(m @ x_true, ctx)
    x = M.DenseVector(rng.random(n), ctx)
    S = getattr(M, solver)(x)
    S.setPreconditioner(_prec(prec, A, n, ctx))
    for rep in range(5):  # testRepeatSolve
        S.solve(A, b, x)
        np.testing.assert_allclose(x.getData(), x_true, atol=1e-4, rtol=0)
        x.upload(rng.random(n))


def _compare_to_oracle(ctx, sysm, method, prec, rtol, restart=30, max_iter=100000):
    n = sysm[0]
    O = oracle.Csr(*sysm)
    x_true = P.uniform(n, 1)
    bh = O.mult(x_true)
    A = M.CompRowMatrix(*sysm, ctx=ctx)
    op = None
    if prec == "diag":
        op = oracle.Precond(oracle.PREC_DIAGONAL, O)
    elif prec == "ilu":
        op = oracle.Precond(oracle.PREC_ILU, O)
    code, xo, it_o, res_o = oracle.solve(method, O, bh, np.zeros(n), op,
                                         oracle.monitor(max_iter, rtol), restart=restart)
    b = M.DenseVector(bh, ctx)
    x = M.DenseVector(n, ctx)
    cls = {"cg": M.CG, "bicgstab": M.BiCGstab, "gmres": M.GMRES, "ir": M.IR}[method]
    S = cls(x, restart) if method == "gmres" else cls(x)
    S.setPreconditioner(_prec(prec, A, n, ctx))
    S.setIterationMonitor(M.DefaultIterationMonitor(max_iter, rtol))
    S.solve(A, b, x)
    it_g = S.getIterationMonitor().iterations()
    xg = x.getData()
    assert code == oracle.OK
    assert abs(it_g - it_o) <= max(1, 0.05 * it_o), (it_g, it_o)
    m = sp.csr_matrix((sysm[4], sysm[3], sysm[2]), shape=(n, n))
    true_rel = np.linalg.norm(bh - m @ xg) / np.linalg.norm(bh)
    ref_rel = np.linalg.norm(bh - m @ xo) / np.linalg.norm(bh)
    assert true_rel <= 10 * max(ref_rel, rtol * 1e-3)
    return it_g, it_o, xg, xo


def test_cg_jacobi_poisson_256(ctx):
    """2-D 5-point Poisson 256^2, Jacobi-CG to 1e-8."""
    it_g, it_o, xg, xo = _compare_to_oracle(ctx, P.poisson2d(256), "cg", "diag", 1e-8)
    assert np.linalg.norm(xg - xo) / np.linalg.norm(xo) < 1e-6


def test_cg_plain_and_ilu(ctx):
    _compare_to_oracle(ctx, P.poisson3d(24), "cg", "none", 1e-8)
    _compare_to_oracle(ctx, P.varcoef_poisson(3, 16, seed=3), "cg", "ilu", 1e-8)


@pytest.mark.parametrize("method,gen", [
    ("cg", lambda: P.poisson3d(82)),
    ("bicgstab", lambda: P.convdiff3d(82, peclet=10.0, seed=4)),
])
def test_split_fused_product(ctx, method, gen):
    """551,368 rows: the fused product's grid (2,154 workgroups) exceeds one
    persistent grid, so it runs as the full-grid product with unticketed
    partial sums plus the reduction kernel, and CG keeps its separate
    direction update; iterations and solution as the oracle's."""
    it_g, it_o, xg, xo = _compare_to_oracle(ctx, gen(), method, "diag", 1e-8)
    assert np.linalg.norm(xg - xo) / np.linalg.norm(xo) < 1e-6


@pytest.mark.parametrize("prec", ["none", "diag", "ilu"])
def test_bicgstab_convdiff(ctx, prec):
    _compare_to_oracle(ctx, P.convdiff3d(24, seed=7), "bicgstab", prec, 1e-8)


@pytest.mark.parametrize("prec", ["none", "diag", "ilu"])
def test_gmres_convdiff(ctx, prec):
    _compare_to_oracle(ctx, P.convdiff3d(20, seed=7), "gmres", prec, 1e-8)


def test_gmres_small_restart(ctx):
    _compare_to_oracle(ctx, P.convdiff3d(12, seed=2), "gmres", "diag", 1e-8, restart=5)


def test_ir_jacobi(ctx):
    _compare_to_oracle(ctx, P.powerlaw(3000, seed=1), "ir", "diag", 1e-10)


def test_mixed_precision_cg(ctx):
    sysm = P.varcoef_poisson(3, 32, seed=5)
    n = sysm[0]
    O = oracle.Csr(*sysm)
    x_true = P.uniform(n, 1)
    bh = O.mult(x_true)
    code, xo, it_o, _ = oracle.solve("cg", O, bh, np.zeros(n), oracle.Precond(oracle.PREC_DIAGONAL, O),
                                     oracle.monitor(100000, 1e-10))
    A = M.CompRowMatrix(*sysm, ctx=ctx)
    b = M.DenseVector(bh, ctx)
    x = M.DenseVector(n, ctx)
    S = M.MixedCG(x, inner_rtol=1e-4, inner_max_iter=2000)
    S.setIterationMonitor(M.DefaultIterationMonitor(100, 1e-10))
    S.solve(A, b, x)
    m = sp.csr_matrix((sysm[4], sysm[3], sysm[2]), shape=(n, n))
    xg = x.getData()
    assert np.linalg.norm(bh - m @ xg) / np.linalg.norm(bh) <= 1e-10
    assert np.linalg.norm(xg - xo) / np.linalg.norm(xo) < 1e-7
    assert S.inner_iterations() > 0


def test_forced_failures(ctx):
    sysm = P.poisson2d(32)
    n = sysm[0]
    A = M.CompRowMatrix(*sysm, ctx=ctx)
    b = M.DenseVector(P.uniform(n, 1), ctx)
    x = M.DenseVector(n, ctx)
    S = M.CG(x)
    S.setIterationMonitor(M.DefaultIterationMonitor(1, 1e-12))
    with pytest.raises(M.IterativeSolverNotConvergedException) as e:
        S.solve(A, b, x)
    assert e.value.getReason() == M.Reason.Iterations and e.value.getIterations() == 1
    bn = P.uniform(n, 1)
    bn[5] = np.nan
    x.zero()
    S.setIterationMonitor(M.DefaultIterationMonitor(100, 1e-12))
    with pytest.raises(M.IterativeSolverNotConvergedException) as e:
        S.solve(A, M.DenseVector(bn, ctx), x)
    assert e.value.getReason() == M.Reason.Divergence
    with pytest.raises(M.IllegalArgumentException):
        S.solve(A, M.DenseVector(n + 1, ctx), x)


def test_reporter_and_determinism(ctx):
    sysm = P.poisson2d(48)
    n = sysm[0]
    A = M.CompRowMatrix(*sysm, ctx=ctx)
    b = M.DenseVector(P.uniform(n, 2), ctx)
    seen = []
    x1 = M.DenseVector(n, ctx)
    S = M.CG(x1)
    S.setIterationMonitor(M.DefaultIterationMonitor(10000, 1e-8))
    S.setIterationReporter(lambda r, i: seen.append((i, r)))
    S.solve(A, b, x1)
    assert len(seen) >= S.getIterationMonitor().iterations()
    S.setIterationReporter(None)
    x2 = M.DenseVector(n, ctx)
    S.solve(A, b, x2)
    np.testing.assert_array_equal(x1.getData(), x2.getData())


def test_ir_with_fp32_cg_preconditioner(ctx):
    """C4's scheme through MTJ's plug-in points: IR (S/IR.java:57-70) with a
    Preconditioner whose apply is the FP32 inner Jacobi-CG, against the
    oracle's IR with the same preconditioner restated: the same refinement
    count, inner iterations within 5%, the FP64 residual at 1e-10, and the
    oracle's solution to FP64 accuracy."""
    sysm = P.varcoef_poisson(3, 32, seed=5)
    n = sysm[0]
    O = oracle.Csr(*sysm)
    x_true = P.uniform(n, 1)
    bh = O.mult(x_true)
    op = oracle.Precond(oracle.PREC_CG32, O, inner_rtol=1e-4)
    code, xo, it_o, _ = oracle.solve("ir", O, bh, np.zeros(n), op, oracle.monitor(100, 1e-10))
    assert code == oracle.OK
    A = M.CompRowMatrix(*sysm, ctx=ctx)
    b = M.DenseVector(bh, ctx)
    x = M.DenseVector(n, ctx)
    Mp = M.FP32CGPreconditioner(n, inner_rtol=1e-4, ctx=ctx)
    Mp.setMatrix(A)
    # one apply on its own: the FP32 solve of the device against the oracle's
    z = M.DenseVector(n, ctx)
    zg = Mp.apply(b, z).getData()
    zo = oracle.Precond(oracle.PREC_CG32, O, inner_rtol=1e-4).apply(bh)
    assert np.linalg.norm(zg - zo) <= 1e-6 * np.linalg.norm(zo)
    S = M.IR(x)
    S.setPreconditioner(Mp)
    S.setIterationMonitor(M.DefaultIterationMonitor(100, 1e-10))
    before = Mp.inner_iterations()
    S.solve(A, b, x)
    inner = Mp.inner_iterations() - before
    xg = x.getData()
    assert S.getIterationMonitor().iterations() == it_o
    assert abs(inner - op.s.inner_total) <= max(2, 0.05 * op.s.inner_total), (inner, op.s.inner_total)
    m = sp.csr_matrix((sysm[4], sysm[3], sysm[2]), shape=(n, n))
    assert np.linalg.norm(bh - m @ xg) / np.linalg.norm(bh) <= 1e-10
    assert np.linalg.norm(xg - xo) / np.linalg.norm(xo) < 1e-9


def test_fp32_cg_preconditioner_errors(ctx):
    A = M.CompRowMatrix(3, 3, [0, 1, 2, 3], [0, 1, 2], [1.0, 0.0, 2.0], ctx=ctx)
    Mp = M.FP32CGPreconditioner(3, ctx=ctx)
    x = M.DenseVector(3, ctx)
    with pytest.raises(M.IllegalArgumentException):
        Mp.apply(M.DenseVector(np.ones(3), ctx), x)
    with pytest.raises(M.ZeroDiagonalException) as e:
        Mp.setMatrix(A)
    assert e.value.row == 1
    A = M.CompRowMatrix(3, 3, [0, 1, 2, 3], [0, 1, 2], [1.0, 4.0, 2.0], ctx=ctx)
    Mp.setMatrix(A)
    x.upload(np.ones(3))
    np.testing.assert_array_equal(Mp.apply(M.DenseVector(np.zeros(3), ctx), x).getData(), np.zeros(3))
    np.testing.assert_allclose(Mp.apply(M.DenseVector(np.array([1.0, 2.0, 4.0]), ctx), x).getData(),
                               [1.0, 0.5, 2.0], rtol=1e-7)
    with pytest.raises(M.IllegalArgumentException):
        M.FP32CGPreconditioner(3, inner_rtol=0.0, ctx=ctx)


@pytest.mark.parametrize("gen,diags", [("varcoef3d_24", 7), ("poisson2d_96", 5), ("poisson3d_20x17x9", 7),
                                       ("varcoef3d_84", 7)])
def test_fp32_stencil_copy_bitwise_sell(ctx, monkeypatch, gen, diags):
    """The FP32 stencil copy (values only, diagonal by diagonal) sums each row
    in the SELL-64 FP32 product's order: the mixed-precision solve and the
    FP32-CG preconditioner give bitwise the same results with the copy as
    without it (MTJG_SPMV_STENCIL=0).  84^3 (592,704 rows) runs the split
    product (more than 2048 workgroups: partials, then k_sell_fin)."""
    sysm = {"varcoef3d_24": lambda: P.varcoef_poisson(3, 24, seed=3),
            "poisson2d_96": lambda: P.poisson2d(96),
            "poisson3d_20x17x9": lambda: P.varcoef_poisson(3, 20, 17, 9, seed=4),
            "varcoef3d_84": lambda: P.varcoef_poisson(3, 84, seed=6)}[gen]()
    n = sysm[0]
    bh = oracle.Csr(*sysm).mult(P.uniform(n, 2))
    out = {}
    for flag in ("1", "0"):
        monkeypatch.setenv("MTJG_SPMV_STENCIL", flag)
        A = M.CompRowMatrix(*sysm, ctx=ctx)
        assert A.fp32_diagonals() == (diags if flag == "1" else 0)
        b = M.DenseVector(bh, ctx)
        x = M.DenseVector(n, ctx)
        S = M.MixedCG(x, inner_rtol=1e-4)
        S.setIterationMonitor(M.DefaultIterationMonitor(100, 1e-10))
        S.solve(A, b, x)
        Mp = M.FP32CGPreconditioner(n, inner_rtol=1e-4, ctx=ctx)
        Mp.setMatrix(A)
        z = M.DenseVector(n, ctx)
        out[flag] = (x.getData(), S.inner_iterations(), Mp.apply(b, z).getData(), Mp.inner_iterations())
        # values changed after the copy was filled reach it before the next product
        A.setValues(sysm[4] * 2.0)
        Mp.setMatrix(A)
        out[flag] += (Mp.apply(b, z).getData(),)
        for o in (Mp, S, z, x, b, A):
            o.close()
    a, c = out["1"], out["0"]
    np.testing.assert_array_equal(a[0], c[0])
    assert a[1] == c[1] and a[3] == c[3]
    np.testing.assert_array_equal(a[2], c[2])
    np.testing.assert_array_equal(a[4], c[4])


def test_fp32_stencil_copy_not_for_wide_patterns(ctx):
    # nine distinct diagonals: the FP32 products keep SELL-64
    n = 200
    rows = [sorted({i, (i + 1) % n, (i + 2) % n, (i + 3) % n, (i + 5) % n, (i + 7) % n, (i + 11) % n,
                    (i + 13) % n, (i + 17) % n}) for i in range(n)]
    rp = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    ci = np.concatenate(rows).astype(np.int32)
    A = M.CompRowMatrix(n, n, rp, ci, np.ones(len(ci)), ctx=ctx)
    assert A.format()[0] == "sell" and A.fp32_diagonals() == 0
    A.close()
