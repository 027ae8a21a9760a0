"""GPU AMG (mtjg_prec_amg_*, through the C ABI) against the oracle's
restatement of S/AMG.java: the same hierarchy, and apply / transApply
bitwise equal to the oracle's (SSOR sweeps, residual, restriction,
prolongation and the coarse LU solve all keep MTJ's operation order) for
the default V-cycle, a W-cycle with two pre-smoothings, the unsmoothed
(omega = 0) variant and the forward-sweep constructor; repeated cycles as a
multigrid iteration; errors; and CG with AMG on the reference's small
systems."""
import numpy as np
import pytest
import scipy.sparse as sp

import mtjgpu as M
import oracle
from mtjgpu import problems as P

pytestmark = pytest.mark.gpu

DEFAULT = ()
SYSTEMS = {
    "p2d_64_default": (lambda: P.poisson2d(64), DEFAULT),
    "p2d_48_w_cycle": (lambda: P.poisson2d(48), (1.2, 1.5, 1.5, 1.2, 2, 1, 2, 30, 0.5)),
    "varcoef3d_12_default": (lambda: P.varcoef_poisson(3, 12, 10, 9, seed=2), DEFAULT),
    "p2d_40_unsmoothed": (lambda: P.poisson2d(40), (1.0, 1.85, 1.85, 1.0, 1, 1, 1, 40, 0.0)),
    "convdiff_10_forward": (lambda: P.convdiff3d(10, peclet=10.0, seed=3), (1.2, 1.2, 1, 2, 1, 20, 2.0 / 3)),
    "powerlaw_3000_default": (lambda: P.powerlaw(3000, seed=4), DEFAULT),
}


def _pair(ctx, name):
    gen, params = SYSTEMS[name]
    sysm = gen()
    n = sysm[0]
    A = M.CompRowMatrix(*sysm, ctx=ctx)
    G = M.AMG(*params, ctx=ctx)
    G.setMatrix(A)
    O = oracle.Amg(oracle.Csr(*sysm), *params)
    assert O.code == oracle.OK
    return n, A, G, O


@pytest.mark.parametrize("name", sorted(SYSTEMS))
def test_apply_bitwise(ctx, name):
    n, A, G, O = _pair(ctx, name)
    assert G.levels() == O.levels()
    bh = P.uniform(n, 4)
    for x0 in (np.zeros(n), P.uniform(n, 5)):
        x = M.DenseVector(x0, ctx)
        G.apply(M.DenseVector(bh, ctx), x)
        rc, ref = O.apply(bh, x0)
        assert rc == oracle.OK
        assert x.getData().tobytes() == ref.tobytes()
        x = M.DenseVector(x0, ctx)
        G.transApply(M.DenseVector(bh, ctx), x)
        rc, ref = O.apply(bh, x0, transpose=True)
        assert x.getData().tobytes() == ref.tobytes()


def test_multigrid_iteration(ctx):
    """Repeated cycles from the previous iterate: a stationary multigrid
    solve, every iterate bitwise the oracle's; 1e-8 within 16 cycles."""
    n, A, G, O = _pair(ctx, "p2d_64_default")
    Oc = oracle.Csr(*P.poisson2d(64))
    bh = Oc.mult(P.uniform(n, 1))
    b = M.DenseVector(bh, ctx)
    x = M.DenseVector(n, ctx)
    xo = np.zeros(n)
    for it in range(1, 30):
        G.apply(b, x)
        _, xo = O.apply(bh, xo)
        assert x.getData().tobytes() == xo.tobytes()
        if np.linalg.norm(bh - Oc.mult(xo)) < 1e-8 * np.linalg.norm(bh):
            break
    assert it <= 16


def test_aliased_and_rebuilt(ctx):
    n, A, G, O = _pair(ctx, "p2d_48_w_cycle")
    bh = P.uniform(n, 7)
    bb = M.DenseVector(bh, ctx)
    G.apply(bb, bb)  # b and x one vector: apply copies both first
    _, ref = O.apply(bh, bh)
    assert bb.getData().tobytes() == ref.tobytes()
    # setMatrix again with another matrix rebuilds the hierarchy
    sysm = P.varcoef_poisson(2, 30, seed=9)
    A2 = M.CompRowMatrix(*sysm, ctx=ctx)
    G.setMatrix(A2)
    O2 = oracle.Amg(oracle.Csr(*sysm), 1.2, 1.5, 1.5, 1.2, 2, 1, 2, 30, 0.5)
    assert G.levels() == O2.levels()
    b2 = P.uniform(sysm[0], 8)
    x = M.DenseVector(sysm[0], ctx)
    G.apply(M.DenseVector(b2, ctx), x)
    assert x.getData().tobytes() == O2.apply(b2)[1].tobytes()


def test_direct_solve_only(ctx):
    rng = np.random.default_rng(3)
    n = 30
    d = rng.random((n, n)) * (rng.random((n, n)) < 0.2) + np.eye(n) * 5
    m = sp.csr_matrix(d)
    A = M.CompRowMatrix(m, ctx=ctx)
    G = M.AMG(ctx=ctx)
    G.setMatrix(A)
    assert G.levels() == [n]
    O = oracle.Amg(oracle.Csr(n, n, m.indptr.astype(np.int32), m.indices.astype(np.int32), m.data))
    b = rng.standard_normal(n)
    for tr in (False, True):
        x = M.DenseVector(n, ctx)
        (G.transApply if tr else G.apply)(M.DenseVector(b, ctx), x)
        assert x.getData().tobytes() == O.apply(b, None, transpose=tr)[1].tobytes()


def test_errors(ctx):
    m = sp.csr_matrix(np.array([[1.0, 1.0], [1.0, 0.0]]))
    m.eliminate_zeros()
    A = M.CompRowMatrix(m, ctx=ctx)
    G = M.AMG(1.0, 1.85, 1.85, 1.0, 1, 1, 1, 1, 2.0 / 3, ctx=ctx)
    with pytest.raises(M.MissingDiagonalException) as e:
        G.setMatrix(A)
    assert e.value.row == 1
    with pytest.raises(M.IllegalArgumentException):  # setMatrix must come first
        G.apply(M.DenseVector(2, ctx), M.DenseVector(2, ctx))
    with pytest.raises(M.IllegalArgumentException):
        M.AMG(1.0, 2.5, 1.85, 1.0, 1, 1, 1, 40, 2.0 / 3, ctx=ctx).setMatrix(
            M.CompRowMatrix(*P.poisson2d(10), ctx=ctx))
    R = M.CompRowMatrix(sp.random(4, 5, density=0.5, format="csr", random_state=1), ctx=ctx)
    with pytest.raises(M.IllegalArgumentException):
        M.AMG(ctx=ctx).setMatrix(R)
    S = M.CompRowMatrix(sp.csr_matrix(np.array([[1.0, 1.0], [1.0, 1.0]])), ctx=ctx)
    G = M.AMG(ctx=ctx)
    G.setMatrix(S)
    x = M.DenseVector(np.array([7.0, 8.0]), ctx)
    with pytest.raises(M.MatrixSingularException):
        G.apply(M.DenseVector(np.ones(2), ctx), x)
    assert list(x.getData()) == [7.0, 8.0]


@pytest.mark.parametrize("seed", range(4))
def test_cg_with_amg_reference_style(ctx, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 41))
    d = np.zeros((n, n))
    for i in range(n):
        cols = rng.choice(n, size=min(n, 5), replace=False)
        d[i, cols] = rng.random(len(cols))
    d = d + d.T + np.eye(n) * 100.0
    m = sp.csr_matrix(d)
    m.sort_indices()
    O = oracle.Csr(n, n, m.indptr.astype(np.int32), m.indices.astype(np.int32), m.data)
    go = oracle.Amg(O)
    xt = rng.random(n)
    x0 = rng.random(n)
    oc, _, oit, _ = oracle.solve("cg", O, d @ xt, x0, go.precond(), oracle.monitor(100000, 1e-5))
    A = M.CompRowMatrix(m, ctx=ctx)
    x = M.DenseVector(x0, ctx)
    S = M.CG(x)
    G = M.AMG(ctx=ctx)
    G.setMatrix(A)
    S.setPreconditioner(G)
    S.setIterationMonitor(M.DefaultIterationMonitor(100000, 1e-5))
    S.solve(A, M.DenseVector(d @ xt, ctx), x)
    assert oc == 0
    assert abs(S.getIterationMonitor().iterations() - oit) <= 1
    np.testing.assert_allclose(x.getData(), xt, atol=1e-4, rtol=0)
