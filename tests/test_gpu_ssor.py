"""GPU SSOR (mtjg_prec_ssor_*, through the C ABI) against the oracle's
restatement of S/SSOR.java: the level-scheduled sweeps do MTJ's per-row work
in MTJ's order, so apply and transApply are bitwise the oracle's for the same
incoming x, on stencil and random-pattern matrices, both relaxations and with
the reverse sweep on or off; setOmega / setMatrix errors; and SSOR inside CG
and BiCGstab on the reference's random SPD / nonsymmetric test systems (same
iteration count as the oracle, solution to 1e-4)."""
import numpy as np
import pytest
import scipy.sparse as sp

import mtjgpu as M
import oracle
from mtjgpu import problems as P

pytestmark = pytest.mark.gpu


def _mats():
    yield P.varcoef_poisson(3, 12, 10, 9, seed=2)
    yield P.convdiff3d(16, peclet=10.0, seed=3)
    n, _, rp, ci, va = P.powerlaw(3000, seed=4)
    yield n, n, rp, ci, va


@pytest.mark.parametrize("wf,wr,reverse", [(1.0, 1.0, True), (1.3, 0.7, True), (1.7, 1.7, False)])
def test_apply_bitwise(ctx, wf, wr, reverse):
    for n, m, rp, ci, va in _mats():
        A = M.CompRowMatrix(n, m, rp, ci, va, ctx=ctx)
        S = M.SSOR(A, reverse=reverse, omegaF=wf, omegaR=wr)
        S.setMatrix(A)
        O = oracle.Csr(n, m, rp, ci, va)
        Op = oracle.Precond(oracle.PREC_SSOR, O, reverse=reverse, omega_f=wf, omega_r=wr)
        b = P.uniform(n, 7)
        for x0 in (np.zeros(n), P.uniform(n, 8)):
            x = M.DenseVector(x0, ctx)
            S.apply(M.DenseVector(b, ctx), x)
            assert x.getData().tobytes() == Op.apply(b, x0).tobytes()
            x = M.DenseVector(x0, ctx)
            S.transApply(M.DenseVector(b, ctx), x)
            assert x.getData().tobytes() == Op.trans_apply(b, x0).tobytes()


def test_errors(ctx):
    n, m, rp, ci, va = P.poisson2d(8)
    A = M.CompRowMatrix(n, m, rp, ci, va, ctx=ctx)
    with pytest.raises(M.IllegalArgumentException):
        M.SSOR(A, omegaF=2.5)
    S = M.SSOR(A)
    with pytest.raises(M.IllegalArgumentException):
        S.setOmega(1.0, -0.5)
    d = sp.csr_matrix((va, ci, rp), shape=(n, n)).tolil()
    d[7, 7] = 0
    d = d.tocsr()
    d.eliminate_zeros()
    B = M.CompRowMatrix(d, ctx=ctx)
    with pytest.raises(M.MissingDiagonalException) as e:
        S.setMatrix(B)
    assert e.value.row == 7
    R = M.CompRowMatrix(sp.random(4, 5, density=0.5, format="csr", random_state=1), ctx=ctx)
    with pytest.raises(M.IllegalArgumentException):
        M.SSOR(R)


def _mtj_random(seed, spd):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 50))
    k = int(rng.integers(0, min(10, n) + 1))
    d = np.zeros((n, n))
    for i in range(n):
        cols = rng.choice(n, size=k, replace=False) if k else []
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


@pytest.mark.parametrize("solver", ["cg", "bicgstab"])
@pytest.mark.parametrize("seed", range(5))
def test_ssor_inside_solvers(ctx, solver, seed):
    A = _mtj_random(seed, spd=(solver == "cg"))
    n = A.shape[0]
    rng = np.random.default_rng(200 + seed)
    xt = rng.random(n)
    b = A @ xt
    x0 = rng.random(n)
    O = oracle.Csr(n, n, A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data)
    oc, ox, oit, _ = oracle.solve(solver, O, b, x0, oracle.Precond(oracle.PREC_SSOR, O),
                                  oracle.monitor(100000, 1e-5))
    Ad = M.CompRowMatrix(A, ctx=ctx)
    x = M.DenseVector(x0, ctx)
    S = {"cg": M.CG, "bicgstab": M.BiCGstab}[solver](x)
    Pc = M.SSOR(Ad)
    Pc.setMatrix(Ad)
    S.setPreconditioner(Pc)
    S.setIterationMonitor(M.DefaultIterationMonitor(100000, 1e-5))
    S.solve(Ad, M.DenseVector(b, ctx), x)
    assert oc == 0
    assert abs(S.getIterationMonitor().iterations() - oit) <= 1
    np.testing.assert_allclose(x.getData(), xt, atol=1e-4, rtol=0)
