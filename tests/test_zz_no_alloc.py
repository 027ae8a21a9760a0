"""GPU test of the allocation-free CG solve (mtjg_cg_solve_noalloc), through
the C ABI, against the existing entry point it varies (mtjg_solver_solve with
an MTJG_CG solver and the same preconditioner): x, the iteration count and the
residual bit for bit, on calls in a row whose systems shrink (so each call
must clear what the previous one left in the shared scratch array), one with
the fused product and one with the split product of large grids; a system
larger than the scratch is rejected before any device work."""
import ctypes as C

import numpy as np
import pytest

import mtjgpu as M
from mtjgpu import _lib as L
from mtjgpu import problems as P

pytestmark = pytest.mark.gpu


def _noalloc(ctx, A, Mp, b, x, rtol):
    mon = L.Monitor(100000, rtol, 1e-50, 1e5)
    res = L.Result()
    code = ctx._lib.mtjg_cg_solve_noalloc(ctx.handle, A.handle, Mp.handle if Mp is not None else None, b.handle,
                                          x.handle, C.byref(mon), C.byref(res))
    return code, res


def _case(ctx, sysm, prec, rtol=1e-8):
    n = sysm[0]
    A = M.CompRowMatrix(*sysm, ctx=ctx)
    bh = P.uniform(n, 3)
    b = M.DenseVector(bh, ctx)
    Mp = None
    if prec == "diag":
        Mp = M.DiagonalPreconditioner(n, ctx)
        Mp.setMatrix(A)
    # the existing entry point
    x_ref = M.DenseVector(n, ctx)
    S = M.CG(x_ref)
    S.setPreconditioner(Mp)
    S.setIterationMonitor(M.DefaultIterationMonitor(100000, rtol))
    S.solve(A, b, x_ref)
    it_ref = S.getIterationMonitor().iterations()
    res_ref = S.getIterationMonitor().residual()
    # the allocation-free one, from a nonzero initial guess's worth of stale
    # memory: x starts at zero as for the reference
    x = M.DenseVector(n, ctx)
    code, res = _noalloc(ctx, A, Mp, b, x, rtol)
    assert code == L.MTJG_OK, ctx._lib.mtjg_last_error()
    assert res.iterations == it_ref and res.residual == res_ref
    np.testing.assert_array_equal(x.getData(), x_ref.getData())
    for o in (S, x_ref, x, b, A) + ((Mp,) if Mp is not None else ()):
        o.close()
    return res.iterations


def test_noalloc_matches_solver_solve_on_shrinking_systems(ctx):
    # 1,000,000 rows: the split fused product (grid above one persistent grid)
    it1 = _case(ctx, P.varcoef_poisson(3, 100, seed=2), "diag")
    # 110,592 rows: the fused product with the direction update folded in
    it2 = _case(ctx, P.varcoef_poisson(3, 48, seed=3), "diag")
    # 4,096 rows, no preconditioner
    it3 = _case(ctx, P.poisson2d(64), "none")
    assert it1 > 0 and it2 > 0 and it3 > 0


def test_noalloc_rejects_oversize_and_other_preconditioners(ctx):
    n = (1 << 20) + 1
    rp = np.arange(n + 1, dtype=np.int32)
    A = M.CompRowMatrix(n, n, rp, np.arange(n, dtype=np.int32), np.full(n, 2.0), ctx=ctx)
    b = M.DenseVector(np.ones(n), ctx)
    x = M.DenseVector(n, ctx)
    code, _ = _noalloc(ctx, A, None, b, x, 1e-8)
    assert code == L.E_INVALID_ARGUMENT
    assert "exceeds" in ctx._lib.mtjg_last_error().decode()
    for o in (x, b, A):
        o.close()
    sysm = P.poisson2d(16)
    A = M.CompRowMatrix(*sysm, ctx=ctx)
    ilu = M.ILU(A)
    ilu.setMatrix(A)
    b = M.DenseVector(np.ones(256), ctx)
    x = M.DenseVector(256, ctx)
    code, _ = _noalloc(ctx, A, ilu, b, x, 1e-8)
    assert code == L.E_UNSUPPORTED
