"""GMRES's Gram-Schmidt loop as one cooperative launch per Arnoldi step
(k_gmres_mgs_all, workgroups holding 1 or 2 tiles of w in registers)
against the per-step launches (MTJG_GMRES_COOP=0): the same grid, tile
order and reduction tree, so the iterates are bitwise equal.  The per-step
form is itself checked against the oracle's GMRES (S/GMRES.java:110-200)
by the solver tests.  Also the barrier's poll-budget error path."""

import numpy as np
import pytest

import mtjgpu as M
from mtjgpu import _lib as L
from mtjgpu import problems as P

pytestmark = pytest.mark.gpu


# (system, preconditioner, iteration cap): 24^3 has a part-filled last tile
# (one tile per workgroup); 128^3 fills the 2048-workgroup grid with one
# tile per workgroup, 1500^2 with two
CASES = {
    "convdiff24_ilu": (lambda: P.convdiff3d(24, peclet=10.0, seed=1), "ilu", 10000),
    "convdiff128": (lambda: P.convdiff3d(128, peclet=10.0, seed=1), None, 40),
    "poisson1500": (lambda: P.poisson2d(1500), None, 40),
}


def _solve(ctx, sysm, pc, cap):
    n = sysm[0]
    A = M.CompRowMatrix(*sysm, ctx=ctx)
    b = M.DenseVector(P.uniform(n, 3), ctx)
    x = M.DenseVector(n, ctx)
    S = M.GMRES(x, 30)
    if pc == "ilu":
        ilu = M.ILU(A)
        ilu.setMatrix(A)
        S.setPreconditioner(ilu)
    if cap < 10000:
        S.setIterationMonitor(M.DefaultIterationMonitor(cap, 1e-30))
        with pytest.raises(M.IterativeSolverNotConvergedException):
            S.solve(A, b, x)
    else:
        S.setIterationMonitor(M.DefaultIterationMonitor(cap, 1e-10))
        S.solve(A, b, x)
    return S.getIterationMonitor().iterations(), x.getData()


@pytest.mark.parametrize("name", list(CASES))
def test_coop_same_bits_as_per_step(ctx, monkeypatch, name):
    make, pc, cap = CASES[name]
    sysm = make()
    out = []
    for coop in ("1", "0"):
        monkeypatch.setenv("MTJG_GMRES_COOP", coop)
        print(f"gmres coop={coop} case={name} n={sysm[0]}", flush=True)
        out.append(_solve(ctx, sysm, pc, cap))
    assert out[0][0] == out[1][0]
    np.testing.assert_array_equal(out[0][1], out[1][1])


def test_poll_budget_error_path(ctx, monkeypatch):
    """A poll budget of 0, set from the host, admits no wait: a workgroup
    that reaches the step barrier before the step is published raises the
    error flag, sets the device-error status and leaves, and the others leave
    at their next wait.  solve then fails with MTJG_E_DEVICE, having cleared
    the flag and the reduction's tickets; a new solver with the default
    budget afterwards is bitwise the per-step form's."""
    sysm = P.poisson2d(1024)
    n = sysm[0]
    A = M.CompRowMatrix(*sysm, ctx=ctx)
    b = M.DenseVector(P.uniform(n, 5), ctx)
    monkeypatch.setenv("MTJG_GMRES_POLL_BUDGET", "0")
    x = M.DenseVector(n, ctx)
    S = M.GMRES(x, 30)
    S.setIterationMonitor(M.DefaultIterationMonitor(40, 1e-30))
    print("gmres poll budget 0", flush=True)
    with pytest.raises(M.MtjgError) as e:
        S.solve(A, b, x)
    assert e.value.code == L.E_DEVICE
    monkeypatch.delenv("MTJG_GMRES_POLL_BUDGET")
    ref = _solve(ctx, sysm, None, 40)
    monkeypatch.setenv("MTJG_GMRES_COOP", "0")
    base = _solve(ctx, sysm, None, 40)
    assert ref[0] == base[0]
    np.testing.assert_array_equal(ref[1], base[1])
