"""Pins the oracle's MTJ-specific control flow in CG, BiCGstab and GMRES
independently of its own code:

* A6, BiCGstab's half-step check (S/BiCGstab.java:102-104) runs while
  iter == 0, so DefaultIterationMonitor re-captures initR = ||s_0||
  (S/DefaultIterationMonitor.java:136-137) and every later check is relative
  to ||s_0||, not ||r_0||.  Exact-arithmetic cases worked by hand below, one
  of which changes the stopping iteration; and SciPy's bicgstab, whose
  checks use a fixed absolute tolerance, gives the same iteration counts only
  when that tolerance is rtol * ||s_0||.
* A7, GMRES counts one iteration per Arnoldi step (S/GMRES.java:144-145)
  and one per completed restart cycle (:137).  A GMRES(1) case worked by
  hand in closed form.
* CG (S/CG.java:58-87) with x0 = 0 stops where SciPy's cg does
  (||r_k|| < rtol ||b||, the same two-rounding p update): the iteration
  counts are equal.

Each case is chosen so that an oracle without the quirk fails it."""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

import oracle
from mtjgpu import problems as P


def _csr(dense):
    A = sp.csr_matrix(np.asarray(dense, dtype=float))
    A.sort_indices()
    return A, oracle.Csr(A.shape[0], A.shape[1], A.indptr.astype(np.int32), A.indices.astype(np.int32),
                         A.data.astype(np.float64))


def test_bicgstab_recaptures_init_r_and_stops_one_iteration_later():
    """A = diag(1, 2, 3), b = (1, 1, 1), x0 = 0, identity preconditioner,
    rtol = 0.5.

    iter 0, top:  r0 = (1, 1, 1), ||r0|| = sqrt(3) = 1.732 -> initR = sqrt(3).
                  rho = rtilde.r = 3, p = r0, v = A p = (1, 2, 3),
                  alpha = 3 / (rtilde.v = 6) = 1/2,
                  s = r0 - v/2 = (1/2, 0, -1/2), x = (1/2, 1/2, 1/2).
    iter 0, half: ||s|| = sqrt(1/2) = 0.7071 with iter == 0, so
                  initR := 0.7071 (A6) and the test is 0.7071 < 0.3536: no.
                  [Without A6: 0.7071 < 0.5 * 1.732 = 0.866 -> stop here,
                  iterations() = 0, x = (1/2, 1/2, 1/2).]
                  t = A s = (1/2, 0, -3/2), t.s = 1, t.t = 5/2, omega = 2/5,
                  x = (1/2 + 1/5, 1/2, 1/2 - 1/5) = (7/10, 1/2, 3/10),
                  r = s - omega t = (3/10, 0, 1/10).
    iter 1, top:  ||r|| = sqrt(10)/10 = 0.3162 < 0.5 * 0.7071 = 0.3536: stop,
                  iterations() = 1, residual sqrt(10)/10."""
    _, O = _csr(np.diag([1.0, 2.0, 3.0]))
    mon = oracle.monitor(100, 0.5)
    code, x, it, res = oracle.solve("bicgstab", O, np.ones(3), np.zeros(3), None, mon)
    assert code == oracle.OK and it == 1
    assert res == pytest.approx(math.sqrt(10) / 10, rel=1e-15)
    assert mon.init_r == pytest.approx(math.sqrt(0.5), rel=1e-15)
    np.testing.assert_allclose(x, [0.7, 0.5, 0.3], rtol=1e-15)


def test_bicgstab_half_step_return_after_recapture():
    """A = diag(1, 2), b = (1, 1), x0 = 0, identity preconditioner, rtol = 0.2.

    iter 0, top:  ||r0|| = sqrt(2) -> initR = sqrt(2); rho = 2, p = (1, 1),
                  v = (1, 2), alpha = 2/3, s = (1/3, -1/3), x = (2/3, 2/3).
    iter 0, half: ||s|| = sqrt(2)/3 = 0.4714 -> initR := 0.4714 (A6), test
                  0.4714 < 0.0943: no.  t = (1/3, -2/3), omega = (1/3)/(5/9)
                  = 3/5, x = (13/15, 7/15), r = (2/15, 1/15).
    iter 1, top:  ||r|| = sqrt(5)/15 = 0.1491 < 0.0943: no.
                  [Without A6 the threshold is 0.2 * sqrt(2) = 0.2828 and the
                  solve stops here with x = (13/15, 7/15), residual 0.1491.]
                  rho = 1/5, beta = (1/5 / 2) (2/3 / 3/5) = 1/9,
                  p = r + beta (p - omega v) = (8/45, 2/45), v = (8/45, 4/45),
                  alpha = (1/5) / (12/45) = 3/4, s = r - alpha v = (0, 0),
                  x = x + alpha p = (1, 1/2): the exact solution.
    iter 1, half: ||s|| = 0 (to rounding) < 0.0943: return, iterations() = 1."""
    _, O = _csr(np.diag([1.0, 2.0]))
    code, x, it, res = oracle.solve("bicgstab", O, np.ones(2), np.zeros(2), None, oracle.monitor(100, 0.2))
    assert code == oracle.OK and it == 1
    assert res < 1e-15
    np.testing.assert_allclose(x, [1.0, 0.5], rtol=1e-15)


def test_gmres_counts_restart_cycles():
    """GMRES(1) on A = diag(1, 2), b = (1, 1), x0 = 0, identity
    preconditioner, rtol = 5e-4.

    A cycle of GMRES(1) is one minimal-residual step along r:
    x += (r.Ar / Ar.Ar) r.  From r0 = (1, 1): Ar0 = (1, 2), step 3/5,
    r1 = (2/5, -1/5); Ar1 = (2/5, -2/5), step (6/25)/(8/25) = 3/4,
    r2 = (1/10, 1/10) = r0 / 10.  So ||r_k|| / ||r0|| runs 0.316, 0.1,
    0.0316, 0.01, 3.2e-3, 1e-3, 3.2e-4: the first below 5e-4 is k = 7, with
    r7 = r1 / 1000 and x7 = A^-1 (b - r7) = (0.9996, 0.5001).

    Counting (S/GMRES.java:130-146): cycle k's outer check sees iter = 2k,
    because each cycle adds one iter.next() for its Arnoldi step and one for
    the restart; the check after 7 cycles is at iterations() = 14.
    [Counting Arnoldi steps only would give 7.]"""
    _, O = _csr(np.diag([1.0, 2.0]))
    code, x, it, res = oracle.solve("gmres", O, np.ones(2), np.zeros(2), None, oracle.monitor(1000, 5e-4),
                                    restart=1)
    assert code == oracle.OK and it == 14
    assert res == pytest.approx(math.sqrt(5) / 5 * 1e-3, rel=1e-12)
    np.testing.assert_allclose(x, [0.9996, 0.5001], rtol=1e-12)


def test_gmres_restart_count_with_a_longer_cycle():
    """The same system with GMRES(2): two Arnoldi steps span R^2, so the
    first cycle ends exact, but the loop still runs both inner steps (the
    inner check of step 1 sees |s_1| > 0), restarts (iter.next()), and the
    outer check at iterations() = 3 sees a residual at rounding level."""
    _, O = _csr(np.diag([1.0, 2.0]))
    code, x, it, res = oracle.solve("gmres", O, np.ones(2), np.zeros(2), None, oracle.monitor(1000, 1e-8),
                                    restart=2)
    assert code == oracle.OK and it == 3
    np.testing.assert_allclose(x, [1.0, 0.5], rtol=1e-14)


def _convdiff(grid, seed):
    n, _, rp, ci, va = P.convdiff3d(grid, peclet=10.0, seed=seed)
    A = sp.csr_matrix((va, ci, rp), shape=(n, n))
    return A, oracle.Csr(n, n, rp, ci, va)


def _count(fn, A, b, **kw):
    cnt = [0]
    x, info = fn(A, b, maxiter=100000, callback=lambda xk: cnt.__setitem__(0, cnt[0] + 1), **kw)
    assert info == 0
    return cnt[0], x


@pytest.mark.parametrize("prec", ["none", "diag"])
@pytest.mark.parametrize("grid,rtol", [(10, 1e-10), (14, 1e-9)])
def test_bicgstab_counts_match_scipy_at_s0_threshold(prec, grid, rtol):
    """SciPy's bicgstab checks ||r|| and ||s|| against one fixed atol.  With
    x0 = 0 MTJ's checks after the first half step are all relative to
    ||s_0|| (A6), so the counts are equal when atol = rtol * ||s_0||, with
    s_0 = b - alpha_0 A M b, alpha_0 = (b.b) / (b.A M b).  On these systems
    atol = rtol * ||b|| gives a different count, so the match pins A6."""
    A, O = _convdiff(grid, 3)
    n = A.shape[0]
    b = A @ P.uniform(n, 2)
    minv = 1.0 / A.diagonal() if prec == "diag" else np.ones(n)
    op = oracle.Precond(oracle.PREC_DIAGONAL, O) if prec == "diag" else None
    code, x, it, _ = oracle.solve("bicgstab", O, b, np.zeros(n), op, oracle.monitor(100000, rtol))
    assert code == oracle.OK
    v = A @ (minv * b)
    s0 = b - (b @ b) / (b @ v) * v
    J = sla.LinearOperator((n, n), matvec=lambda u: minv * u, dtype=float)
    it_s0, xs = _count(sla.bicgstab, A, b, rtol=0.0, atol=rtol * np.linalg.norm(s0), M=J)
    it_b, _ = _count(sla.bicgstab, A, b, rtol=0.0, atol=rtol * np.linalg.norm(b), M=J)
    assert it == it_s0
    assert it_b != it_s0
    assert np.linalg.norm(x - xs) <= 1e-6 * np.linalg.norm(xs)


@pytest.mark.parametrize("prec", ["none", "diag"])
@pytest.mark.parametrize("gen", ["poisson3d_12", "varcoef_3d_10", "powerlaw_spd_2000"])
def test_cg_counts_match_scipy(prec, gen):
    sysm = {"poisson3d_12": lambda: P.poisson3d(12), "varcoef_3d_10": lambda: P.varcoef_poisson(3, 10, seed=4),
            "powerlaw_spd_2000": lambda: P.powerlaw_spd(2000, seed=5)}[gen]()
    n, _, rp, ci, va = sysm
    A = sp.csr_matrix((va, ci, rp), shape=(n, n))
    O = oracle.Csr(*sysm)
    b = O.mult(P.uniform(n, 1))
    op = oracle.Precond(oracle.PREC_DIAGONAL, O) if prec == "diag" else None
    code, x, it, _ = oracle.solve("cg", O, b, np.zeros(n), op, oracle.monitor(100000, 1e-10))
    assert code == oracle.OK
    minv = 1.0 / A.diagonal() if prec == "diag" else np.ones(n)
    J = sla.LinearOperator((n, n), matvec=lambda u: u * minv, dtype=float)
    it_s, xs = _count(sla.cg, A, b, rtol=1e-10, atol=0.0, M=J)
    assert it == it_s
    assert np.linalg.norm(x - xs) <= 1e-8 * np.linalg.norm(xs)
